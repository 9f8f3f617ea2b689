// Equal-time measurements of the Hubbard plugin on the device, accumulated per chain.
//
// Restates Obser for N_FL = 1 (Prog/Hamiltonians/Hamiltonian_Hubbard_smod.F90:718-799) with
// GRC(i, j) = <c^+_i c_j> = delta_ij - G(j, i), ZP = PHASE / Re PHASE, ZS = sign(Re PHASE):
//   Kin  = N_SUN sum_ij O_ij GRC(i, j)        (Predefined_Hop_mod.F90:1850-1959; O = H_T's matrix)
//   Pot  = U sum_i GRC(i, i)^2,  Part = N_SUN sum_i GRC(i, i),  Ener = Kin + Pot
//   Green(d, a, b) += N_SUN GRC(I1, J1)                      (Predefined_Obs_mod.F90:226-266)
//   SpinZ(d, a, b) += N_SUN GRC(I1, J1) G(I1, J1)            (Predefined_Obs_mod.F90:77-120)
//   Den(d, a, b)   += Z_I Z_J + N_SUN GRC(I1, J1) G(I1, J1),  Den0(a) += Z_I, Z_I = N_SUN GRC(I1, I1)
// for every pair of sites I1 = (cell I, orbital a), J1 = (cell J, orbital b) with d = imj(I, J),
// each term times ZP ZS; counts and sign sums as the reference keeps them.
//
// One launch per measured slice: workgroup d < ncell sums the correlation terms of distance
// class d over all cells (J = jm(I, d)) in a fixed order, workgroup ncell the scalars over the
// nonzero hopping entries and the sites; no atomics, so the accumulators are reproducible.
#include <vector>

#include "alfq_common.h"
#include "alfq_obs.h"

namespace {

constexpr int OT = 256;

__device__ __forceinline__ void block_sum(cplx* v, int m, cplx* red) {
  // v[0..m) of every thread summed into red[0..m) (fixed tree order)
  const int tid = threadIdx.x;
  for (int q = 0; q < m; ++q) red[q * OT + tid] = v[q];
  __syncthreads();
  for (int s = OT / 2; s > 0; s >>= 1) {
    if (tid < s)
      for (int q = 0; q < m; ++q) red[q * OT + tid] = cadd(red[q * OT + tid], red[q * OT + tid + s]);
    __syncthreads();
  }
}

__global__ __launch_bounds__(OT) void obs_kernel(const cplx* __restrict__ G, const ChainState* __restrict__ st,
                                                 ObsDev o, long long cs, long long ocs) {
  __shared__ cplx red[4 * OT];
  G = chain_ptr(G, cs);
  st = chain_ptr(st, cs);
  char* ob = o.buf + (long long)blockIdx.z * ocs;
  ObsHead* h = (ObsHead*)(ob + o.off_head);
  const double U = *(const double*)(ob + o.off_u);
  cplx* den0 = (cplx*)(ob + o.off_den0);
  cplx* latt = (cplx*)(ob + o.off_latt);
  const cplx* ov = (const cplx*)(ob + o.off_ov);
  const int tid = threadIdx.x;
  const int n = o.n, nc = o.ncell, no = o.norb;
  const double ns = (double)o.n_sun;
  const cplx ph = st->phase;
  const double zs = ph.x > 0.0 ? 1.0 : (ph.x < 0.0 ? -1.0 : 0.0);
  // ZP ZS = ZS (1, Im PHASE / Re PHASE)
  const cplx w = ph.x != 0.0 ? make_double2(zs, zs * ph.y / ph.x) : make_double2(0.0, 0.0);
  const int d = blockIdx.x;
  if (d < nc) {
    for (int p = 0; p < no * no; ++p) {
      const int a = p / no, b = p % no;
      cplx acc[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
      for (int I = tid; I < nc; I += OT) {
        const int i1 = o.site_of[I * no + a];
        const int J = o.jm[(long long)I * nc + d];
        const int j1 = o.site_of[J * no + b];
        const cplx g = G[i1 + (long long)j1 * n];
        const cplx gt = G[j1 + (long long)i1 * n];
        const cplx grc = make_double2((i1 == j1 ? 1.0 : 0.0) - gt.x, -gt.y);
        const cplx gii = G[i1 + (long long)i1 * n], gjj = G[j1 + (long long)j1 * n];
        const cplx zi = make_double2(ns * (1.0 - gii.x), -ns * gii.y);
        const cplx zj = make_double2(ns * (1.0 - gjj.x), -ns * gjj.y);
        const cplx sp = cscale(cmul(grc, g), ns);
        acc[0] = cadd(acc[0], cscale(grc, ns));
        acc[1] = cadd(acc[1], sp);
        acc[2] = cadd(acc[2], cadd(cmul(zi, zj), sp));
      }
      block_sum(acc, 3, red);
      if (tid == 0) {
        for (int q = 0; q < 3; ++q) {
          cplx* t = latt + ((long long)q * nc + d) * no * no + p;
          *t = cadd(*t, cmul(red[q * OT], w));
        }
      }
      __syncthreads();
    }
    return;
  }
  // scalars and the density background
  cplx acc[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
  for (int e = tid; e < o.nnz; e += OT) {
    const int i = o.oi[e], j = o.oj[e];
    const cplx gji = G[j + (long long)i * n];
    const cplx grc = make_double2((i == j ? 1.0 : 0.0) - gji.x, -gji.y);
    acc[0] = cadd(acc[0], cmul(ov[e], grc));
  }
  for (int i = tid; i < n; i += OT) {
    const cplx gii = G[i + (long long)i * n];
    const cplx c = make_double2(1.0 - gii.x, -gii.y);
    acc[1] = cadd(acc[1], cmul(c, c));
    acc[2] = cadd(acc[2], c);
  }
  block_sum(acc, 3, red);
  cplx kin = make_double2(0, 0), pot = make_double2(0, 0), part = make_double2(0, 0);
  if (tid == 0) {
    kin = cscale(red[0], ns);
    pot = cscale(red[OT], U);
    part = cscale(red[2 * OT], ns);
  }
  __syncthreads();
  for (int a = 0; a < no; ++a) {
    cplx v[1] = {make_double2(0, 0)};
    for (int I = tid; I < nc; I += OT) {
      const cplx gii = G[o.site_of[I * no + a] * (long long)(n + 1)];
      v[0] = cadd(v[0], make_double2(ns * (1.0 - gii.x), -ns * gii.y));
    }
    block_sum(v, 1, red);
    if (tid == 0) den0[a] = cadd(den0[a], cmul(red[0], w));
    __syncthreads();
  }
  if (tid == 0) {
    const cplx s4[4] = {kin, pot, part, cadd(kin, pot)};
    for (int q = 0; q < 4; ++q) h->scal[q] = cadd(h->scal[q], cmul(s4[q], w));
    h->n_scal += 1;
    h->sign_scal += zs;
    h->n_eq += 1;
    h->sign_eq += zs;
  }
}

}  // namespace

int obs_measure(const ObsDev& o, const cplx* G, const ChainState* st, int nchain, long long cs, long long ocs,
                hipStream_t s) {
  hipLaunchKernelGGL(obs_kernel, dim3((unsigned)(o.ncell + 1), 1, (unsigned)nchain), dim3(OT), 0, s, G, st, o, cs,
                     ocs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}
