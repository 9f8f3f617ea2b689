// Time-displaced Green's functions (TAU_M with CGR2_2) and measurements (ObserT) of the
// Hubbard plugin on the device, N_FL = 1, Symm = .F.
//
// TAU_M (Prog/tau_m_mod.F90:56-195) propagates G(tau, 0), G(0, tau) and G(tau, tau) slice
// by slice with the hopping and diagonal products of the sweep; at every stabilisation
// point CGR2_2 (Prog/cgr2_2_mod.F90:196-428) recomputes all four from the right chain
// B(tau, 0) = UR DR VR and the stored left state B(beta, tau)^H = UL DL VL^H as blocks of the
// inverse of O = [[1, B(beta, tau)], [-B(tau, 0), 1]].  The host side (alfq.hip) runs the
// products and the QR; this file holds the element-wise steps and the measurement.
#include "alfq_tau.h"

namespace {

constexpr int TT = 256;

__global__ void tau_init_kernel(const cplx* __restrict__ G, cplx* __restrict__ G00, cplx* __restrict__ GT0,
                                cplx* __restrict__ GTT, cplx* __restrict__ G0T, int n, long long cs) {
  G = chain_ptr(G, cs);
  G00 = chain_ptr(G00, cs);
  GT0 = chain_ptr(GT0, cs);
  GTT = chain_ptr(GTT, cs);
  G0T = chain_ptr(G0T, cs);
  const long long nn = (long long)n * n;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn) return;
  const cplx g = G[e];
  G00[e] = g;
  GT0[e] = g;
  GTT[e] = g;
  const bool diag = (e % n) == (e / n);
  G0T[e] = make_double2(g.x - (diag ? 1.0 : 0.0), g.y);         // -(1 - G)
}

// D+^-1 and D- of the STAB3 split (cgr2_2_mod.F90:285-299): D <= 1 stays in D-, D > 1
// moves into the rows of V^-1 (left) or U^H (right).
__device__ __forceinline__ cplx dplus_inv(cplx d) { return d.x > 1.0 ? cinv(d) : make_double2(1.0, 0.0); }
__device__ __forceinline__ cplx dminus(cplx d) { return d.x <= 1.0 ? d : make_double2(1.0, 0.0); }

__global__ void cgr22_build_kernel(const cplx* __restrict__ X, const int* __restrict__ pinv,
                                   const cplx* __restrict__ UL, const cplx* __restrict__ DL,
                                   const cplx* __restrict__ UR, const cplx* __restrict__ DR,
                                   const cplx* __restrict__ VR, cplx* __restrict__ H, cplx* __restrict__ F1,
                                   cplx* __restrict__ F2, int* __restrict__ flag, int n, long long cs) {
  X = chain_ptr(X, cs);
  pinv = chain_ptr(pinv, cs);
  UL = chain_ptr(UL, cs);
  DL = chain_ptr(DL, cs);
  UR = chain_ptr(UR, cs);
  DR = chain_ptr(DR, cs);
  VR = chain_ptr(VR, cs);
  H = chain_ptr(H, cs);
  F1 = chain_ptr(F1, cs);
  F2 = chain_ptr(F2, cs);
  flag = chain_ptr(flag, cs);
  const int m = 2 * n;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)m * m) return;
  const int i = (int)(e % m), j = (int)(e / m);
  const bool first = DL[0].x > DR[0].x;
  if (e == 0) *flag = first ? 1 : 0;
  // block (bi, bj) of the left-first order; the other order swaps both block indices
  int bi = i / n, bj = j / n;
  const int r = i - bi * n, c = j - bj * n;
  if (!first) { bi ^= 1; bj ^= 1; }
  cplx v;
  if (bi == 0 && bj == 0) {
    v = cmul(dplus_inv(DL[r]), X[pinv[r] + (long long)c * n]);          // D_L+^-1 V_L^-1
  } else if (bi == 0) {
    v = cmul(dminus(DL[r]), cconj(UL[c + (long long)r * n]));           // D_L- U_L^H
  } else if (bj == 0) {
    const cplx t = cmul(dminus(DR[r]), VR[r + (long long)c * n]);       // -D_R- V_R
    v = make_double2(-t.x, -t.y);
  } else {
    v = cmul(dplus_inv(DR[r]), cconj(UR[c + (long long)r * n]));        // D_R+^-1 U_R^H
  }
  H[e] = v;
  if (i < n && j < n) F1[r + (long long)c * n] = v;
  if (i >= n && j >= n) F2[r + (long long)c * n] = v;
}

constexpr int XT = 256;
constexpr int XB = CGR22_PARTS;   // workgroups per chain of the block extraction

// Grid-wide pass: the blocks of O^-1 into G00, G0T, GT0, GTT and, per workgroup, the
// maxima and sums of |old - new| of the four Control_Precision_tau comparisons into
// part[blockIdx.x][8].
__global__ __launch_bounds__(XT) void cgr22_extract_kernel(const cplx* __restrict__ Oinv,
                                                           const int* __restrict__ flag,
                                                           const cplx* __restrict__ GR, cplx* __restrict__ G00,
                                                           cplx* __restrict__ G0T, cplx* __restrict__ GT0,
                                                           cplx* __restrict__ GTT, double* __restrict__ part, int n,
                                                           long long cs) {
  __shared__ double smax[4][XT];
  __shared__ double ssum[4][XT];
  Oinv = chain_ptr(Oinv, cs);
  flag = chain_ptr(flag, cs);
  GR = chain_ptr(GR, cs);
  G00 = chain_ptr(G00, cs);
  G0T = chain_ptr(G0T, cs);
  GT0 = chain_ptr(GT0, cs);
  GTT = chain_ptr(GTT, cs);
  part = chain_ptr(part, cs);
  const int tid = threadIdx.x;
  const bool first = *flag != 0;
  const int m = 2 * n;
  double mx[4] = {0, 0, 0, 0}, sm[4] = {0, 0, 0, 0};
  const long long nn = (long long)n * n;
  for (long long e = (long long)blockIdx.x * XT + tid; e < nn; e += (long long)XB * XT) {
    const int i = (int)(e % n), j = (int)(e / n);
    const cplx a = Oinv[i + (long long)j * m];               // upper-left
    const cplx b = Oinv[i + (long long)(j + n) * m];         // upper-right
    const cplx c = Oinv[(i + n) + (long long)j * m];         // lower-left
    const cplx d = Oinv[(i + n) + (long long)(j + n) * m];   // lower-right
    // get_blocks(GR00, GR0T, GRT0, GRTT) or, swapped, get_blocks(GRTT, GRT0, GR0T, GR00)
    const cplx n00 = first ? a : d, n0t = first ? b : c, nt0 = first ? c : b, ntt = first ? d : a;
    const double d0 = cabs_(csub(GR[e], n00));
    const double d1 = cabs_(csub(GTT[e], ntt));
    const double d2 = cabs_(csub(GT0[e], nt0));
    const double d3 = cabs_(csub(G0T[e], n0t));
    mx[0] = fmax(mx[0], d0); sm[0] += d0;
    mx[1] = fmax(mx[1], d1); sm[1] += d1;
    mx[2] = fmax(mx[2], d2); sm[2] += d2;
    mx[3] = fmax(mx[3], d3); sm[3] += d3;
    G00[e] = n00;
    G0T[e] = n0t;
    GT0[e] = nt0;
    GTT[e] = ntt;
  }
  for (int q = 0; q < 4; ++q) {
    smax[q][tid] = mx[q];
    ssum[q][tid] = sm[q];
  }
  __syncthreads();
  for (int s = XT / 2; s > 0; s >>= 1) {
    if (tid < s)
      for (int q = 0; q < 4; ++q) {
        smax[q][tid] = fmax(smax[q][tid], smax[q][tid + s]);
        ssum[q][tid] += ssum[q][tid + s];
      }
    __syncthreads();
  }
  if (tid < 4) {
    part[blockIdx.x * 8 + tid] = smax[tid][0];
    part[blockIdx.x * 8 + 4 + tid] = ssum[tid][0];
  }
}

// The workgroups' partial results in a fixed order, then Control_Precision_tau in the
// reference's order: (GR, G00), (old GTT, GTT), (old GT0, GT0), (old G0T, G0T).
__global__ void cgr22_precision_kernel(const double* __restrict__ part, TauCtl* ctl, int n, long long cs) {
  part = chain_ptr(part, cs);
  ctl = chain_ptr(ctl, cs);
  if (threadIdx.x != 0) return;
  const double nn = (double)n * (double)n;
  for (int q = 0; q < 4; ++q) {
    double mx = 0.0, sm = 0.0;
    for (int b = 0; b < XB; ++b) {
      mx = fmax(mx, part[b * 8 + q]);
      sm += part[b * 8 + 4 + q];
    }
    ctl->ncg += 1;
    ctl->xmax = fmax(ctl->xmax, mx);
    ctl->xmean += sm / nn;
  }
}

__device__ __forceinline__ void block_sum3(cplx (&v)[3], cplx* red) {
  const int tid = threadIdx.x;
  for (int q = 0; q < 3; ++q) red[q * TT + tid] = v[q];
  __syncthreads();
  for (int s = TT / 2; s > 0; s >>= 1) {
    if (tid < s)
      for (int q = 0; q < 3; ++q) red[q * TT + tid] = cadd(red[q * TT + tid], red[q * TT + tid + s]);
    __syncthreads();
  }
}

__global__ __launch_bounds__(TT) void tau_obs_kernel(const cplx* __restrict__ GT0, const cplx* __restrict__ G0T,
                                                     const cplx* __restrict__ G00, const cplx* __restrict__ GTT,
                                                     const ChainState* __restrict__ st, ObsDev o, char* acc,
                                                     long long tau_cs, long long off_den0, long long off_latt, int nt,
                                                     int ntau, long long cs) {
  __shared__ cplx red[3 * TT];
  GT0 = chain_ptr(GT0, cs);
  G0T = chain_ptr(G0T, cs);
  G00 = chain_ptr(G00, cs);
  GTT = chain_ptr(GTT, cs);
  st = chain_ptr(st, cs);
  char* blk = acc + (long long)blockIdx.z * tau_cs;
  TauHead* h = (TauHead*)blk;
  cplx* den0 = (cplx*)(blk + off_den0);
  cplx* latt = (cplx*)(blk + off_latt);
  const int tid = threadIdx.x;
  const int n = o.n, nc = o.ncell, no = o.norb;
  const double ns = (double)o.n_sun;
  const cplx ph = st->phase;
  const double zs = ph.x > 0.0 ? 1.0 : (ph.x < 0.0 ? -1.0 : 0.0);
  const cplx w = ph.x != 0.0 ? make_double2(zs, zs * ph.y / ph.x) : make_double2(0.0, 0.0);
  const int d = blockIdx.x;
  if (d < nc) {
    for (int p = 0; p < no * no; ++p) {
      const int a = p / no, b = p % no;
      cplx v[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
      for (int I = tid; I < nc; I += TT) {
        const int i1 = o.site_of[I * no + a];
        const int J = o.jm[(long long)I * nc + d];
        const int j1 = o.site_of[J * no + b];
        const cplx gt0 = GT0[i1 + (long long)j1 * n];
        const cplx g0t = G0T[j1 + (long long)i1 * n];
        const cplx gtt = GTT[i1 + (long long)i1 * n], g00 = G00[j1 + (long long)j1 * n];
        const cplx zi = make_double2(ns * (1.0 - gtt.x), -ns * gtt.y);
        const cplx zj = make_double2(ns * (1.0 - g00.x), -ns * g00.y);
        const cplx sp = cscale(cmul(g0t, gt0), -ns);
        v[0] = cadd(v[0], gt0);
        v[1] = cadd(v[1], sp);
        v[2] = cadd(v[2], cadd(cmul(zi, zj), sp));
      }
      block_sum3(v, red);
      if (tid == 0) {
        for (int q = 0; q < 3; ++q) {
          cplx* t = latt + (((long long)q * ntau + nt) * nc + d) * no * no + p;
          *t = cadd(*t, cmul(red[q * TT], w));
        }
      }
      __syncthreads();
    }
    return;
  }
  for (int a = 0; a < no; ++a) {
    cplx v[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    for (int I = tid; I < nc; I += TT) {
      const cplx gtt = GTT[o.site_of[I * no + a] * (long long)(n + 1)];
      v[0] = cadd(v[0], make_double2(ns * (1.0 - gtt.x), -ns * gtt.y));
    }
    block_sum3(v, red);
    if (tid == 0) den0[a] = cadd(den0[a], cmul(red[0], w));
    __syncthreads();
  }
  if (tid == 0 && nt == 0) {
    h->n += 1;
    h->sign += zs;
  }
}

}  // namespace

int tau_init(const cplx* G, cplx* G00, cplx* GT0, cplx* GTT, cplx* G0T, int n, int nchain, long long cs,
             hipStream_t s) {
  const long long nn = (long long)n * n;
  hipLaunchKernelGGL(tau_init_kernel, dim3((unsigned)((nn + TT - 1) / TT), 1, nchain), dim3(TT), 0, s, G, G00, GT0,
                     GTT, G0T, n, cs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

int cgr22_build(const cplx* X, const int* pinv, const cplx* UL, const cplx* DL, const cplx* UR, const cplx* DR,
                const cplx* VR, cplx* H, cplx* F1, cplx* F2, int* flag, int n, int nchain, long long cs,
                hipStream_t s) {
  const long long mm = 4LL * n * n;
  hipLaunchKernelGGL(cgr22_build_kernel, dim3((unsigned)((mm + TT - 1) / TT), 1, nchain), dim3(TT), 0, s, X, pinv,
                     UL, DL, UR, DR, VR, H, F1, F2, flag, n, cs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

int cgr22_extract(const cplx* Oinv, const int* flag, const cplx* GR, cplx* G00, cplx* G0T, cplx* GT0, cplx* GTT,
                  double* part, TauCtl* ctl, int n, int nchain, long long cs, hipStream_t s) {
  hipLaunchKernelGGL(cgr22_extract_kernel, dim3(XB, 1, nchain), dim3(XT), 0, s, Oinv, flag, GR, G00, G0T, GT0, GTT,
                     part, n, cs);
  hipLaunchKernelGGL(cgr22_precision_kernel, dim3(1, 1, nchain), dim3(64), 0, s, (const double*)part, ctl, n, cs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

int tau_measure(const ObsDev& o, const cplx* GT0, const cplx* G0T, const cplx* G00, const cplx* GTT,
                const ChainState* st, char* acc, long long tau_cs, long long off_den0, long long off_latt, int nt,
                int ntau, int nchain, long long cs, hipStream_t s) {
  hipLaunchKernelGGL(tau_obs_kernel, dim3((unsigned)(o.ncell + 1), 1, (unsigned)nchain), dim3(TT), 0, s, GT0, G0T,
                     G00, GTT, st, o, acc, tau_cs, off_den0, off_latt, nt, ntau, cs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}
