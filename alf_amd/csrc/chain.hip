// Metropolis sweep over the fields of one time slice with delayed
// Sherman-Morrison updates (the per-op loop of WRAPGRUP/WRAPGRDO with Upgrade2,
// Prog/Wrapgr_mod.F90:115-147, 198-236; Prog/upgrade_mod.F90:105-302), plus
// the diagonal-exponential vectors of a slice and the random-number stream.
//
// Delayed update: after m accepted flips the current Green's function is
//   G_cur = G + X(:, 0:m) Y(:, 0:m)^T,
//   X(:, j) = -x_j G_cur_j(:, P_j),  Y(:, j) = e_P - G_cur_j(P_j, :)^T,
// which is the rank-1 update of upgrade_mod.F90:266-269 (ZGERU) kept
// unapplied.  The block kernel decides from a submatrix of G_cur held in LDS
// and forms the columns of X and Y in bulk per block of visits; the generic
// kernel (deep delays) forms them one accepted flip at a time.  A kernel stops
// after K accepted flips; the flush G += X Y^T is an MFMA GEMM.
//
// Diagonal operator exponentials (the type-1 Op_Wrapup/Op_Wrapdo scalings,
// Operator_mod.F90:767-782, 884-899) are not applied here: for diagonal
// operators they are a similarity transform G -> D G D^-1 that commutes with
// the rank-1 update and leaves G_PP unchanged, so all of a slice's scalings are
// fused into the epilogue of the adjacent hopping GEMM.
#include <hip/hip_ext.h>

#include <mutex>

#include "alfq_chain.h"

namespace {

__device__ inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

// Philox4x32 with 10 rounds (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ inline uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(0xD2511F53u, c.x, hi0, lo0);
    mulhilo(0xCD9E8D57u, c.z, hi1, lo1);
    c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    k.x += 0x9E3779B9u;
    k.y += 0xBB67AE85u;
  }
  return c;
}

// uniforms u[q - base] for global indices q in [base, base + count): one Philox block per index,
// 53-bit mantissa from two words, u in [0, 1).
__global__ void philox_fill_kernel(double* __restrict__ u, unsigned long long base, long long count,
                                   unsigned long long seed, unsigned stream) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const unsigned long long q = base + (unsigned long long)i;
  const uint4 c = make_uint4((uint32_t)q, (uint32_t)(q >> 32), stream, 0x414c4651u);
  const uint2 k = make_uint2((uint32_t)seed, (uint32_t)(seed >> 32));
  const uint4 r = philox4x32_10(c, k);
  const uint64_t w = ((uint64_t)r.x << 32) | r.y;
  u[i] = (double)(w >> 11) * (1.0 / 9007199254740992.0);
}

__device__ inline cplx wave_csum(cplx v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o, 64);
    v.y += __shfl_xor(v.y, o, 64);
  }
  return v;
}

__device__ inline int nranf3(double u) {
  // nranf(3) = clamp(nint(3u + 0.5), 1, 3)   (random_wrap_mod.F90:159-168)
  const double x = u * 3.0 + 0.5;
  int k = (int)floor(x + 0.5);
  return k < 1 ? 1 : (k > 3 ? 3 : k);
}

__device__ inline int flip_st(int s, int k) {
  // FLIP_st(s, k) of Fields_mod.F90:287-301
  switch (s) {
    case -2: return k == 1 ? -1 : (k == 2 ? 1 : 2);
    case -1: return k == 1 ? 1 : (k == 2 ? 2 : -2);
    case 1: return k == 1 ? 2 : (k == 2 ? -2 : -1);
    default: return k == 1 ? -2 : (k == 2 ? -1 : 1);
  }
}

__global__ __launch_bounds__(256) void chain_kernel(ChainArgs a) {
  __shared__ cplx s_xr[CHAIN_KMAX];
  __shared__ cplx s_yr[CHAIN_KMAX];
  __shared__ int s_acc, s_P, s_pos, s_m, s_stop;
  __shared__ cplx s_x;

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int n = a.n;
  ChainState* st = a.st;
  if (tid == 0) {
    s_pos = st->pos;
    s_m = 0;
    s_stop = 0;
  }
  __syncthreads();

  while (true) {
    const int pos = s_pos;
    const int m = s_m;
    if (pos >= a.n_op || m >= a.Kmax || s_stop) break;
    const int nop = a.dir > 0 ? pos : a.n_op - 1 - pos;
    const int P = a.op_site[nop];
    if (wave == 0) {
      cplx d = make_double2(0.0, 0.0);
      for (int j = lane; j < m; j += 64) {
        const cplx xv = a.X[P + (long long)j * n];
        const cplx yv = a.Y[P + (long long)j * n];
        s_xr[j] = xv;
        s_yr[j] = yv;
        d = cfma(xv, yv, d);
      }
      d = wave_csum(d);
      if (lane == 0) {
        const cplx gpp = cadd(a.G[P + (long long)P * n], d);
        const int typ = a.op_type[nop];
        int8_t* fp = a.f + (long long)a.slice * a.n_op + nop;
        const int s_old = *fp;
        unsigned long long q = st->rng_pos;
        const unsigned long long need = (typ == 2) ? 3ull : 2ull;
        int acc = 0;
        if (q < a.u_base || q + need > a.u_base + (unsigned long long)a.u_len) {
          st->status = 1;   // random stream exhausted: stop without touching G
          s_stop = 1;
        } else {
          const double* u = a.u + (q - a.u_base);
          int s_new, k = 0;
          if (typ == 2) { s_new = flip_st(s_old, nranf3(u[0])); k = 1; } else { s_new = -s_old; }
          // u[k]: T0_proposal = 1.5 > ranf, always taken
          const double u3 = u[k + 1];
          st->rng_pos = q + need;
          const int ti = nop * 25 + (s_old + 2) * 5 + (s_new + 2);
          const cplx myexp = a.op_myexp[ti];
          const cplx delta = make_double2(myexp.x - 1.0, myexp.y);
          const cplx dmat = csub(myexp, cmul(delta, gpp));
          const cplx ratio = cmul(dmat, a.op_expalpha[ti]);
          cplx rt = ratio;
          for (int p = 1; p < a.n_sun; ++p) rt = cmul(rt, ratio);
          rt = cscale(rt, a.gama_ratio[ti]);
          const cplx ph = st->phase;
          const double weight = fabs(cmul(ph, rt).x / ph.x);
          st->nc += 1;
          if (weight > u3) {
            acc = 1;
            st->acc += 1;
            st->phase = cmul(ph, cscale(rt, 1.0 / sqrt(cabs2(rt))));
            *fp = (int8_t)s_new;
            // x = Delta / (1 + v_P Delta), v_P = 1 - G_cur(P,P)
            const cplx vp = make_double2(1.0 - gpp.x, -gpp.y);
            const cplx den = cadd(make_double2(1.0, 0.0), cmul(vp, delta));
            s_x = cdiv(delta, den);
          }
        }
        s_acc = acc;
        s_P = P;
      }
    }
    __syncthreads();
    if (s_stop) break;
    if (s_acc) {
      const cplx x = s_x;
      const cplx mx = make_double2(-x.x, -x.y);
      for (int r = tid; r < n; r += blockDim.x) {
        cplx col = a.G[r + (long long)P * n];
        cplx row = a.G[P + (long long)r * n];
        for (int j = 0; j < m; ++j) {
          col = cfma(a.X[r + (long long)j * n], s_yr[j], col);
          row = cfma(s_xr[j], a.Y[r + (long long)j * n], row);
        }
        a.X[r + (long long)m * n] = cmul(mx, col);
        a.Y[r + (long long)m * n] = make_double2((r == P ? 1.0 : 0.0) - row.x, -row.y);
      }
    }
    __syncthreads();
    if (tid == 0) {
      s_pos = pos + 1;
      if (s_acc) s_m = m + 1;
    }
    __syncthreads();
  }
  if (tid == 0) {
    st->pos = s_pos;
    st->m = s_m;
  }
}

// ---------------------------------------------------------------------------------
// Submatrix block kernel.  The visits of a round are taken in blocks of up to SB_V
// consecutive visits.  For the sites S of a block the kernel holds g = G_cur[S, S] in
// LDS and decides each visit from g alone (G_cur(P,P) = g(p,p)); every thread computes
// the decision itself, so a rejected proposal needs no barrier and no global load.
// An accepted flip is the rank-1 update of upgrade_mod.F90:266-269 restricted to S,
//   g += u w^T,  u = -x g(:, p),  w = e_p - g(p, :)^T,
// and u, w are the restrictions to S of the delayed-update columns X(:, j), Y(:, j).
// At the end of the block the full columns follow from the Green's function at the
// block start, G_bs = G + X(:, 0:m0) Y(:, 0:m0)^T, one row per thread:
//   X(:, j) = -x_j [G_bs(:, P_j) + sum_{i<j in block} X(:, i) w_i(p_j)],
//   Y(:, j) = e_{P_j} - [G_bs(P_j, :)^T + sum_{i<j in block} Y(:, i) u_i(p_j)],
// and the next block starts from g = G[S', S'] + X[S', :] Y[S', :]^T.
constexpr int SB_V = 32;   // visits per block: g is SB_V x SB_V
constexpr int SB_A = 16;   // accepted flips per block
constexpr int SB_K = 32;   // accepted flips per round (delayed-update depth)

struct BVisit {
  cplx myexp, expal;        // exp(Z1 E) and exp(Z1 alpha) of the proposal
  double gr, u3;            // Gama(new)/Gama(old), Metropolis uniform
  int P, snew, nop, loc;    // site, proposed field, operator, index of the site in the block
};

__device__ __forceinline__ void block_sync_lds() {
  __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_s_barrier();
  __asm__ volatile("" ::: "memory");
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void chain_block_kernel(ChainArgs a) {
  static_assert(SB_V == 32 && SB_A <= SB_V, "block layout assumes 32 visits and 8 column groups");
  __shared__ cplx gs[2][SB_V * SB_V];           // g(r, c) at [r + SB_V c], double-buffered
  __shared__ cplx Xb[SB_K][SB_V], Yb[SB_K][SB_V]; // X(P_r, i), Y(P_r, i) of the round's earlier flips
  __shared__ cplx Us[SB_A][SB_V], Ws[SB_A][SB_V]; // u_j(r), w_j(r) of the block's flips
  __shared__ BVisit vb[SB_V];
  __shared__ cplx s_xj[SB_A];
  __shared__ int s_aloc[SB_A], s_aP[SB_A];
  __shared__ int s_nvis, s_pos0;
  __shared__ unsigned long long s_base;
  __shared__ cplx s_phase;
  __shared__ long long s_nc, s_accn;

  const int tid = threadIdx.x;
  const int n = a.n;
  const int ga = tid & (SB_V - 1), gq = tid >> 5;   // g entries (ga, gq + 8 q), q = 0..3
  ChainState* st = a.st;
  if (tid == 0) {
    s_pos0 = a.first_round ? 0 : st->pos;
    s_base = a.first_round ? st->rng_pos : st->rng_slice_base;
    s_phase = st->phase;
    s_nc = st->nc;
    s_accn = st->acc;
    if (a.first_round) st->slice_acc = 0;
  }
  __syncthreads();
  const unsigned long long base = s_base;
  int pos = s_pos0, m = 0;
  cplx ph = s_phase;
  double iph = 1.0 / ph.x;
  long long nc = s_nc, accn = s_accn;
  const long long acc_at_start = accn;
  const int kmax = a.Kmax < SB_K ? a.Kmax : SB_K;
  bool stopped = false;

  while (pos < a.n_op) {
    if (m >= kmax) {
      if (!a.final_round) break;
      // last round of the slice and the round is full: flush G += X Y^T in place
      for (int r = tid; r < n; r += 256) {
        for (int c = 0; c < n; ++c) {
          cplx acc = a.G[r + (long long)c * n];
          for (int j = 0; j < m; ++j) acc = cfma(a.X[r + (long long)j * n], a.Y[c + (long long)j * n], acc);
          a.G[r + (long long)c * n] = acc;
        }
      }
      __syncthreads();
      m = 0;
    }
    const int nb = a.n_op - pos < SB_V ? a.n_op - pos : SB_V;
    // ---- the block's proposals, table entries and uniforms (a field changes only at its own visit) ----
    if (tid < SB_V) {
      BVisit w;
      w.myexp = make_double2(1.0, 0.0); w.expal = make_double2(1.0, 0.0);
      w.gr = 1.0; w.u3 = 2.0; w.P = 0; w.snew = 0; w.nop = 0;
      bool ok = false;
      if (tid < nb) {
        const int v = pos + tid;
        const int nop = a.dir > 0 ? v : a.n_op - 1 - v;
        const int typ = a.op_type[nop];
        const int s_old = a.f[(long long)a.slice * a.n_op + nop];
        const unsigned long long q = base + (unsigned long long)a.uoff[v];
        const unsigned long long need = (unsigned long long)(a.uoff[v + 1] - a.uoff[v]);
        ok = !(q < a.u_base || q + need > a.u_base + (unsigned long long)a.u_len);
        int s_new = -s_old;
        if (ok) {
          const double* u = a.u + (q - a.u_base);
          if (typ == 2) { s_new = flip_st(s_old, nranf3(u[0])); w.u3 = u[2]; } else { w.u3 = u[1]; }
        }
        const int ti = nop * 25 + (s_old + 2) * 5 + (s_new + 2);
        w.myexp = a.op_myexp[ti];
        w.expal = a.op_expalpha[ti];
        w.gr = a.gama_ratio[ti];
        w.P = a.op_site[nop];
        w.snew = s_new;
        w.nop = nop;
      }
      // index of the site in the block: its first visit in the block
      int loc = tid;
      for (int t = 0; t < SB_V; ++t) {
        const int Pt = __builtin_amdgcn_readlane(w.P, t);
        if (t < tid && t < nb && tid < nb && Pt == w.P && loc == tid) loc = t;
      }
      w.loc = tid < nb ? loc : 0;
      vb[tid] = w;
      const unsigned long long bad = __ballot(tid < nb && !ok);
      if (tid == 0) s_nvis = bad ? (int)__builtin_ctzll(bad) : nb;
    }
    __syncthreads();
    const int nvis = s_nvis;
    // rows of X, Y at the block's sites for the round's earlier flips, and G at the block's sites
    for (int e = tid; e < SB_V * m; e += 256) {
      const int r = e & (SB_V - 1), i = e >> 5;
      cplx xv = make_double2(0.0, 0.0), yv = make_double2(0.0, 0.0);
      if (r < nb) {
        const int P = vb[r].P;
        xv = a.X[P + (long long)i * n];
        yv = a.Y[P + (long long)i * n];
      }
      Xb[i][r] = xv;
      Yb[i][r] = yv;
    }
    cplx gown[4];
    {
      const int Pa = vb[ga].P;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int b = gq + 8 * q;
        gown[q] = (ga < nb && b < nb) ? a.G[Pa + (long long)vb[b].P * n] : make_double2(0.0, 0.0);
      }
    }
    __syncthreads();
    for (int i = 0; i < m; ++i) {
      const cplx xa = Xb[i][ga];
#pragma unroll
      for (int q = 0; q < 4; ++q) gown[q] = cfma(xa, Yb[i][gq + 8 * q], gown[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) gs[0][ga + SB_V * (gq + 8 * q)] = gown[q];
    __syncthreads();

    // ---- the visits: decisions from g, rank-1 updates of g on acceptance ----
    int buf = 0, mb = 0, k = 0;
    BVisit w = vb[0];
    BVisit wn = vb[1];
    cplx d = gs[0][w.loc * (SB_V + 1)];
    for (; k < nvis; ++k) {
      if (m + mb >= kmax || mb >= SB_A) break;
      const BVisit wnn = vb[k + 2 < SB_V ? k + 2 : SB_V - 1];
      const int p = w.loc, pn = wn.loc;
      const cplx dn_rej = gs[buf][pn * (SB_V + 1)];
      const cplx delta = make_double2(w.myexp.x - 1.0, w.myexp.y);
      const cplx dmat = csub(w.myexp, cmul(delta, d));
      const cplx ratio = cmul(dmat, w.expal);
      cplx rt = ratio;
      for (int s = 1; s < a.n_sun; ++s) rt = cmul(rt, ratio);
      rt = cscale(rt, w.gr);
      const double weight = fabs(cmul(ph, rt).x * iph);
      nc += 1;
      const int accept = __builtin_amdgcn_readfirstlane(weight > w.u3 ? 1 : 0);
      cplx dn = dn_rej;
      if (accept) {
        accn += 1;
        ph = cmul(ph, cscale(rt, 1.0 / sqrt(cabs2(rt))));
        iph = 1.0 / ph.x;
        // x = Delta / (1 + v_P Delta), v_P = 1 - G_cur(P,P)
        const cplx vp = make_double2(1.0 - d.x, -d.y);
        const cplx x = cdiv(delta, cadd(make_double2(1.0, 0.0), cmul(vp, delta)));
        const cplx mx = make_double2(-x.x, -x.y);
        if (tid == 0) {
          a.f[(long long)a.slice * a.n_op + w.nop] = (int8_t)w.snew;
          s_xj[mb] = x;
          s_aloc[mb] = p;
          s_aP[mb] = w.P;
        }
        const cplx ua = cmul(mx, gs[buf][ga + SB_V * p]);
        const cplx g_np = gs[buf][pn + SB_V * p];
        const cplx g_pn = gs[buf][p + SB_V * pn];
        if (gq == 0) Us[mb][ga] = ua;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int b = gq + 8 * q;
          const cplx gpb = gs[buf][p + SB_V * b];
          const cplx wb = make_double2((b == p ? 1.0 : 0.0) - gpb.x, -gpb.y);
          if (ga == 0) Ws[mb][b] = wb;
          gown[q] = cfma(ua, wb, gown[q]);
          gs[buf ^ 1][ga + SB_V * b] = gown[q];
        }
        // the next visit's diagonal entry after this update, as its owner computes it
        dn = cfma(cmul(mx, g_np), make_double2((p == pn ? 1.0 : 0.0) - g_pn.x, -g_pn.y), dn_rej);
        mb += 1;
        buf ^= 1;
        block_sync_lds();
      }
      w = wn;
      wn = wnn;
      d = dn;
    }
    pos += k;
    if (k == nvis && nvis < nb) stopped = true;   // random stream exhausted at this visit

    // ---- full columns X(:, m:m+mb), Y(:, m:m+mb) of the block's flips ----
    if (mb > 0) {
      for (int r = tid; r < n; r += 256) {
        cplx xc[SB_A], yc[SB_A];
#pragma unroll
        for (int j = 0; j < SB_A; ++j) {
          if (j < mb) {
            const int Pj = s_aP[j];
            xc[j] = a.G[r + (long long)Pj * n];
            yc[j] = a.G[Pj + (long long)r * n];
          }
        }
        for (int i = 0; i < m; ++i) {
          const cplx xr = a.X[r + (long long)i * n], yr = a.Y[r + (long long)i * n];
#pragma unroll
          for (int j = 0; j < SB_A; ++j) {
            if (j < mb) {
              const int pj = s_aloc[j];
              xc[j] = cfma(xr, Yb[i][pj], xc[j]);
              yc[j] = cfma(Xb[i][pj], yr, yc[j]);
            }
          }
        }
#pragma unroll
        for (int j = 0; j < SB_A; ++j) {
          if (j < mb) {
            const int pj = s_aloc[j];
#pragma unroll
            for (int i = 0; i < j; ++i) {
              xc[j] = cfma(xc[i], Ws[i][pj], xc[j]);
              yc[j] = cfma(Us[i][pj], yc[i], yc[j]);
            }
            const cplx x = s_xj[j];
            xc[j] = cmul(make_double2(-x.x, -x.y), xc[j]);
            yc[j] = make_double2((r == s_aP[j] ? 1.0 : 0.0) - yc[j].x, -yc[j].y);
            a.X[r + (long long)(m + j) * n] = xc[j];
            a.Y[r + (long long)(m + j) * n] = yc[j];
          }
        }
      }
      m += mb;
    }
    __syncthreads();   // the new columns are visible to the next block, the LDS tables are free
    if (stopped) break;
  }
  if (tid == 0) {
    st->phase = ph;
    st->nc = nc;
    st->acc = accn;
    st->pos = pos;
    st->m = m;
    st->rng_slice_base = base;
    st->rng_pos = base + (unsigned long long)a.uoff[pos];
    st->slice_acc += (int)(accn - acc_at_start);
    if (stopped) {
      st->status = 1;
      st->pos = a.n_op;    // do not revisit
    }
    if (st->pos >= a.n_op && st->slice_acc > st->max_slice_acc) st->max_slice_acc = st->slice_acc;
  }
}

// Diagonal exponentials of slices [s0, s0 + ns): for each site i,
//   d(i) = prod_{ops at i} E_exp(op, f),  dinv(i) = prod E_exp(op, -f),  dconj(i) = conj(d(i)).
__global__ void dvec_kernel(const int8_t* __restrict__ f, int n_op, int s0, int ns, int n,
                            const int* __restrict__ site_ptr, const int* __restrict__ site_ops,
                            const cplx* __restrict__ eexp, cplx* __restrict__ d, cplx* __restrict__ dinv,
                            cplx* __restrict__ dconj) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * ns) return;
  const int i = idx % n, s = idx / n;
  const int slice = s0 + s;
  cplx p = make_double2(1.0, 0.0), q = make_double2(1.0, 0.0);
  for (int t = site_ptr[i]; t < site_ptr[i + 1]; ++t) {
    const int op = site_ops[t];
    const int sp = f[(long long)slice * n_op + op];
    p = cmul(p, eexp[op * 5 + (sp + 2)]);
    q = cmul(q, eexp[op * 5 + (2 - sp)]);
  }
  d[idx] = p;
  dinv[idx] = q;
  dconj[idx] = cconj(p);
}

// Op_phase: prod over (op, slice) of exp(i Im(g alpha phi(f))) as a block product.
__global__ void opphase_kernel(const int8_t* __restrict__ f, long long count, int n_op,
                               const cplx* __restrict__ phasefac, cplx* __restrict__ out) {
  __shared__ cplx sh[256];
  cplx p = make_double2(1.0, 0.0);
  for (long long i = threadIdx.x; i < count; i += blockDim.x) {
    const int op = (int)(i % n_op);
    const int sp = f[i];
    p = cmul(p, phasefac[op * 5 + (sp + 2)]);
  }
  sh[threadIdx.x] = p;
  __syncthreads();
  for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = cmul(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

}  // namespace

int launch_philox_fill(double* u, unsigned long long base, long long count, unsigned long long seed,
                       unsigned stream, hipStream_t s) {
  if (count <= 0) return ST_NONE;
  hipLaunchKernelGGL(philox_fill_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, u, base, count,
                     seed, stream);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

int chain_max_delay(int n) {
  (void)n;
  return SB_K;
}

int launch_chain(const ChainArgs& a, hipStream_t s, KTimer* t) {
  if (a.Kmax < 1 || a.Kmax > CHAIN_KMAX) {
    alfq_set_error("launch_chain: K outside [1, CHAIN_KMAX]");
    return ST_GENERIC;
  }
  if (a.Kmax <= SB_K && a.uoff) {
    if (t && t->on) {
      if (t->used == t->ev.size()) {
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        t->ev.emplace_back(e0, e1);
      }
      auto& pr = t->ev[t->used++];
      hipExtLaunchKernelGGL(chain_block_kernel, dim3(1), dim3(256), 0, s, pr.first, pr.second, 0, a);
    } else {
      hipLaunchKernelGGL(chain_block_kernel, dim3(1), dim3(256), 0, s, a);
    }
    HIP_OK(hipGetLastError());
    return ST_NONE;
  }
  hipLaunchKernelGGL(chain_kernel, dim3(1), dim3(256), 0, s, a);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

int launch_dvec(const int8_t* f, int n_op, int s0, int ns, int n, const int* site_ptr, const int* site_ops,
                const cplx* eexp, cplx* d, cplx* dinv, cplx* dconj, hipStream_t s) {
  const int total = n * ns;
  hipLaunchKernelGGL(dvec_kernel, dim3((total + 255) / 256), dim3(256), 0, s, f, n_op, s0, ns, n, site_ptr,
                     site_ops, eexp, d, dinv, dconj);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

int launch_opphase(const int8_t* f, long long count, int n_op, const cplx* phasefac, cplx* out, hipStream_t s) {
  hipLaunchKernelGGL(opphase_kernel, dim3(1), dim3(256), 0, s, f, count, n_op, phasefac, out);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}
