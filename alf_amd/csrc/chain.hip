// Metropolis sweep over the fields of one time slice with delayed
// Sherman-Morrison updates (the per-op loop of WRAPGRUP/WRAPGRDO with Upgrade2,
// Prog/Wrapgr_mod.F90:115-147, 198-236; Prog/upgrade_mod.F90:105-302), plus
// the diagonal-exponential vectors of a slice and the random-number stream.
//
// Delayed update: after m accepted flips the current Green's function is
//   G_cur = G + X(:, 0:m) Y(:, 0:m)^T,
//   X(:, j) = -x_j G_cur_j(:, P_j),  Y(:, j) = e_P - G_cur_j(P_j, :)^T,
// which is the rank-1 update of upgrade_mod.F90:266-269 (ZGERU) kept
// unapplied.  A rejected proposal only needs G_cur(P,P) = G(P,P) + X(P,:).Y(P,:);
// an accepted one materialises the new column/row pair in O(n m).  The kernel
// stops after K accepted flips; the flush G += X Y^T is an MFMA GEMM.
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
// Register-resident chain kernel.  Thread t owns rows t, t + 256, ... (RPT rows) and
// keeps X(r, 0:KT), Y(r, 0:KT) of its rows in VGPRs.  Everything about a visit that
// does not depend on earlier decisions -- the proposal (the field of an operator
// changes only at its own visit), the table entries, the uniforms, the stale G_PP --
// is precomputed into LDS in a parallel prologue.  The sequential part per visit:
// the owner of row P forms G_cur(P,P) from its registers, decides, and publishes
// (accept, x, X(P,:), Y(P,:)) in a double-buffered LDS slot; one barrier; on accept
// every thread extends its rows of X and Y.  Column and row P of the stale G are
// prefetched one visit ahead.
struct Visit {
  cplx myexp, expal, gpp;
  double gr, u3;
  int P, snew, nop, ok;
};

__device__ __forceinline__ void block_sync_lds() {
  __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_s_barrier();
  __asm__ volatile("" ::: "memory");
}

template <int RPT, int KT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void chain2_kernel(ChainArgs a) {
  extern __shared__ char smem_raw[];
  Visit* vis = (Visit*)smem_raw;                        // n_op entries
  __shared__ cplx sX[2][KT], sY[2][KT];
  __shared__ cplx s_x[2];
  __shared__ int s_acc[2];
  __shared__ cplx s_phase;
  __shared__ double s_iph;                              // 1 / Re(PHASE)
  __shared__ long long s_nc, s_accn;
  __shared__ int s_stop_at, s_pos0;
  __shared__ unsigned long long s_base;

  const int tid = threadIdx.x;
  const int n = a.n;
  ChainState* st = a.st;
  long long acc_at_start = 0;
  if (tid == 0) {
    s_pos0 = a.first_round ? 0 : st->pos;
    s_base = a.first_round ? st->rng_pos : st->rng_slice_base;
    s_phase = st->phase;
    s_iph = 1.0 / st->phase.x;
    s_nc = st->nc;
    s_accn = st->acc;
    acc_at_start = st->acc;
    if (a.first_round) st->slice_acc = 0;
  }
  __syncthreads();
  const unsigned long long base = s_base;
  int pos = s_pos0, m = 0;
  const int kmax = a.Kmax < KT ? a.Kmax : KT;

  cplx Xr[RPT][KT], Yr[RPT][KT], dcorr[RPT];
#pragma unroll
  for (int rr = 0; rr < RPT; ++rr) {
    dcorr[rr] = make_double2(0.0, 0.0);
#pragma unroll
    for (int j = 0; j < KT; ++j) { Xr[rr][j] = make_double2(0.0, 0.0); Yr[rr][j] = make_double2(0.0, 0.0); }
  }
  // column and row P of the stale G for the next PD visits, one ring slot per visit
  constexpr int PD = (RPT == 1) ? 4 : 2;
  cplx rc[PD][RPT], rw[PD][RPT];
  int stop_at = a.n_op;
  int buf = 0;
  bool prologue = true;
  while (true) {
    if (prologue) {
      // ---- per-visit proposal, tables, uniforms and stale G_PP for visits pos.. ----
      for (int v = pos + tid; v < a.n_op; v += blockDim.x) {
        Visit w;
        const int nop = a.dir > 0 ? v : a.n_op - 1 - v;
        const int typ = a.op_type[nop];
        const int s_old = a.f[(long long)a.slice * a.n_op + nop];
        const unsigned long long q = base + (unsigned long long)a.uoff[v];
        const unsigned long long need = (unsigned long long)(a.uoff[v + 1] - a.uoff[v]);
        w.ok = !(q < a.u_base || q + need > a.u_base + (unsigned long long)a.u_len);
        int s_new = -s_old;
        double u3 = 2.0;
        if (w.ok) {
          const double* u = a.u + (q - a.u_base);
          if (typ == 2) { s_new = flip_st(s_old, nranf3(u[0])); u3 = u[2]; } else { u3 = u[1]; }
        }
        const int ti = nop * 25 + (s_old + 2) * 5 + (s_new + 2);
        const int P = a.op_site[nop];
        w.myexp = a.op_myexp[ti];
        w.expal = a.op_expalpha[ti];
        w.gr = a.gama_ratio[ti];
        w.u3 = u3;
        w.P = P;
        w.snew = s_new;
        w.nop = nop;
        w.gpp = a.G[P + (long long)P * n];
        vis[v] = w;
      }
      __syncthreads();
      if (tid == 0) {
        s_stop_at = a.n_op;
        for (int v = pos; v < a.n_op; ++v)
          if (!vis[v].ok) { s_stop_at = v; break; }
      }
      __syncthreads();
      stop_at = s_stop_at;
      if (pos < stop_at) {
#pragma unroll
        for (int S = 0; S < PD; ++S) {
          const int v = pos + S < stop_at ? pos + S : stop_at - 1;
          const int P = vis[v].P;
#pragma unroll
          for (int rr = 0; rr < RPT; ++rr) {
            const int r = tid + 256 * rr;
            const int rcl = r < n ? r : n - 1;
            rc[S][rr] = a.G[rcl + (long long)P * n];
            rw[S][rr] = a.G[P + (long long)rcl * n];
          }
        }
      }
      prologue = false;
    }
    // One visit: the owner of row P decides; on acceptance every thread extends its rows
    // of X, Y (and the running diagonal correction) using column/row P of the stale G
    // held in COLV/ROWV.
#define ALFQ_CHAIN_VISIT(COLV, ROWV)                                                            \
    {                                                                                         \
      const Visit w = vis[pos];                                                               \
      const int P = w.P;                                                                      \
      const int owner = P & 255, orr = P >> 8;                                                \
      if (tid == owner) {                                                                     \
        cplx d = w.gpp;                                                                       \
        _Pragma("unroll") for (int rr = 0; rr < RPT; ++rr) if (rr == orr) d = cadd(d, dcorr[rr]); \
        const cplx delta = make_double2(w.myexp.x - 1.0, w.myexp.y);                          \
        const cplx dmat = csub(w.myexp, cmul(delta, d));                                      \
        const cplx ratio = cmul(dmat, w.expal);                                               \
        cplx rt = ratio;                                                                      \
        for (int p = 1; p < a.n_sun; ++p) rt = cmul(rt, ratio);                               \
        rt = cscale(rt, w.gr);                                                                \
        const cplx ph = s_phase;                                                              \
        const double weight = fabs(cmul(ph, rt).x * s_iph);                                   \
        s_nc += 1;                                                                            \
        int acc = 0;                                                                          \
        if (weight > w.u3) {                                                                  \
          acc = 1;                                                                            \
          s_accn += 1;                                                                        \
          const cplx nph = cmul(ph, cscale(rt, 1.0 / sqrt(cabs2(rt))));                       \
          s_phase = nph;                                                                      \
          s_iph = 1.0 / nph.x;                                                                \
          a.f[(long long)a.slice * a.n_op + w.nop] = (int8_t)w.snew;                          \
          const cplx vp = make_double2(1.0 - d.x, -d.y);                                      \
          s_x[buf] = cdiv(delta, cadd(make_double2(1.0, 0.0), cmul(vp, delta)));              \
          _Pragma("unroll") for (int rr = 0; rr < RPT; ++rr) {                                \
            if (rr == orr) {                                                                  \
              _Pragma("unroll") for (int j = 0; j < KT; ++j) {                                \
                if (j < m) { sX[buf][j] = Xr[rr][j]; sY[buf][j] = Yr[rr][j]; }                \
              }                                                                               \
            }                                                                                 \
          }                                                                                   \
        }                                                                                     \
        s_acc[buf] = acc;                                                                     \
      }                                                                                       \
      block_sync_lds();                                                                       \
      if (s_acc[buf]) {                                                                       \
        const cplx x = s_x[buf];                                                              \
        const cplx mx = make_double2(-x.x, -x.y);                                             \
        _Pragma("unroll") for (int rr = 0; rr < RPT; ++rr) {                                  \
          const int r = tid + 256 * rr;                                                       \
          cplx col = COLV[rr], row = ROWV[rr];                                                \
          _Pragma("unroll") for (int j = 0; j < KT; ++j) {                                    \
            if (j < m) {                                                                      \
              col = cfma(Xr[rr][j], sY[buf][j], col);                                         \
              row = cfma(sX[buf][j], Yr[rr][j], row);                                         \
            }                                                                                 \
          }                                                                                   \
          const cplx xn = cmul(mx, col);                                                      \
          const cplx yn = make_double2((r == P ? 1.0 : 0.0) - row.x, -row.y);                 \
          dcorr[rr] = cfma(xn, yn, dcorr[rr]);                                                \
          _Pragma("unroll") for (int j = 0; j < KT; ++j)                                      \
            if (j == m) { Xr[rr][j] = xn; Yr[rr][j] = yn; }                                   \
        }                                                                                     \
        m = __builtin_amdgcn_readfirstlane(m + 1);                                            \
      }                                                                                       \
      pos = __builtin_amdgcn_readfirstlane(pos + 1);                                          \
      buf ^= 1;                                                                               \
    }

    // fast path: PD visits per iteration with no per-visit guard, so the refills of the
    // ring stay in flight across iterations; taken while PD more visits and accepts fit
    while (pos + PD <= stop_at && m + PD <= kmax) {
#pragma unroll
      for (int S = 0; S < PD; ++S) {
        ALFQ_CHAIN_VISIT(rc[S], rw[S])
        {
          const int v = pos - 1 + PD < stop_at ? pos - 1 + PD : stop_at - 1;
          const int Pn = vis[v].P;
#pragma unroll
          for (int rr = 0; rr < RPT; ++rr) {
            const int rcl = (tid + 256 * rr) < n ? tid + 256 * rr : n - 1;
            rc[S][rr] = a.G[rcl + (long long)Pn * n];
            rw[S][rr] = a.G[Pn + (long long)rcl * n];
          }
        }
      }
    }
    // slow path for the last few visits or accepts of the round: direct loads
    while (pos < stop_at && m < kmax) {
      cplx dc[RPT], dr[RPT];
      const int P = vis[pos].P;
#pragma unroll
      for (int rr = 0; rr < RPT; ++rr) {
        const int rcl = (tid + 256 * rr) < n ? tid + 256 * rr : n - 1;
        dc[rr] = a.G[rcl + (long long)P * n];
        dr[rr] = a.G[P + (long long)rcl * n];
      }
      ALFQ_CHAIN_VISIT(dc, dr)
    }
#undef ALFQ_CHAIN_VISIT
    if (pos >= stop_at || !a.final_round) break;
    // ---- last round of the slice and K accepts pending: flush G += X Y^T in place ----
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
      const int r = tid + 256 * rr;
      if (r < n) {
#pragma unroll
        for (int j = 0; j < KT; ++j)
          if (j < m) a.Y[r + (long long)j * n] = Yr[rr][j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
      const int r = tid + 256 * rr;
      if (r < n) {
        for (int c = 0; c < n; ++c) {
          cplx acc = a.G[r + (long long)c * n];
#pragma unroll
          for (int j = 0; j < KT; ++j)
            if (j < m) acc = cfma(Xr[rr][j], a.Y[c + (long long)j * n], acc);
          a.G[r + (long long)c * n] = acc;
        }
        dcorr[rr] = make_double2(0.0, 0.0);
      }
    }
    __syncthreads();
    m = 0;
    prologue = true;
  }
  // ---- hand the pending columns to the flush GEMM and save the chain state ----
#pragma unroll
  for (int rr = 0; rr < RPT; ++rr) {
    const int r = tid + 256 * rr;
    if (r < n) {
#pragma unroll
      for (int j = 0; j < KT; ++j) {
        if (j < m) {
          a.X[r + (long long)j * n] = Xr[rr][j];
          a.Y[r + (long long)j * n] = Yr[rr][j];
        }
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    st->phase = s_phase;
    st->nc = s_nc;
    st->acc = s_accn;
    st->pos = pos;
    st->m = m;
    st->rng_slice_base = base;
    st->rng_pos = base + (unsigned long long)a.uoff[pos];
    st->slice_acc += (int)(s_accn - acc_at_start);
    if (pos == stop_at && stop_at < a.n_op) {
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
  if (n <= 256) return 32;
  if (n <= 512) return 16;
  if (n <= 1024) return 8;
  if (n <= 2048) return 4;
  return 0;
}

template <int RPT, int KT>
static int launch_chain2(const ChainArgs& a, hipStream_t s, KTimer* t) {
  const size_t lds = sizeof(Visit) * (size_t)(a.n_op > 0 ? a.n_op : 1);
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [] {
    attr_err = hipFuncSetAttribute((const void*)chain2_kernel<RPT, KT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   96 * 1024);
  });
  HIP_OK(attr_err);
  if (lds > 96 * 1024) { alfq_set_error("launch_chain: too many operators per slice for the LDS tables"); return ST_GENERIC; }
  if (t && t->on) {
    if (t->used == t->ev.size()) {
      hipEvent_t e0, e1;
      HIP_OK(hipEventCreate(&e0));
      HIP_OK(hipEventCreate(&e1));
      t->ev.emplace_back(e0, e1);
    }
    auto& pr = t->ev[t->used++];
    hipExtLaunchKernelGGL((chain2_kernel<RPT, KT>), dim3(1), dim3(256), lds, s, pr.first, pr.second, 0, a);
  } else {
    hipLaunchKernelGGL((chain2_kernel<RPT, KT>), dim3(1), dim3(256), lds, s, a);
  }
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

int launch_chain(const ChainArgs& a, hipStream_t s, KTimer* t) {
  if (a.Kmax < 1 || a.Kmax > CHAIN_KMAX) {
    alfq_set_error("launch_chain: K outside [1, CHAIN_KMAX]");
    return ST_GENERIC;
  }
  const int kr = chain_max_delay(a.n);
  if (kr > 0 && a.Kmax <= kr && a.uoff) {
    if (a.n <= 256) return launch_chain2<1, 32>(a, s, t);
    if (a.n <= 512) return launch_chain2<2, 16>(a, s, t);
    if (a.n <= 1024) return launch_chain2<4, 8>(a, s, t);
    return launch_chain2<8, 4>(a, s, t);
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
