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
#include <type_traits>

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

// uniforms u[q - base] for global indices q in [base, base + count), base = st->rng_pos: one
// Philox block per index, 53-bit mantissa from two words, u in [0, 1).  The window is read
// from and recorded in device memory, so a captured sweep replays with the stream's position.
__global__ void philox_fill_kernel(double* __restrict__ u, ChainState* st, long long count, long long cs) {
  u = chain_ptr(u, cs);
  st = chain_ptr(st, cs);
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long base = st->rng_pos, seed = st->seed;
  const unsigned stream = st->stream;
  if (i == 0) {
    st->u_base = base;
    st->u_len = count;
  }
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
        if (q < st->u_base || q + need > st->u_base + (unsigned long long)st->u_len) {
          st->status = 1;   // random stream exhausted: stop without touching G
          s_stop = 1;
        } else {
          const double* u = a.u + (q - st->u_base);
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
//
// The full columns are never formed one flip at a time.  With A = (P_0, P_1, ...) the
// sites of the round's accepted flips in order and G the Green's function at the start
// of the round,
//   X = G(:, A) C,   Y = Yhat D,   Yhat(:, j) = e_{P_j} - G(P_j, :)^T,
//   C(:, j) = -x_j (e_j + sum_{i<j} C(:, i) Y(P_j, i)),
//   D(:, j) =        e_j - sum_{i<j} D(:, i) X(P_j, i),
// with C, D upper triangular of the round's size; the entries X(P_j, i), Y(P_j, i)
// are already in LDS (u, w of the block, or the block-start rows below).  A block
// starts from X[S, :] = G[S, A] C, Y[S, :] = Yhat[S, :] D and
// g = G[S, S] + X[S, :] Y[S, :]^T.  At the end of the round the kernel leaves
// W = C D^T, the site list A and Yhat; the flush is G += (G(:, A) W) Yhat^T, two GEMMs
// over the whole device.  All round-sized tables are kept zero outside the round's
// m x m corner, so their sums run in unconditional groups of four.
constexpr int SB_V = 32;    // visits per block: g is SB_V x SB_V
constexpr int SB_A = 16;    // accepted flips per block
constexpr int SB_K = 32;    // accepted flips per round (delayed-update depth)
constexpr int SB_R = 256;   // visit records prepared at once
static_assert(SB_K == CHAIN_W_LD, "W is stored with the round's size as leading dimension");

// One visit.  With Delta = exp(Z1 E) - 1 the single-colour ratio is
// exp(Z1 alpha) (exp(Z1 E) - Delta G(P,P)) = ca - cb G(P,P); the Metropolis test
// |Re(PHASE ratio^N_SUN Gama(new)/Gama(old))| / |Re(PHASE)| > u becomes
// |Re(z ratio^N_SUN)| > thr |Re(z)| with thr = u / |Gama ratio| and z any positive
// multiple of PHASE, so the decision needs no division.
struct BVisit {
  cplx ca, cb, delta;
  double thr;
  int loc;                  // index of the site in its block (the site's first visit there)
  int P;                    // site
};

// z * 2^-e with |z| in [0.5, 1): an exact rescaling that keeps the running phase in range
__device__ __forceinline__ cplx renorm_pow2(cplx z) {
  const double mx = fmax(fabs(z.x), fabs(z.y));
  const int e = __builtin_amdgcn_frexp_exp(mx);
  return make_double2(ldexp(z.x, -e), ldexp(z.y, -e));
}

// the complex value of lane l (l uniform) in every lane
__device__ __forceinline__ cplx lane_c(cplx v, int l) {
  const long long re = __builtin_bit_cast(long long, v.x), im = __builtin_bit_cast(long long, v.y);
  const long long r = ((long long)(unsigned)__builtin_amdgcn_readlane((int)(re >> 32), l) << 32) |
                      (unsigned)__builtin_amdgcn_readlane((int)re, l);
  const long long i = ((long long)(unsigned)__builtin_amdgcn_readlane((int)(im >> 32), l) << 32) |
                      (unsigned)__builtin_amdgcn_readlane((int)im, l);
  return make_double2(__builtin_bit_cast(double, r), __builtin_bit_cast(double, i));
}

__device__ __forceinline__ void block_sync_lds() {
  __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_s_barrier();
  __asm__ volatile("" ::: "memory");
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void chain_block_kernel(ChainArgs a) {
  static_assert(SB_V == 32 && SB_K == 32 && SB_A == 16 && SB_R == 256, "layout of the block kernel");
  __shared__ cplx gs[2][SB_V * SB_V];              // g(r, c) at [r + SB_V c]; gs[1] also stages G[S, A]
  __shared__ cplx Xb[SB_K][SB_V], Yb[SB_K][SB_V];  // X(P_s, i), Y(P_s, i) of the round's earlier flips
  __shared__ cplx UW[2 * SB_A * SB_V];             // u_j(s), w_j(s) of the block's flips; also stages Yhat[S, A]
  __shared__ cplx Cm[SB_K][SB_K], Dm[SB_K][SB_K];  // C(l, j) at Cm[j][l], D(l, j) at Dm[j][l]
  __shared__ BVisit vr[SB_R];                      // visits [vbase, vbase + SB_R)
  __shared__ int2 vaux[SB_R];                      // (proposed field, operator)
  __shared__ cplx s_x[SB_K];
  __shared__ int s_aP[SB_K], s_aloc[SB_A];
  __shared__ int s_pos0, s_stop;
  __shared__ unsigned long long s_base, s_ubase;
  __shared__ long long s_ulen;
  __shared__ cplx s_phase;
  __shared__ long long s_nc, s_accn;
  cplx* const Us = UW;
  cplx* const Ws = UW + SB_A * SB_V;

  const int tid = threadIdx.x;
  const int n = a.n;
  const int ga = tid & (SB_V - 1), gq = tid >> 5;   // g entries (ga, gq + 8 q), q = 0..3
  a.G = chain_ptr(a.G, a.cs);
  a.X = chain_ptr(a.X, a.cs);
  a.Y = chain_ptr(a.Y, a.cs);
  a.f = chain_ptr(a.f, a.cs);
  a.op_myexp = chain_ptr(a.op_myexp, a.cs);
  a.op_expalpha = chain_ptr(a.op_expalpha, a.cs);
  a.gama_ratio = chain_ptr(a.gama_ratio, a.cs);
  a.u = chain_ptr(a.u, a.cs);
  a.st = chain_ptr(a.st, a.cs);
  a.W = chain_ptr(a.W, a.cs);
  a.amap = chain_ptr(a.amap, a.cs);
  ChainState* st = a.st;
  if (a.lz && a.slice >= a.lz[blockIdx.z]) {
    // a slice beyond this chain's Ltrot: no visits, no random numbers, nothing to flush
    if (tid == 0) st->m = 0;
    return;
  }
  if (tid == 0) {
    s_pos0 = a.first_round ? 0 : st->pos;
    s_base = a.first_round ? st->rng_pos : st->rng_slice_base;
    s_phase = st->phase;
    s_nc = st->nc;
    s_accn = st->acc;
    s_ubase = st->u_base;
    s_ulen = st->u_len;
    if (a.first_round) st->slice_acc = 0;
  }
  __syncthreads();
  const unsigned long long base = s_base, u_base = s_ubase;
  const long long u_len = s_ulen;
  int pos = s_pos0, m = 0;
  cplx z = s_phase;         // the phase up to a positive factor
  long long nvisits = 0, naccepts = 0;
  const int kmax = a.Kmax < SB_K ? a.Kmax : SB_K;
  int vbase = 0, stop_at = a.n_op;
  bool have_records = false, stopped = false;

  auto zero_cd = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      Cm[gq + 8 * q][ga] = make_double2(0.0, 0.0);
      Dm[gq + 8 * q][ga] = make_double2(0.0, 0.0);
    }
  };
  // records of visits [vbase, vbase + SB_R): proposal, table entries, uniforms (a field
  // changes only at its own visit, so these do not depend on earlier decisions)
  auto build_records = [&]() {
    if (tid == 0) s_stop = a.n_op;
    __syncthreads();
    const int v = vbase + tid;
    BVisit w;
    w.ca = make_double2(1.0, 0.0); w.cb = make_double2(0.0, 0.0); w.delta = make_double2(0.0, 0.0);
    w.thr = 2.0; w.P = 0; w.loc = 0;
    int2 ax = make_int2(0, 0);
    if (v < a.n_op) {
      const int nop = a.dir > 0 ? v : a.n_op - 1 - v;
      const int typ = a.op_type[nop];
      const int s_old = a.f[(long long)a.slice * a.n_op + nop];
      const unsigned long long q = base + (unsigned long long)a.uoff[v];
      const unsigned long long need = (unsigned long long)(a.uoff[v + 1] - a.uoff[v]);
      const bool ok = !(q < u_base || q + need > u_base + (unsigned long long)u_len);
      int s_new = -s_old;
      double u3 = 2.0;
      if (ok) {
        const double* u = a.u + (q - u_base);
        if (typ == 2) { s_new = flip_st(s_old, nranf3(u[0])); u3 = u[2]; } else { u3 = u[1]; }
      } else {
        atomicMin(&s_stop, v);
      }
      const int ti = nop * 25 + (s_old + 2) * 5 + (s_new + 2);
      const cplx myexp = a.op_myexp[ti], expal = a.op_expalpha[ti];
      w.delta = make_double2(myexp.x - 1.0, myexp.y);
      w.ca = cmul(expal, myexp);
      w.cb = cmul(expal, w.delta);
      w.thr = u3 / fabs(a.gama_ratio[ti]);
      w.P = a.op_site[nop];
      ax = make_int2(s_new, nop);
    }
    vr[tid] = w;
    vaux[tid] = ax;
    __syncthreads();
    stop_at = s_stop;
    have_records = true;
  };
  // W = C D^T into gs[1] (zero outside m x m): W(l, j) = sum_i C(l, i) D(j, i)
  // Column j of C and D is zero for j >= m, so of a thread's columns gq + 8 q only q < live_q(m)
  // can be nonzero in a product with them (the same bound for every lane of the wave).  The
  // products below are instantiated per bound (with_count), so their loops stay unrolled with
  // every load of a step issued ahead of its FMAs.
  auto live_q = [&](int mm) {
    const int q = (mm - 2 * (tid >> 6) + 7) >> 3;
    return __builtin_amdgcn_readfirstlane(q < 0 ? 0 : (q > 4 ? 4 : q));
  };
  auto with_count = [](int c, auto&& f) {
    switch (c) {
      case 1: f(std::integral_constant<int, 1>{}); break;
      case 2: f(std::integral_constant<int, 2>{}); break;
      case 3: f(std::integral_constant<int, 3>{}); break;
      case 4: f(std::integral_constant<int, 4>{}); break;
      default: break;
    }
  };
  auto form_w = [&]() {
    cplx acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = make_double2(0.0, 0.0);
    with_count(live_q(m), [&](auto QC) {
      constexpr int QN = decltype(QC)::value;
      for (int i0 = 0; i0 < m; i0 += 4) {
        cplx cli[4], dv[4][QN];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          cli[t] = Cm[i0 + t][ga];
#pragma unroll
          for (int q = 0; q < QN; ++q) dv[t][q] = Dm[i0 + t][gq + 8 * q];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
          for (int q = 0; q < QN; ++q) acc[q] = cfma(cli[t], dv[t][q], acc[q]);
        }
      }
    });
#pragma unroll
    for (int q = 0; q < 4; ++q) gs[1][ga + SB_K * (gq + 8 * q)] = acc[q];
  };
  // Yhat(r, j) = delta(r, P_j) - G(P_j, r) into Y for j < m, rows of this thread
  auto form_yhat = [&]() {
    for (int r = tid; r < n; r += 256) {
      for (int j0 = 0; j0 < m; j0 += 8) {
        cplx gv[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) gv[t] = a.G[(j0 + t < m ? s_aP[j0 + t] : 0) + (long long)r * n];
#pragma unroll
        for (int t = 0; t < 8; ++t)
          if (j0 + t < m)
            a.Y[r + (long long)(j0 + t) * n] = make_double2((r == s_aP[j0 + t] ? 1.0 : 0.0) - gv[t].x, -gv[t].y);
      }
    }
  };

  zero_cd();
  while (pos < a.n_op) {
    if (m >= kmax) {
      if (!a.final_round) break;
      // last round of the slice and the round is full: G += (G(:, A) W) Yhat^T in place
      form_w();
      form_yhat();
      __syncthreads();
      for (int r = tid; r < n; r += 256) {
        for (int j = 0; j < m; ++j) {
          cplx acc = make_double2(0.0, 0.0);
          for (int l = 0; l < m; ++l) acc = cfma(a.G[r + (long long)s_aP[l] * n], gs[1][l + SB_K * j], acc);
          a.X[r + (long long)j * n] = acc;
        }
      }
      __syncthreads();
      for (int r = tid; r < n; r += 256) {
        for (int c = 0; c < n; ++c) {
          cplx acc = a.G[r + (long long)c * n];
          for (int j = 0; j < m; ++j) acc = cfma(a.X[r + (long long)j * n], a.Y[c + (long long)j * n], acc);
          a.G[r + (long long)c * n] = acc;
        }
      }
      m = 0;
      zero_cd();
      __syncthreads();
    }
    int nb = a.n_op - pos < SB_V ? a.n_op - pos : SB_V;
    if (!have_records || pos + nb > vbase + SB_R) {
      vbase = pos;
      build_records();
    }
    const int r0 = pos - vbase;
    int nvis = stop_at - pos < nb ? stop_at - pos : nb;
    if (nvis < 0) nvis = 0;
    // ---- block start: the sites' indices in the block (wave 0), G at the block's sites,
    //      G[S, A] and Yhat[S, A] for the round's earlier flips ----
    if (tid < SB_V) {
      const int Pw = vr[r0 + (tid < nb ? tid : 0)].P;
      int loc = tid;
      for (int t = 0; t < SB_V; ++t) {
        const int Pt = __builtin_amdgcn_readlane(Pw, t);
        if (t < tid && t < nb && tid < nb && Pt == Pw && loc == tid) loc = t;
      }
      if (tid < nb) vr[r0 + tid].loc = loc;
    }
    cplx gown[4];
    {
      const int Pa = vr[r0 + (ga < nb ? ga : 0)].P;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int b = gq + 8 * q;
        const cplx gv = a.G[Pa + (long long)vr[r0 + (b < nb ? b : 0)].P * n];
        gown[q] = (ga < nb && b < nb) ? gv : make_double2(0.0, 0.0);
      }
    }
    {
      const int s = ga, Ps = vr[r0 + (s < nb ? s : 0)].P;
      cplx gsa[4], gas[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int l = gq + 8 * q;
        const int Pl = l < m ? s_aP[l] : 0;
        gsa[q] = a.G[Ps + (long long)Pl * n];
        gas[q] = a.G[Pl + (long long)Ps * n];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int l = gq + 8 * q;
        const bool live = s < nb && l < m;
        const int Pl = l < m ? s_aP[l] : 0;
        gs[1][s + SB_V * l] = live ? gsa[q] : make_double2(0.0, 0.0);
        UW[s + SB_V * l] = live ? make_double2((Ps == Pl ? 1.0 : 0.0) - gas[q].x, -gas[q].y) : make_double2(0.0, 0.0);
      }
    }
    __syncthreads();
    // X[S, j] = sum_l G[S, A_l] C(l, j),  Y[S, j] = sum_l Yhat[S, A_l] D(l, j)   (zero for j >= m)
    {
      cplx ax[4], ay[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) { ax[q] = make_double2(0.0, 0.0); ay[q] = make_double2(0.0, 0.0); }
      with_count(live_q(m), [&](auto QC) {
        constexpr int QN = decltype(QC)::value;
        for (int l0 = 0; l0 < m; l0 += 4) {
          cplx gl[4], yl[4], cv[4][QN], dv[4][QN];
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            gl[t] = gs[1][ga + SB_V * (l0 + t)];
            yl[t] = UW[ga + SB_V * (l0 + t)];
#pragma unroll
            for (int q = 0; q < QN; ++q) {
              cv[t][q] = Cm[gq + 8 * q][l0 + t];
              dv[t][q] = Dm[gq + 8 * q][l0 + t];
            }
          }
#pragma unroll
          for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int q = 0; q < QN; ++q) {
              ax[q] = cfma(gl[t], cv[t][q], ax[q]);
              ay[q] = cfma(yl[t], dv[t][q], ay[q]);
            }
          }
        }
      });
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        Xb[gq + 8 * q][ga] = ax[q];
        Yb[gq + 8 * q][ga] = ay[q];
      }
    }
    __syncthreads();
    for (int i0 = 0; i0 < m; i0 += 4) {
      cplx xa[4], yv[4][4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        xa[t] = Xb[i0 + t][ga];
#pragma unroll
        for (int q = 0; q < 4; ++q) yv[t][q] = Yb[i0 + t][gq + 8 * q];
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int q = 0; q < 4; ++q) gown[q] = cfma(xa[t], yv[t][q], gown[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) gs[0][ga + SB_V * (gq + 8 * q)] = gown[q];
    __syncthreads();

    // ---- the visits: decisions from g, rank-1 updates of g on acceptance ----
    // A rejected proposal leaves g and the phase unchanged, so the decisions of all visits
    // up to the next acceptance follow from the same g: lane j (= ga) decides visit j of
    // the block with the record and diagonal entry of that visit, and the first visit at
    // or after k whose test passes is the next acceptance (every one before it is
    // rejected).  Each lane also forms its visit's x = Delta / (1 + v_P Delta), so the
    // division runs beside the test.  The arithmetic of each decision is that of a
    // sequential visit with the same g(p, p) and phase.
    int buf = 0, mb = 0, k = 0;
    const BVisit rec = vr[r0 + (ga < nb ? ga : nb - 1)];
    const int2 rec_aux = vaux[r0 + (ga < nb ? ga : nb - 1)];
    const bool lane_live = ga < nvis;
    while (k < nvis) {
      const cplx d = gs[buf][rec.loc * (SB_V + 1)];
      const cplx ratio = make_double2(fma(-rec.cb.x, d.x, fma(rec.cb.y, d.y, rec.ca.x)),
                                      fma(-rec.cb.x, d.y, fma(-rec.cb.y, d.x, rec.ca.y)));
      cplx rt = ratio;
      if (a.n_sun == 2) rt = cmul(ratio, ratio);
      else for (int s = 1; s < a.n_sun; ++s) rt = cmul(rt, ratio);
      const double lhs = fabs(fma(z.x, rt.x, -z.y * rt.y));
      // x = Delta / (1 + v_P Delta), v_P = 1 - G_cur(P,P)
      const cplx delta = rec.delta;
      const cplx den = make_double2(1.0 + (1.0 - d.x) * delta.x + d.y * delta.y,
                                    (1.0 - d.x) * delta.y - d.y * delta.x);
      const double rden = 1.0 / cabs2(den);
      const cplx xl = cscale(cmul(delta, cconj(den)), rden);
      // keep the division here, beside the test, instead of after the ballot
      __asm__ volatile("" ::"v"(xl.x), "v"(xl.y));
      const bool pass = lane_live && ga >= k && lhs > rec.thr * fabs(z.x);
      const unsigned lo = (unsigned)__ballot(pass);   // lanes 0..31: visits 0..31 of the block
      if (lo == 0) {
        k = nvis;
        break;
      }
      const int js = __builtin_amdgcn_readfirstlane(__builtin_ctz(lo));
      k = js + 1;
      z = renorm_pow2(cmul(z, lane_c(rt, js)));
      const cplx x = lane_c(xl, js);
      const int p = __builtin_amdgcn_readlane(rec.loc, js);
      const int P_acc = __builtin_amdgcn_readlane(rec.P, js);
      const int f_new = __builtin_amdgcn_readlane(rec_aux.x, js), op_acc = __builtin_amdgcn_readlane(rec_aux.y, js);
      const cplx mx = make_double2(-x.x, -x.y);
      const cplx gcol = gs[buf][ga + SB_V * p];
      cplx gpb[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) gpb[q] = gs[buf][p + SB_V * (gq + 8 * q)];
      const cplx ua = cmul(mx, gcol);
      cplx wb[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int b = gq + 8 * q;
        wb[q] = make_double2((b == p ? 1.0 : 0.0) - gpb[q].x, -gpb[q].y);
        gown[q] = cfma(ua, wb[q], gown[q]);
        gs[buf ^ 1][ga + SB_V * b] = gown[q];
      }
      if (gq == 0) Us[mb * SB_V + ga] = ua;
      if (ga == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) Ws[mb * SB_V + gq + 8 * q] = wb[q];
        if (tid == 0) {
          a.f[(long long)a.slice * a.n_op + op_acc] = (int8_t)f_new;
          s_x[m + mb] = x;
          s_aP[m + mb] = P_acc;
          s_aloc[mb] = p;
        }
      }
      mb += 1;
      buf ^= 1;
      block_sync_lds();
      if (m + mb >= kmax || mb >= SB_A) break;
    }
    nvisits += k;
    naccepts += mb;
    pos += k;
    if (k == nvis && pos == stop_at && stop_at < a.n_op) stopped = true;   // random stream exhausted here

    // ---- extend C and D by the block's flips ----
    if (mb > 0) {
      __syncthreads();   // every wave is past its last visit (gs[1] is reused below)
      // (1) the part over the round's earlier flips, for every (row l, new column jj):
      //     pre(l, jj) = sum_{i < m} M(l, i) L(i, p_jj), L = Y(P, i) for C and X(P, i) for D;
      //     threads: matrix (tid / 128), row l, four columns jj = c4 + 4 t each
      cplx* const Pre = gs[1];   // [matrix][jj][l]
      {
        const int mat = tid >> 7, l = (tid >> 2) & (SB_K - 1), c4 = tid & 3;
        cplx(*Mm)[SB_K] = mat ? Dm : Cm;
        cplx(*Lold)[SB_V] = mat ? Xb : Yb;
        // columns jj < 4 TN, TN = ceil(mb / 4), cover the block's flips
        with_count(__builtin_amdgcn_readfirstlane((mb + 3) >> 2), [&](auto TC) {
          constexpr int TN = decltype(TC)::value;
          cplx acc[TN];
          int pj[TN];
#pragma unroll
          for (int t = 0; t < TN; ++t) {
            acc[t] = make_double2(0.0, 0.0);
            const int jj = c4 + 4 * t;
            pj[t] = s_aloc[jj < mb ? jj : 0];
          }
          for (int i0 = 0; i0 < m; i0 += 4) {
            cplx mli[4], lv[4][TN];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              mli[u] = Mm[i0 + u][l];
#pragma unroll
              for (int t = 0; t < TN; ++t) lv[u][t] = Lold[i0 + u][pj[t]];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
#pragma unroll
              for (int t = 0; t < TN; ++t) acc[t] = cfma(mli[u], lv[u][t], acc[t]);
            }
          }
#pragma unroll
          for (int t = 0; t < TN; ++t) Pre[(mat * SB_A + c4 + 4 * t) * SB_K + l] = acc[t];
        });
      }
      __syncthreads();
      // (2) the in-block recurrence, one thread per (matrix, row), new entries in registers.
      //     Right-looking: once entry jj is final its terms go into the sums of every later
      //     column, and each table entry of row jj, once used, is replaced by the same column's
      //     entry of row jj + 1, so the next row's reads are in flight while this row's terms
      //     are added.  Each sum still adds its terms in the order ii = 0, 1, ...  Columns up to
      //     4 ceil(mb / 4) are carried without branches; those at or beyond mb are not stored.
      if (tid < 2 * SB_K) {
        const int mat = tid >> 5, l = tid & (SB_K - 1);
        const cplx* Lnew = mat ? Us : Ws;   // X(P_j, i) = u_i(p_j) / Y(P_j, i) = w_i(p_j)
        cplx(*Mm)[SB_K] = mat ? Dm : Cm;
        with_count(__builtin_amdgcn_readfirstlane((mb + 3) >> 2), [&](auto TC) {
          constexpr int JN = 4 * decltype(TC)::value;
          cplx acc[JN], lr[JN];
          int pj[JN];
#pragma unroll
          for (int jj = 0; jj < JN; ++jj) {
            pj[jj] = s_aloc[jj < mb ? jj : 0];
            acc[jj] = Pre[(mat * SB_A + jj) * SB_K + l];
          }
#pragma unroll
          for (int kk = 1; kk < JN; ++kk) lr[kk] = Lnew[pj[kk]];
#pragma unroll
          for (int jj = 0; jj < JN; ++jj) {
            const int j = m + jj;
            const cplx xj = s_x[j < SB_K ? j : SB_K - 1];
            const double dl = l == j ? 1.0 : 0.0;
            const cplx vd = make_double2(dl - acc[jj].x, -acc[jj].y);
            const cplx vc = cmul(make_double2(-xj.x, -xj.y), make_double2(dl + acc[jj].x, acc[jj].y));
            const cplx v = (jj < mb && l <= j) ? (mat ? vd : vc) : make_double2(0.0, 0.0);
            if (jj < mb) Mm[j][l] = v;
#pragma unroll
            for (int kk = jj + 1; kk < JN; ++kk) {
              acc[kk] = cfma(v, lr[kk], acc[kk]);
              if (kk > jj + 1) lr[kk] = Lnew[(jj + 1) * SB_V + pj[kk]];
            }
            __builtin_amdgcn_sched_barrier(0);
          }
        });
      }
      m += mb;
    }
    __syncthreads();   // the LDS tables of the block are free, C and D are complete
    if (stopped) break;
  }
  // ---- the round's flush operands: W = C D^T, the site list, Yhat ----
  form_w();
  if (!a.yhat_external) form_yhat();
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) a.W[ga + SB_K * (gq + 8 * q)] = gs[1][ga + SB_K * (gq + 8 * q)];
  if (tid < SB_K) a.amap[tid] = tid < m ? s_aP[tid] : 0;
  if (tid == 0) {
    st->phase = cscale(z, 1.0 / sqrt(cabs2(z)));
    st->nc = s_nc + nvisits;
    st->acc = s_accn + naccepts;
    st->pos = pos;
    st->m = m;
    st->rng_slice_base = base;
    st->rng_pos = base + (unsigned long long)a.uoff[pos];
    st->slice_acc += (int)naccepts;
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
                            cplx* __restrict__ dconj, long long cs) {
  f = chain_ptr(f, cs);
  eexp = chain_ptr(eexp, cs);
  d = chain_ptr(d, cs);
  dinv = chain_ptr(dinv, cs);
  dconj = chain_ptr(dconj, cs);
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

// Op_phase: prod over (op, slice) of exp(i Im(g alpha phi(f))) as a block product.  A thread
// takes operators op = tid, tid + 256, ... over all count / n_op slices, as four partial
// products over the slices (so the factor loads run ahead of the multiplications), combined
// in a fixed order.
__global__ void opphase_kernel(const int8_t* __restrict__ f, long long count, int n_op,
                               const cplx* __restrict__ phasefac, cplx* __restrict__ out, long long cs) {
  __shared__ cplx sh[256];
  f = chain_ptr(f, cs);
  phasefac = chain_ptr(phasefac, cs);
  out = chain_ptr(out, cs);
  const long long nsl = count / n_op;
  cplx p = make_double2(1.0, 0.0);
  for (int op = threadIdx.x; op < n_op; op += blockDim.x) {
    const cplx* pf = phasefac + op * 5 + 2;
    cplx q[4] = {make_double2(1.0, 0.0), make_double2(1.0, 0.0), make_double2(1.0, 0.0), make_double2(1.0, 0.0)};
    long long sl = 0;
    for (; sl + 4 <= nsl; sl += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) q[u] = cmul(q[u], pf[f[(sl + u) * n_op + op]]);
    }
    for (; sl < nsl; ++sl) q[0] = cmul(q[0], pf[f[sl * n_op + op]]);
    p = cmul(p, cmul(cmul(q[0], q[1]), cmul(q[2], q[3])));
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

int launch_philox_fill(double* u, ChainState* st, long long count, int nchain, long long cs, hipStream_t s) {
  if (count <= 0) return ST_NONE;
  hipLaunchKernelGGL(philox_fill_kernel, dim3((unsigned)((count + 255) / 256), 1, nchain), dim3(256), 0, s, u, st,
                     count, cs);
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
        // no system-scope release when the timing events are recorded: they bracket one
        // kernel on the device and nothing on the host reads its results in between
        HIP_OK(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
        HIP_OK(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
        t->ev.emplace_back(e0, e1);
      }
      auto& pr = t->ev[t->used++];
      hipExtLaunchKernelGGL(chain_block_kernel, dim3(1, 1, a.nchain), dim3(256), 0, s, pr.first, pr.second, 0, a);
    } else {
      hipLaunchKernelGGL(chain_block_kernel, dim3(1, 1, a.nchain), dim3(256), 0, s, a);
    }
    HIP_OK(hipGetLastError());
    return ST_NONE;
  }
  if (a.nchain > 1) {
    alfq_set_error("launch_chain: delays beyond the block kernel's depth are supported for single chains only");
    return ST_GENERIC;
  }
  hipLaunchKernelGGL(chain_kernel, dim3(1), dim3(256), 0, s, a);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

int launch_dvec(const int8_t* f, int n_op, int s0, int ns, int n, const int* site_ptr, const int* site_ops,
                const cplx* eexp, cplx* d, cplx* dinv, cplx* dconj, int nchain, long long cs, hipStream_t s) {
  const int total = n * ns;
  hipLaunchKernelGGL(dvec_kernel, dim3((total + 255) / 256, 1, nchain), dim3(256), 0, s, f, n_op, s0, ns, n,
                     site_ptr, site_ops, eexp, d, dinv, dconj, cs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

int launch_opphase(const int8_t* f, long long count, int n_op, const cplx* phasefac, cplx* out, int nchain,
                   long long cs, hipStream_t s) {
  hipLaunchKernelGGL(opphase_kernel, dim3(1, 1, nchain), dim3(256), 0, s, f, count, n_op, phasefac, out, cs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}
