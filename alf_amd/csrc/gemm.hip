// FP64 MFMA GEMM for gfx950: real x complex and complex x complex products with
// fused diagonal scalings, gathers and conjugation.
//
// Tiling: a workgroup of 4 waves computes a (16 WM) x (16 WN) complex output tile;
// the waves are arranged WM x WN x KS, so with KS > 1 the waves of one tile split
// K and their partial tiles are summed through LDS in a fixed order (the result
// does not depend on scheduling).  Each wave holds a 16 x 16 sub-tile as two
// (re, im) v_mfma_f64_16x16x4f64 accumulators (three for complex x complex).  K advances in stages of 16 KS
// through LDS, where the operands are split into real and imaginary planes so each
// MFMA consumes one plane; a complex x complex product costs 3 MFMAs per k-step of 4 (3M:
// Ar Br, Ai Bi and (Ar + Ai)(Br + Bi), combined after the K loop), a real x complex one 2.
// The host picks the tile shape per launch: at Ndim = 256 a 32 x 32 tiling leaves most of
// the 256 CUs idle, a 16 x 16 tiling with KS = 4 fills them with a quarter of the per-wave K.
//
// This one kernel carries every dense product of the sweep:
//   * hopping wraps  G <- D exp(-dtau T) G exp(dtau T) D^-1       (real T: 2 MFMA/step)
//   * UDV chain      U <- D exp(-dtau T) U,  U <- exp(-dtau T) conj(D) U
//   * delayed-update flush  G += X Y^T
//   * QR trailing updates / Q formation / TRMM / CGR products
#include "alfq_common.h"

namespace {

constexpr int NT = 256;

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// The two MFMA arithmetics of the kernels below (GemmArgs::prec 0 and 1).  Both take A as
// lane = row + 16 k and B as lane = column + 16 k (k = 0..3); they differ in where the
// 16 x 16 result lands:
//   FP64 v_mfma_f64_16x16x4f64: D(i, j) in lane j + 16 (i % 4), register i / 4
//   FP32 v_mfma_f32_16x16x4f32: D(i, j) in lane j + 16 (i / 4), register i % 4
// With FP32 the operands are rounded to float when staged and the products accumulate in
// float; scalings, alpha and beta are applied to the result in double.
template <bool F32> struct Mma;
template <> struct Mma<false> {
  typedef double E;
  typedef v4d Acc;
  __device__ __forceinline__ static Acc mac(E a, E b, Acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ __forceinline__ static int row(int r, int lk) { return 4 * r + lk; }
};
template <> struct Mma<true> {
  typedef float E;
  typedef v4f Acc;
  __device__ __forceinline__ static Acc mac(E a, E b, Acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ __forceinline__ static int row(int r, int lk) { return 4 * lk + r; }
};

// moves every operand pointer to the arena of the chain bz / batch; true when the product is
// the identity slice of a chain beyond its own Ltrot (the hopping operand is the identity) and
// the workgroup can write its tile as a copy of the other operand (identity_tile)
__device__ __forceinline__ bool chain_operands(GemmArgs& g, int bz) {
  if (g.nchain <= 1) return false;
  const long long off = (long long)(bz / (g.batch > 1 ? g.batch : 1)) * g.cstride;
  auto mv = [off](auto p) { return p ? (decltype(p))((const char*)p + off) : p; };
  g.A = mv(g.A); g.B = mv(g.B); g.C = (void*)mv((const char*)g.C);
  g.rs = mv(g.rs); g.cs = mv(g.cs); g.bs = mv(g.bs); g.rs_real = mv(g.rs_real);
  g.amap = mv(g.amap); g.bmap = mv(g.bmap); g.kdev = mv(g.kdev);
  g.yhat = mv(g.yhat);
  if (g.lz && g.slice >= g.lz[bz / (g.batch > 1 ? g.batch : 1)]) {
    if (g.hop_side == 1) g.A = g.ident;
    else if (g.hop_side == 2) g.B = g.ident;
    // the identity is K x K: M = K when it is opA, N = K when it is opB
    return ((g.hop_side == 1 && g.M == g.K) || (g.hop_side == 2 && g.N == g.K)) && g.batch <= 1 && !g.amap &&
           !g.bmap && !g.kdev && !g.yhat;
  }
  return false;
}

// The tile [i0, i0 + tm) x [j0, j0 + tn) of C = alpha rs (I X) cs + beta C when the hopping
// operand is the identity: X = diag(bs) opB for hop_side 1, opA diag(bs) for hop_side 2, each element
// with the epilogue of the product.  The product's value without its MFMA work: in FP64 the
// MFMA product with an identity adds only zeros, and with FP32 operands (prec 1) the copy
// keeps the double values the product would have rounded.
__device__ void identity_tile(const GemmArgs& g, int i0, int j0, int tm, int tn, int tid, int nt) {
  cplx* C = (cplx*)g.C;
  for (int e = tid; e < tm * tn; e += nt) {
    const int i = i0 + e % tm, j = j0 + e / tm;
    if (i >= g.M || j >= g.N) continue;
    cplx v;
    if (g.hop_side == 1) {
      const long long off = (g.opB == OP_N) ? i + (long long)j * g.ldb : j + (long long)i * g.ldb;
      v = g.b_cplx ? ((const cplx*)g.B)[off] : make_double2(((const double*)g.B)[off], 0.0);
      if (g.opB == OP_C) v.y = -v.y;
      if (g.bs) v = cmul(g.bs[i], v);
    } else {
      const long long off = (g.opA == OP_N) ? i + (long long)j * g.lda : j + (long long)i * g.lda;
      v = g.a_cplx ? ((const cplx*)g.A)[off] : make_double2(((const double*)g.A)[off], 0.0);
      if (g.opA == OP_C) v.y = -v.y;
      if (g.bs) v = cmul(v, g.bs[j]);
    }
    v = cmul(g.alpha, v);
    if (g.rs) v = cmul(g.rs[i], v);
    if (g.rs_real) v = cscale(v, g.rs_real[i]);
    if (g.cs) v = cmul(v, g.cs[j]);
    const long long o = i + (long long)j * g.ldc;
    if (g.beta.x != 0.0 || g.beta.y != 0.0) v = cadd(v, cmul(g.beta, C[o]));
    C[o] = v;
  }
}

// One K stage of an LDS-tiled kernel's operands, held in registers from its global loads to
// its LDS stores.  Gather indices and row scales are read first and every data load is then
// issued back to back: no wait on one load sits between the issue of two others (a wait at
// the join of an optional map lookup would otherwise serialise the whole stage).
// Conjugation is applied when the stage is stored; the row scales of opB (one per k) go to
// an LDS vector of the stage and are applied to the B fragments at multiplication time.
template <int TM, int TN, int BK, bool A_CPLX, bool B_CPLX>
struct Stage {
  static constexpr int LA = TM * BK / NT, LB = BK * TN / NT;   // elements per thread
  static_assert(LA >= 1 && LB >= 1 && BK <= NT, "a stage covers the workgroup");
  double a_re[LA], a_im[LA], b_re[LB], b_im[LB];
  cplx scl;                 // bs(k0 + tid) for tid < BK

  __device__ __forceinline__ static void a_index(const GemmArgs& g, int tid, int p, int& i, int& kk) {
    const int idx = tid + NT * p;
    if (g.opA == OP_N) { i = idx % TM; kk = idx / TM; } else { kk = idx % BK; i = idx / BK; }
  }
  __device__ __forceinline__ static void b_index(const GemmArgs& g, int tid, int p, int& kk, int& j) {
    const int idx = tid + NT * p;
    if (g.opB == OP_N) { kk = idx % BK; j = idx / BK; } else { j = idx % TN; kk = idx / TN; }
  }

  __device__ __forceinline__ void load(const GemmArgs& g, const void* A, const void* B, int i0, int j0,
                                       int K, int k0, int tid) {
    int ka[LA], kb[LB];
#pragma unroll
    for (int p = 0; p < LA; ++p) {
      int i, kk;
      a_index(g, tid, p, i, kk);
      ka[p] = (i0 + i < g.M && k0 + kk < K) ? k0 + kk : -1;
    }
#pragma unroll
    for (int p = 0; p < LB; ++p) {
      int kk, j;
      b_index(g, tid, p, kk, j);
      kb[p] = (k0 + kk < K && j0 + j < g.N) ? k0 + kk : -1;
    }
    if (g.bs) scl = (tid < BK && k0 + tid < K) ? g.bs[k0 + tid] : make_double2(0.0, 0.0);
    if (g.amap) {
#pragma unroll
      for (int p = 0; p < LA; ++p) if (ka[p] >= 0) ka[p] = g.amap[ka[p]];
    }
    if (g.bmap) {
#pragma unroll
      for (int p = 0; p < LB; ++p) if (kb[p] >= 0) kb[p] = g.bmap[kb[p]];
    }
#pragma unroll
    for (int p = 0; p < LA; ++p) {
      int i, kk;
      a_index(g, tid, p, i, kk);
      const long long gi = i0 + i;
      double re = 0.0, im = 0.0;
      if (ka[p] >= 0) {
        const long long off = (g.opA == OP_N) ? gi + (long long)ka[p] * g.lda : ka[p] + gi * g.lda;
        if (A_CPLX) {
          const cplx v = ((const cplx*)A)[off];
          re = v.x;
          im = v.y;
        } else {
          re = ((const double*)A)[off];
        }
      }
      a_re[p] = re;
      a_im[p] = im;
    }
#pragma unroll
    for (int p = 0; p < LB; ++p) {
      int kk, j;
      b_index(g, tid, p, kk, j);
      const long long gj = j0 + j;
      double re = 0.0, im = 0.0;
      if (kb[p] >= 0) {
        const long long off = (g.opB == OP_N) ? kb[p] + gj * g.ldb : gj + (long long)kb[p] * g.ldb;
        if (g.b_cplx) {
          const cplx v = ((const cplx*)B)[off];
          re = v.x;
          im = v.y;
        } else {
          re = ((const double*)B)[off];
        }
      }
      b_re[p] = re;
      b_im[p] = im;
    }
  }

  template <typename E>
  __device__ __forceinline__ void store(const GemmArgs& g, int tid, E (*As_re)[TM + 1], E (*As_im)[TM + 1],
                                        E (*Bs_re)[TN + 1], E (*Bs_im)[TN + 1], cplx* bsl) const {
#pragma unroll
    for (int p = 0; p < LA; ++p) {
      int i, kk;
      a_index(g, tid, p, i, kk);
      As_re[kk][i] = (E)a_re[p];
      if (A_CPLX) As_im[kk][i] = (E)((g.opA == OP_C) ? -a_im[p] : a_im[p]);
    }
#pragma unroll
    for (int p = 0; p < LB; ++p) {
      int kk, j;
      b_index(g, tid, p, kk, j);
      Bs_re[kk][j] = (E)b_re[p];
      if (B_CPLX) Bs_im[kk][j] = (E)((g.opB == OP_C) ? -b_im[p] : b_im[p]);
    }
    if (g.bs && tid < BK) bsl[tid] = scl;
  }
};

// B fragment times the row scale of its k (the scales of a stage sit in LDS)
template <typename E>
__device__ __forceinline__ void scale_b(const cplx s, E& re, E& im) {
  const E r2 = re * (E)s.x - im * (E)s.y;
  im = re * (E)s.y + im * (E)s.x;
  re = r2;
}

template <bool A_CPLX, bool B_CPLX, int WM, int WN, bool F32>
__global__ __launch_bounds__(NT) void gemm_kernel(GemmArgs g) {
  typedef Mma<F32> M;
  typedef typename M::E E;
  constexpr int KS = 4 / (WM * WN);
  constexpr int TM = 16 * WM, TN = 16 * WN, BK = 16 * KS;
  constexpr int LA = TM * BK / NT, LB = BK * TN / NT;   // elements per thread per stage
  static_assert(WM * WN * KS == 4 && LA >= 1 && LB >= 1, "4 waves per tile");
  __shared__ E As_re[BK][TM + 1];
  __shared__ E As_im[A_CPLX ? BK : 1][TM + 1];
  __shared__ E Bs_re[BK][TN + 1];
  __shared__ E Bs_im[B_CPLX ? BK : 1][TN + 1];
  __shared__ cplx red[KS > 1 ? KS - 1 : 1][KS > 1 ? TM * TN : 1];
  __shared__ cplx bsl[BK];

  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * TM;
  const int j0 = blockIdx.y * TN;
  if (chain_operands(g, blockIdx.z)) { identity_tile(g, i0, j0, TM, TN, tid, blockDim.x); return; }
  const int z = blockIdx.z % (g.batch > 1 ? g.batch : 1);
  int K = g.K;
  if (g.kdev) {
    int kd = *g.kdev;
    K = kd < K ? kd : K;
  }
  if (K <= 0 && g.beta.x == 1.0 && g.beta.y == 0.0) return;  // nothing to add

  const long long za = g.batch > 1 ? z * g.strideA : 0, zb = g.batch > 1 ? z * g.strideB : 0;
  const void* A = A_CPLX ? (const void*)((const cplx*)g.A + za) : (const void*)((const double*)g.A + za);
  const void* B = g.b_cplx ? (const void*)((const cplx*)g.B + zb) : (const void*)((const double*)g.B + zb);
  cplx* C = (cplx*)g.C + (g.batch > 1 ? z * g.strideC : 0);

  const int wave = tid >> 6, lane = tid & 63;
  const int wm = wave % WM, wn = (wave / WM) % WN, ks = wave / (WM * WN);
  const int wr = 16 * wm, wc = 16 * wn;
  const int lr = lane & 15, lk = lane >> 4;

  typename M::Acc acc_re = {0, 0, 0, 0};
  typename M::Acc acc_im = {0, 0, 0, 0};
  typename M::Acc acc_t2 = {0, 0, 0, 0};   // complex x complex: Ai Bi (3M)

  Stage<TM, TN, BK, A_CPLX, B_CPLX> st;

  if (K > 0) st.load(g, A, B, i0, j0, K, 0, tid);
  for (int k0 = 0; k0 < K; k0 += BK) {
    st.store(g, tid, As_re, As_im, Bs_re, Bs_im, bsl);
    __syncthreads();
    if (k0 + BK < K) st.load(g, A, B, i0, j0, K, k0 + BK, tid);   // in flight during this stage's MFMAs
#pragma unroll
    for (int kq = 0; kq < 16; kq += 4) {
      const int kk = 16 * ks + kq;
      const E a_re = As_re[kk + lk][wr + lr];
      E b_re = Bs_re[kk + lk][wc + lr];
      E b_im = B_CPLX ? Bs_im[kk + lk][wc + lr] : (E)0;
      if (B_CPLX && g.bs) scale_b(bsl[kk + lk], b_re, b_im);
      acc_re = M::mac(a_re, b_re, acc_re);
      if (A_CPLX && B_CPLX) {
        // 3M: acc_t2 = Ai Bi, acc_im = (Ar + Ai)(Br + Bi), combined after the K loop
        const E a_im = As_im[kk + lk][wr + lr];
        acc_t2 = M::mac(a_im, b_im, acc_t2);
        acc_im = M::mac(a_re + a_im, b_re + b_im, acc_im);
      } else if (B_CPLX) {
        acc_im = M::mac(a_re, b_im, acc_im);
      } else if (A_CPLX) {
        const E a_im = As_im[kk + lk][wr + lr];
        acc_im = M::mac(a_im, b_re, acc_im);
      }
    }
    __syncthreads();
  }
  if (A_CPLX && B_CPLX) {
    acc_im = acc_im - acc_re - acc_t2;
    acc_re = acc_re - acc_t2;
  }

  // ---- epilogue: register r of a lane holds row wr + M::row(r, lane / 16) of column
  //      wc + lane % 16 ----
  if (KS > 1) {
    // waves ks >= 1 park their partial tiles; the ks = 0 waves add them in order ks = 1, 2, ...
    if (ks > 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        red[ks - 1][(wr + M::row(r, lk)) + TM * (wc + lr)] = make_double2(acc_re[r], acc_im[r]);
    }
    __syncthreads();
    if (ks > 0) return;
#pragma unroll
    for (int q = 0; q < KS - 1; ++q) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const cplx p = red[q][(wr + M::row(r, lk)) + TM * (wc + lr)];
        acc_re[r] += p.x;
        acc_im[r] += p.y;
      }
    }
  }
  const int gj = j0 + wc + lr;
  if (gj >= g.N) return;
  cplx csj = g.cs ? g.cs[gj] : make_double2(1.0, 0.0);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int gi = i0 + wr + M::row(r, lk);
    if (gi >= g.M) continue;
    cplx v = make_double2(acc_re[r], (A_CPLX || B_CPLX) ? acc_im[r] : 0.0);
    v = cmul(g.alpha, v);
    if (g.rs) v = cmul(g.rs[gi], v);
    if (g.rs_real) v = cscale(v, g.rs_real[gi]);
    if (g.cs) v = cmul(v, csj);
    const long long off = gi + (long long)gj * g.ldc;
    if (g.beta.x != 0.0 || g.beta.y != 0.0) v = cadd(v, cmul(g.beta, C[off]));
    C[off] = v;
  }
}

// Register-blocked 64 x 64 tile for grids with several tiles per CU (batched chains): 4
// waves in a 2 x 2 arrangement, each wave a 32 x 32 sub-tile held as 2 x 2 MFMA blocks, so
// every operand fragment read from LDS feeds two blocks and a K stage of 16 carries 4x the
// MFMA work of the 32 x 32 tile per barrier.  The next stage's global loads are in flight
// (in registers) while the current stage is multiplied out of LDS.
template <bool A_CPLX, bool B_CPLX, bool F32, int BK>
__global__ __launch_bounds__(NT, 2) void gemm_big_kernel(GemmArgs g) {
  typedef Mma<F32> M;
  typedef typename M::E E;
  constexpr int TM = 64, TN = 64;
  __shared__ E As_re[BK][TM + 1];
  __shared__ E As_im[A_CPLX ? BK : 1][TM + 1];
  __shared__ E Bs_re[BK][TN + 1];
  __shared__ E Bs_im[B_CPLX ? BK : 1][TN + 1];
  __shared__ cplx bsl[BK];

  const int tid = threadIdx.x;
  int bx, by, bz;
  xcd_order(bx, by, bz);
  const int i0 = bx * TM;
  const int j0 = by * TN;
  if (chain_operands(g, bz)) { identity_tile(g, i0, j0, TM, TN, tid, blockDim.x); return; }
  const int z = bz % (g.batch > 1 ? g.batch : 1);
  int K = g.K;
  if (g.kdev) {
    const int kd = *g.kdev;
    K = kd < K ? kd : K;
  }
  if (K <= 0 && g.beta.x == 1.0 && g.beta.y == 0.0) return;  // nothing to add

  const long long za = g.batch > 1 ? z * g.strideA : 0, zb = g.batch > 1 ? z * g.strideB : 0;
  const void* A = A_CPLX ? (const void*)((const cplx*)g.A + za) : (const void*)((const double*)g.A + za);
  const void* B = g.b_cplx ? (const void*)((const cplx*)g.B + zb) : (const void*)((const double*)g.B + zb);
  cplx* C = (cplx*)g.C + (g.batch > 1 ? z * g.strideC : 0);

  const int wave = tid >> 6, lane = tid & 63;
  const int wr = 32 * (wave & 1), wc = 32 * (wave >> 1);
  const int lr = lane & 15, lk = lane >> 4;

  // complex x complex in three real products (3M): acc_re = Ar Br, acc_t2 = Ai Bi,
  // acc_im = (Ar + Ai)(Br + Bi); re = acc_re - acc_t2, im = acc_im - acc_re - acc_t2
  constexpr bool M3 = A_CPLX && B_CPLX;
  typename M::Acc acc_re[2][2], acc_im[2][2], acc_t2[M3 ? 2 : 1][M3 ? 2 : 1];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      acc_re[a][b] = typename M::Acc{0, 0, 0, 0};
      acc_im[a][b] = typename M::Acc{0, 0, 0, 0};
      if (M3) acc_t2[M3 ? a : 0][M3 ? b : 0] = typename M::Acc{0, 0, 0, 0};
    }

  Stage<TM, TN, BK, A_CPLX, B_CPLX> st;

  if (K > 0) st.load(g, A, B, i0, j0, K, 0, tid);
  for (int k0 = 0; k0 < K; k0 += BK) {
    st.store(g, tid, As_re, As_im, Bs_re, Bs_im, bsl);
    __syncthreads();
    if (k0 + BK < K) st.load(g, A, B, i0, j0, K, k0 + BK, tid);
#pragma unroll
    for (int kq = 0; kq < BK; kq += 4) {
      const int kk = kq + lk;
      E a_re[2], a_im[2], b_re[2], b_im[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        a_re[a] = As_re[kk][wr + 16 * a + lr];
        if (A_CPLX) a_im[a] = As_im[kk][wr + 16 * a + lr];
      }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        b_re[b] = Bs_re[kk][wc + 16 * b + lr];
        if (B_CPLX) b_im[b] = Bs_im[kk][wc + 16 * b + lr];
      }
      if (B_CPLX && g.bs) {
        const cplx sk = bsl[kk];
#pragma unroll
        for (int b = 0; b < 2; ++b) scale_b(sk, b_re[b], b_im[b]);
      }
      if (M3) {
        E as[2], bsum[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) as[a] = a_re[a] + a_im[a];
#pragma unroll
        for (int b = 0; b < 2; ++b) bsum[b] = b_re[b] + b_im[b];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            acc_re[a][b] = M::mac(a_re[a], b_re[b], acc_re[a][b]);
            acc_t2[M3 ? a : 0][M3 ? b : 0] = M::mac(a_im[a], b_im[b], acc_t2[M3 ? a : 0][M3 ? b : 0]);
            acc_im[a][b] = M::mac(as[a], bsum[b], acc_im[a][b]);
          }
      } else {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            acc_re[a][b] = M::mac(a_re[a], b_re[b], acc_re[a][b]);
            if (B_CPLX) acc_im[a][b] = M::mac(a_re[a], b_im[b], acc_im[a][b]);
          }
        if (A_CPLX) {
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc_im[a][b] = M::mac(a_im[a], b_re[b], acc_im[a][b]);
        }
      }
    }
    __syncthreads();
  }
  if (M3) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const typename M::Acc t2 = acc_t2[M3 ? a : 0][M3 ? b : 0];
        acc_im[a][b] = acc_im[a][b] - acc_re[a][b] - t2;
        acc_re[a][b] = acc_re[a][b] - t2;
      }
  }

  // register r of a lane holds row M::row(r, lane / 16) of its 16 x 16 block
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int gj = j0 + wc + 16 * b + lr;
    if (gj >= g.N) continue;
    const cplx csj = g.cs ? g.cs[gj] : make_double2(1.0, 0.0);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = i0 + wr + 16 * a + M::row(r, lk);
        if (gi >= g.M) continue;
        cplx v = make_double2(acc_re[a][b][r], (A_CPLX || B_CPLX) ? acc_im[a][b][r] : 0.0);
        v = cmul(g.alpha, v);
        if (g.rs) v = cmul(g.rs[gi], v);
        if (g.rs_real) v = cscale(v, g.rs_real[gi]);
        if (g.cs) v = cmul(v, csj);
        const long long off = gi + (long long)gj * g.ldc;
        if (g.beta.x != 0.0 || g.beta.y != 0.0) v = cadd(v, cmul(g.beta, C[off]));
        C[off] = v;
      }
    }
  }
}

// Delayed-update flush G += X Y^T of rank m <= 32 (the block kernel's round).  A workgroup
// owns a 64 x 64 tile of one chain's G; its operands (the m columns of X for the tile's rows
// and of Y for its columns) and the tile of G itself, read directly in the MFMA result
// layout, are all requested at the start, so a tile costs one memory round trip; the
// product takes ceil(m / 4) k-steps of 2 x 2 MFMA blocks per wave and G is written once.
constexpr int FL_T = 64, FL_K = 32;
__global__ __launch_bounds__(NT, 2) void flush_kernel(cplx* G, const cplx* X, const cplx* Y, const int* mdev,
                                                      int n, int kmax, long long cs) {
  typedef Mma<false> M;
  __shared__ double Xr[FL_K][FL_T + 1], Xi[FL_K][FL_T + 1], Yr[FL_K][FL_T + 1], Yi[FL_K][FL_T + 1];
  int bx, by, bz;
  xcd_order(bx, by, bz);
  G = chain_ptr_at(G, cs, bz);
  X = chain_ptr_at(X, cs, bz);
  Y = chain_ptr_at(Y, cs, bz);
  mdev = chain_ptr_at(mdev, cs, bz);
  int m = *mdev;
  m = m < kmax ? m : kmax;
  if (m <= 0) return;
  const int tid = threadIdx.x;
  const int i0 = bx * FL_T, j0 = by * FL_T;
  const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  const int wr = 32 * (wave & 1), wc = 32 * (wave >> 1);
  const int li = tid & 63, lkb = tid >> 6;   // operand rows: index li, columns k = lkb + 4 p

  cplx xv[FL_K / 4], yv[FL_K / 4];
#pragma unroll
  for (int p = 0; p < FL_K / 4; ++p) {
    const int k = lkb + 4 * p;
    xv[p] = make_double2(0.0, 0.0);
    yv[p] = make_double2(0.0, 0.0);
    if (k < m && i0 + li < n) xv[p] = X[i0 + li + (long long)k * n];
    if (k < m && j0 + li < n) yv[p] = Y[j0 + li + (long long)k * n];
  }
  cplx gv[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = i0 + wr + 16 * a + M::row(r, lk), gj = j0 + wc + 16 * b + lr;
        gv[a][b][r] = make_double2(0.0, 0.0);
        if (gi < n && gj < n) gv[a][b][r] = G[gi + (long long)gj * n];
      }
#pragma unroll
  for (int p = 0; p < FL_K / 4; ++p) {
    const int k = lkb + 4 * p;
    Xr[k][li] = xv[p].x;
    Xi[k][li] = xv[p].y;
    Yr[k][li] = yv[p].x;
    Yi[k][li] = yv[p].y;
  }
  __syncthreads();

  v4d acc_re[2][2], acc_im[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      acc_re[a][b] = v4d{0, 0, 0, 0};
      acc_im[a][b] = v4d{0, 0, 0, 0};
    }
  for (int k0 = 0; k0 < m; k0 += 4) {
    const int kk = k0 + lk;
    double a_re[2], a_im[2], b_re[2], b_im[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      a_re[a] = Xr[kk][wr + 16 * a + lr];
      a_im[a] = Xi[kk][wr + 16 * a + lr];
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      b_re[b] = Yr[kk][wc + 16 * b + lr];
      b_im[b] = Yi[kk][wc + 16 * b + lr];
    }
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const double na = -a_im[a];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc_re[a][b] = M::mac(a_re[a], b_re[b], acc_re[a][b]);
        acc_im[a][b] = M::mac(a_re[a], b_im[b], acc_im[a][b]);
        acc_re[a][b] = M::mac(na, b_im[b], acc_re[a][b]);
        acc_im[a][b] = M::mac(a_im[a], b_re[b], acc_im[a][b]);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = i0 + wr + 16 * a + M::row(r, lk), gj = j0 + wc + 16 * b + lr;
        if (gi < n && gj < n)
          G[gi + (long long)gj * n] = make_double2(gv[a][b][r].x + acc_re[a][b][r], gv[a][b][r].y + acc_im[a][b][r]);
      }
}

inline long long cdivl(long long a, long long b) { return (a + b - 1) / b; }
inline unsigned gemm_zdim(const GemmArgs& g) {
  return (unsigned)((g.batch > 1 ? g.batch : 1) * (g.nchain > 1 ? g.nchain : 1));
}

template <int WM, int WN, bool F32>
void launch_shape_p(const GemmArgs& g, hipStream_t s) {
  dim3 grid((unsigned)cdivl(g.M, 16 * WM), (unsigned)cdivl(g.N, 16 * WN), gemm_zdim(g));
  const bool bc = g.b_cplx || g.bs != nullptr;
  if (g.a_cplx && bc) hipLaunchKernelGGL((gemm_kernel<true, true, WM, WN, F32>), grid, dim3(NT), 0, s, g);
  else if (g.a_cplx) hipLaunchKernelGGL((gemm_kernel<true, false, WM, WN, F32>), grid, dim3(NT), 0, s, g);
  else if (bc) hipLaunchKernelGGL((gemm_kernel<false, true, WM, WN, F32>), grid, dim3(NT), 0, s, g);
  else hipLaunchKernelGGL((gemm_kernel<false, false, WM, WN, F32>), grid, dim3(NT), 0, s, g);
}
template <int WM, int WN>
void launch_shape(const GemmArgs& g, hipStream_t s) {
  if (g.prec) launch_shape_p<WM, WN, true>(g, s);
  else launch_shape_p<WM, WN, false>(g, s);
}

template <bool F32, int BK>
void launch_big_p(const GemmArgs& g, hipStream_t s) {
  dim3 grid((unsigned)cdivl(g.M, 64), (unsigned)cdivl(g.N, 64), gemm_zdim(g));
  const bool bc = g.b_cplx || g.bs != nullptr;
  // complex x complex always in K stages of 16 (32 would not fit two waves per SIMD)
  if (g.a_cplx && bc) hipLaunchKernelGGL((gemm_big_kernel<true, true, F32, 16>), grid, dim3(NT), 0, s, g);
  else if (g.a_cplx) hipLaunchKernelGGL((gemm_big_kernel<true, false, F32, BK>), grid, dim3(NT), 0, s, g);
  else if (bc) hipLaunchKernelGGL((gemm_big_kernel<false, true, F32, BK>), grid, dim3(NT), 0, s, g);
  else hipLaunchKernelGGL((gemm_big_kernel<false, false, F32, BK>), grid, dim3(NT), 0, s, g);
}
// K stages of 32 for long products with a real operand (the hopping and UDV wraps): twice
// the MFMA work per barrier and per load latency; complex x complex stays at 16, where 32
// would not fit two waves per SIMD without spilling
void launch_big(const GemmArgs& g, hipStream_t s) {
  const bool deep = g.K >= 64 && !(g.a_cplx && (g.b_cplx || g.bs != nullptr));
  if (g.prec) deep ? launch_big_p<true, 32>(g, s) : launch_big_p<true, 16>(g, s);
  else deep ? launch_big_p<false, 32>(g, s) : launch_big_p<false, 16>(g, s);
}


// Direct-load split-K kernel for grids that would not fill the device: a workgroup of
// NW waves owns one 16 x 16 output tile; wave w takes the w-th contiguous quarter
// (eighth, ...) of K and feeds its MFMAs straight from global memory in the
// v_mfma_f64_16x16x4f64 operand layout (lane = row + 16 k for A, column + 16 k for B),
// four or eight k-steps per batch with the next batch in flight.  The NW partial tiles are
// summed through LDS in wave order.
template <bool A_CPLX, bool B_CPLX, int NW, bool F32>
__global__ __launch_bounds__(64 * NW) void gemm_direct_kernel(GemmArgs g) {
  typedef Mma<F32> M;
  typedef typename M::E E;
  __shared__ cplx red[NW > 1 ? NW - 1 : 1][256];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
  if (chain_operands(g, blockIdx.z)) { identity_tile(g, i0, j0, 16, 16, tid, blockDim.x); return; }
  const int z = blockIdx.z % (g.batch > 1 ? g.batch : 1);
  int K = g.K;
  if (g.kdev) {
    const int kd = *g.kdev;
    K = kd < K ? kd : K;
  }
  if (A_CPLX && g.yhat && g.amap) {
    // the side copy of rows amap of A, spread over every thread of the grid
    const int nwg = gridDim.x * gridDim.y;
    const long long total = (long long)g.M * (K > 0 ? K : 0);
    const cplx* Ag = (const cplx*)g.A;
    for (long long e = (long long)(blockIdx.x + gridDim.x * blockIdx.y) * blockDim.x + tid; e < total;
         e += (long long)nwg * blockDim.x) {
      const int r = (int)(e % g.M), j = (int)(e / g.M);
      const int P = g.amap[j];
      const cplx v = Ag[P + (long long)r * g.lda];
      g.yhat[e] = make_double2((r == P ? 1.0 : 0.0) - v.x, -v.y);
    }
  }
  if (K <= 0 && g.beta.x == 1.0 && g.beta.y == 0.0) return;  // nothing to add
  const double* Ar = (const double*)g.A;
  const cplx* Ac = (const cplx*)g.A;
  const double* Br = (const double*)g.B;
  const cplx* Bc = (const cplx*)g.B;
  if (g.batch > 1) {
    Ar += z * g.strideA; Ac += z * g.strideA;
    Br += z * g.strideB; Bc += z * g.strideB;
  }
  cplx* C = (cplx*)g.C + (g.batch > 1 ? z * g.strideC : 0);
  // K range of this wave, in whole k-steps of 4
  const int ksteps = (K + 3) >> 2;
  const int per = (ksteps + NW - 1) / NW;
  const int kb = wave * per * 4, ke0 = (wave + 1) * per * 4;
  const int ke = ke0 < K ? ke0 : K;
  const int gi = i0 + lr, gj = j0 + lr;
  const bool row_ok = gi < g.M, col_ok = gj < g.N;

  // One batch of U k-steps in the MFMA operand layout.  Gather indices and row scales are
  // read first, then every data load is issued back to back; conjugation and scaling are
  // applied when the batch moves into the MFMA operands.
  constexpr int U = NW >= 4 ? 8 : 4;   // k-steps per batch: one batch covers a wave's K at Ndim = 256
  auto load_batch = [&](int kbase, cplx* a, cplx* b, cplx* sc) {
    int ia[U], ib[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = kbase + 4 * u + lk;
      ia[u] = (row_ok && k < ke) ? k : -1;
      ib[u] = (col_ok && k < ke) ? k : -1;
    }
    if (g.bs) {
#pragma unroll
      for (int u = 0; u < U; ++u) sc[u] = ib[u] >= 0 ? g.bs[ib[u]] : make_double2(0.0, 0.0);
    }
    if (g.amap) {
#pragma unroll
      for (int u = 0; u < U; ++u) if (ia[u] >= 0) ia[u] = g.amap[ia[u]];
    }
    if (g.bmap) {
#pragma unroll
      for (int u = 0; u < U; ++u) if (ib[u] >= 0) ib[u] = g.bmap[ib[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a[u] = make_double2(0.0, 0.0);
      if (ia[u] >= 0) {
        const long long off = (g.opA == OP_N) ? gi + (long long)ia[u] * g.lda : ia[u] + (long long)gi * g.lda;
        if (A_CPLX) a[u] = Ac[off];
        else a[u].x = Ar[off];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      b[u] = make_double2(0.0, 0.0);
      if (ib[u] >= 0) {
        const long long off = (g.opB == OP_N) ? ib[u] + (long long)gj * g.ldb : gj + (long long)ib[u] * g.ldb;
        if (g.b_cplx) b[u] = Bc[off];
        else b[u].x = Br[off];
      }
    }
  };
  auto finish = [&](cplx* a, cplx* b, const cplx* sc) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (A_CPLX && g.opA == OP_C) a[u].y = -a[u].y;
      if (g.opB == OP_C) b[u].y = -b[u].y;
      if (g.bs) b[u] = cmul(b[u], sc[u]);
    }
  };

  // complex x complex in three real products (3M): acc_re = Ar Br, acc_t2 = Ai Bi,
  // acc_im = (Ar + Ai)(Br + Bi); re = acc_re - acc_t2, im = acc_im - acc_re - acc_t2
  constexpr bool M3 = A_CPLX && B_CPLX;
  typename M::Acc acc_re = {0, 0, 0, 0}, acc_im = {0, 0, 0, 0}, acc_t2 = {0, 0, 0, 0};
  cplx av[U], bv[U], sv[U];
  load_batch(kb, av, bv, sv);
  finish(av, bv, sv);
  for (int k0 = kb; k0 < ke; k0 += 4 * U) {
    cplx an[U], bn[U], sn[U];
    load_batch(k0 + 4 * U, an, bn, sn);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (M3) {
        acc_re = M::mac((E)av[u].x, (E)bv[u].x, acc_re);
        acc_t2 = M::mac((E)av[u].y, (E)bv[u].y, acc_t2);
        acc_im = M::mac((E)(av[u].x + av[u].y), (E)(bv[u].x + bv[u].y), acc_im);
      } else {
        acc_re = M::mac((E)av[u].x, (E)bv[u].x, acc_re);
        if (B_CPLX) acc_im = M::mac((E)av[u].x, (E)bv[u].y, acc_im);
        if (A_CPLX) acc_im = M::mac((E)av[u].y, (E)bv[u].x, acc_im);
      }
    }
    finish(an, bn, sn);
#pragma unroll
    for (int u = 0; u < U; ++u) { av[u] = an[u]; bv[u] = bn[u]; }
  }
  if (M3) {
    acc_im = acc_im - acc_re - acc_t2;
    acc_re = acc_re - acc_t2;
  }
  // register r of a lane holds row M::row(r, lane / 16) of the tile
  if (NW > 1) {
    if (wave > 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) red[wave - 1][M::row(r, lk) + 16 * lr] = make_double2(acc_re[r], acc_im[r]);
    }
    __syncthreads();
    if (wave > 0) return;
#pragma unroll
    for (int w = 0; w < NW - 1; ++w) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const cplx p = red[w][M::row(r, lk) + 16 * lr];
        acc_re[r] += p.x;
        acc_im[r] += p.y;
      }
    }
  }
  if (!col_ok) return;
  const cplx csj = g.cs ? g.cs[gj] : make_double2(1.0, 0.0);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int ri = i0 + M::row(r, lk);
    if (ri >= g.M) continue;
    cplx v = make_double2(acc_re[r], (A_CPLX || B_CPLX) ? acc_im[r] : 0.0);
    v = cmul(g.alpha, v);
    if (g.rs) v = cmul(g.rs[ri], v);
    if (g.rs_real) v = cscale(v, g.rs_real[ri]);
    if (g.cs) v = cmul(v, csj);
    const long long off = ri + (long long)gj * g.ldc;
    if (g.beta.x != 0.0 || g.beta.y != 0.0) v = cadd(v, cmul(g.beta, C[off]));
    C[off] = v;
  }
}

template <int NW, bool F32>
void launch_direct_p(const GemmArgs& g, hipStream_t s) {
  dim3 grid((unsigned)cdivl(g.M, 16), (unsigned)cdivl(g.N, 16), gemm_zdim(g));
  const bool bc = g.b_cplx || g.bs != nullptr;
  if (g.a_cplx && bc) hipLaunchKernelGGL((gemm_direct_kernel<true, true, NW, F32>), grid, dim3(64 * NW), 0, s, g);
  else if (g.a_cplx) hipLaunchKernelGGL((gemm_direct_kernel<true, false, NW, F32>), grid, dim3(64 * NW), 0, s, g);
  else if (bc) hipLaunchKernelGGL((gemm_direct_kernel<false, true, NW, F32>), grid, dim3(64 * NW), 0, s, g);
  else hipLaunchKernelGGL((gemm_direct_kernel<false, false, NW, F32>), grid, dim3(64 * NW), 0, s, g);
}
template <int NW>
void launch_direct(const GemmArgs& g, hipStream_t s) {
  if (g.prec) launch_direct_p<NW, true>(g, s);
  else launch_direct_p<NW, false>(g, s);
}
}  // namespace

// The end of a delayed-update round for a few chains, in one launch:
//   Go = Gi + (Gi(:, A) W) Yhat^T,   Yhat(c, j) = delta(c, P_j) - Gi(P_j, c),
// A = (P_0 .. P_{m-1}) the round's accepted sites, W = C D^T (m x m, ld FL_K), m = *mdev.
// Gi is only read and Go only written (two buffers), so a workgroup can gather the columns A
// and rows A of Gi that it needs while others write their tiles.  A workgroup of two waves
// owns a 16 x 16 tile: wave w forms columns 16w .. 16w + 15 of its rows of X = Gi(:, A) W,
// then the waves split the rank-m product over k and add their halves through LDS.  Every
// global load of the tile is issued before the first MFMA.
constexpr int RF_T = 16;
__global__ __launch_bounds__(128) void round_flush_kernel(const cplx* __restrict__ Gi, cplx* __restrict__ Go,
                                                          const cplx* __restrict__ W, const int* __restrict__ amap,
                                                          const int* mdev, int n, int kmax, long long cs) {
  typedef Mma<false> M;
  __shared__ cplx Xs[RF_T][FL_K + 1];
  __shared__ cplx red[RF_T * RF_T];
  const int z = blockIdx.z;
  Gi = chain_ptr_at(Gi, cs, z);
  Go = chain_ptr_at(Go, cs, z);
  W = chain_ptr_at(W, cs, z);
  amap = chain_ptr_at(amap, cs, z);
  mdev = chain_ptr_at(mdev, cs, z);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int i0 = blockIdx.x * RF_T, j0 = blockIdx.y * RF_T;
  const int gi = i0 + lr, gj = j0 + lr;
  const bool row_ok = gi < n, col_ok = gj < n;
  // the rank m and the sites are loaded together (the sites' buffer holds at least kmax
  // entries), so the gathers below wait for one load instead of two
  int pa[FL_K / 4], py[4];
#pragma unroll
  for (int u = 0; u < FL_K / 4; ++u) pa[u] = 4 * u + lk < kmax ? amap[4 * u + lk] : 0;
#pragma unroll
  for (int u = 0; u < 4; ++u) py[u] = 16 * wave + 4 * u + lk < kmax ? amap[16 * wave + 4 * u + lk] : 0;
  int m = *mdev;
  m = m < kmax ? m : kmax;
  m = m > 0 ? m : 0;

  // X(gi, 16 wave + lr') = sum_k Gi(gi, P_k) W(k, 16 wave + lr'): A operand lane = row + 16 k,
  // B operand lane = column + 16 k, 8 k-steps of 4
  cplx xa[FL_K / 4], xb[FL_K / 4];
#pragma unroll
  for (int u = 0; u < FL_K / 4; ++u) {
    const int k = 4 * u + lk;
    const int P = pa[u];
    xa[u] = (k < m && row_ok) ? Gi[gi + (long long)P * n] : make_double2(0.0, 0.0);
    xb[u] = k < m ? W[k + (long long)(16 * wave + lr) * FL_K] : make_double2(0.0, 0.0);
  }
  // this wave's half of the rank-m product: Yhat(gj, k) for k in [16 wave, 16 wave + 16)
  cplx yb[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int k = 16 * wave + 4 * u + lk;
    yb[u] = make_double2(0.0, 0.0);
    if (k < m && col_ok) {
      const int P = py[u];
      const cplx v = Gi[P + (long long)gj * n];
      yb[u] = make_double2((gj == P ? 1.0 : 0.0) - v.x, -v.y);
    }
  }
  // the tile of Gi in the MFMA result layout (wave 0 adds it)
  cplx gt[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int ri = i0 + M::row(r, lk);
    gt[r] = (wave == 0 && ri < n && col_ok) ? Gi[ri + (long long)gj * n] : make_double2(0.0, 0.0);
  }

  v4d xr = {0, 0, 0, 0}, xi = {0, 0, 0, 0};
#pragma unroll
  for (int u = 0; u < FL_K / 4; ++u) {
    xr = M::mac(xa[u].x, xb[u].x, xr);
    xi = M::mac(xa[u].x, xb[u].y, xi);
    xr = M::mac(-xa[u].y, xb[u].y, xr);
    xi = M::mac(xa[u].y, xb[u].x, xi);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) Xs[M::row(r, lk)][16 * wave + lr] = make_double2(xr[r], xi[r]);
  __syncthreads();

  v4d gr = {0, 0, 0, 0}, gim = {0, 0, 0, 0};
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const cplx a = Xs[lr][16 * wave + 4 * u + lk];
    gr = M::mac(a.x, yb[u].x, gr);
    gim = M::mac(a.x, yb[u].y, gim);
    gr = M::mac(-a.y, yb[u].y, gr);
    gim = M::mac(a.y, yb[u].x, gim);
  }
  if (wave == 1) {
#pragma unroll
    for (int r = 0; r < 4; ++r) red[M::row(r, lk) + RF_T * lr] = make_double2(gr[r], gim[r]);
  }
  __syncthreads();
  if (wave == 1 || !col_ok) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int ri = i0 + M::row(r, lk);
    if (ri >= n) continue;
    const cplx h = red[M::row(r, lk) + RF_T * lr];
    Go[ri + (long long)gj * n] = make_double2(gt[r].x + (gr[r] + h.x), gt[r].y + (gim[r] + h.y));
  }
}

int round_flush_launch(const cplx* Gi, cplx* Go, const cplx* W, const int* amap, const int* mdev, int n, int kmax,
                       int nchain, long long cs, hipStream_t s) {
  if (kmax < 1 || kmax > FL_K) { alfq_set_error("round flush: rank outside [1, 32]"); return ST_GENERIC; }
  if (nchain < 1 || nchain > 65535 || n < 1) { alfq_set_error("round flush: bad batch"); return ST_GENERIC; }
  const unsigned t = (unsigned)cdivl(n, RF_T);
  hipLaunchKernelGGL(round_flush_kernel, dim3(t, t, (unsigned)nchain), dim3(128), 0, s, Gi, Go, W, amap, mdev, n,
                     kmax, cs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

int flush_launch(cplx* G, const cplx* X, const cplx* Y, const int* mdev, int n, int kmax, int nchain,
                 long long cs, hipStream_t s) {
  if (kmax < 1 || kmax > FL_K) { alfq_set_error("flush: rank outside [1, 32]"); return ST_GENERIC; }
  if (nchain < 1 || nchain > 65535) { alfq_set_error("flush: bad batch"); return ST_GENERIC; }
  const unsigned t = (unsigned)cdivl(n, FL_T);
  hipLaunchKernelGGL(flush_kernel, dim3(t, t, (unsigned)nchain), dim3(NT), 0, s, G, X, Y, mdev, n, kmax, cs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

GemmArgs gemm_args(int M, int N, int K, const void* A, int lda, int opA, bool a_cplx,
                   const void* B, int ldb, int opB, bool b_cplx, void* C, int ldc,
                   cplx alpha, cplx beta) {
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K; g.kdev = nullptr;
  g.A = A; g.lda = lda; g.opA = opA; g.a_cplx = a_cplx;
  g.B = B; g.ldb = ldb; g.opB = opB; g.b_cplx = b_cplx;
  g.C = C; g.ldc = ldc;
  g.alpha = alpha; g.beta = beta;
  g.batch = 1;
  g.nchain = 1;
  g.cstride = 0;
  g.prec = 0;
  g.policy = 0;
  g.lz = nullptr;
  g.slice = 0;
  g.hop_side = 0;
  g.ident = nullptr;
  return g;
}

// Tile shape with the least estimated time: every SIMD of the 256 CUs holds one wave of a
// full round, and a wave's time is proportional to the K it covers (K / KS).
static int gemm_tile_shape(int M, int N, int K, int batch) {
  const long long b = batch > 1 ? batch : 1;
  const long long slots = 256LL * 4;
  struct Opt { int code, ks; long long tm, tn; };
  const Opt opts[3] = {{0, 1, 32, 32}, {1, 2, 16, 32}, {2, 4, 16, 16}};
  int best = 0;
  double best_cost = 0.0;
  for (int o = 0; o < 3; ++o) {
    const long long waves = 4 * b * cdivl(M, opts[o].tm) * cdivl(N, opts[o].tn);
    const double cost = (double)cdivl(waves, slots) * (double)cdivl(K > 0 ? K : 1, 16 * opts[o].ks);
    if (o == 0 || cost < best_cost) { best = o; best_cost = cost; }
  }
  return opts[best].code;
}

// grid sizes of the tiling policies: direct-load kernel below direct_below tiles of 32 x 32,
// 64 x 64 kernel from big_from tiles of 64 x 64 (with 3/4 of their area inside the product)
struct TilePolicy { long long direct_below, big_from; };
constexpr TilePolicy kTilePolicy[GEMM_POLICIES] = {{256, 512}, {256, 256}, {256, 1LL << 62}, {512, 512}};

static const TilePolicy& tile_policy(const GemmArgs& g) {
  return kTilePolicy[(g.policy >= 0 && g.policy < GEMM_POLICIES) ? g.policy : 0];
}

bool gemm_direct_path(const GemmArgs& g) {
  return g.M > 0 && g.N > 0 && cdivl(g.M, 32) * cdivl(g.N, 32) * (long long)gemm_zdim(g) < tile_policy(g).direct_below;
}

int gemm_launch(const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return ST_NONE;
  if (gemm_zdim(g) > 65535) { alfq_set_error("gemm_launch: batch too large"); return ST_GENERIC; }
  if (g.yhat && !(gemm_direct_path(g) && g.amap && g.a_cplx && g.opA == OP_N && g.batch <= 1)) {
    alfq_set_error("gemm_launch: the row copy yhat needs the direct kernel, a complex opA = A and amap");
    return ST_GENERIC;
  }
  const long long b = gemm_zdim(g);
  if (gemm_direct_path(g)) {
    // too few 32 x 32 tiles to fill the device: 16 x 16 tiles, K split over up to 8 waves
    const int ksteps = g.K > 0 ? (g.K + 3) / 4 : 1;
    if (ksteps >= 32) launch_direct<8>(g, s);
    else if (ksteps >= 16) launch_direct<4>(g, s);
    else if (ksteps >= 8) launch_direct<2>(g, s);
    else launch_direct<1>(g, s);
    HIP_OK(hipGetLastError());
    return ST_NONE;
  }
  {
    // two or more 64 x 64 tiles per CU, at least 3/4 of each tile's area inside the product
    const long long tm = cdivl(g.M, 64), tn = cdivl(g.N, 64);
    if (tm * tn * b >= tile_policy(g).big_from && 4LL * g.M * g.N >= 3LL * (64 * tm) * (64 * tn)) {
      launch_big(g, s);
      HIP_OK(hipGetLastError());
      return ST_NONE;
    }
  }
  switch (gemm_tile_shape(g.M, g.N, g.K, (int)gemm_zdim(g))) {
    case 0: launch_shape<2, 2>(g, s); break;
    case 1: launch_shape<1, 2>(g, s); break;
    default: launch_shape<1, 1>(g, s); break;
  }
  HIP_OK(hipGetLastError());
  return ST_NONE;
}
