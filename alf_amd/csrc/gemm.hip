// FP64 MFMA GEMM for gfx950: real x complex and complex x complex products with
// fused diagonal scalings, gathers and conjugation.
//
// Tiling: one workgroup (4 waves) computes a 32 x 32 complex output tile; each
// wave owns a 16 x 16 sub-tile held as two (re, im) v_mfma_f64_16x16x4f64
// accumulators.  K advances in steps of 16 through LDS, where the operands are
// split into real and imaginary planes so each MFMA consumes one plane.
// A complex product costs 4 MFMAs per k-step of 4, a real x complex one 2.
//
// This one kernel carries every dense product of the sweep:
//   * hopping wraps  G <- D exp(-dtau T) G exp(dtau T) D^-1       (K1, real T: 2 MFMA/step)
//   * UDV chain      U <- D exp(-dtau T) U,  U <- exp(-dtau T) conj(D) U
//   * delayed-update flush  G += X Y^T                              (K2b)
//   * QR trailing updates / Q formation / TRMM / CGR products       (K3, K4)
#include "alfq_common.h"

namespace {

constexpr int BM = 32, BN = 32, BK = 16, NT = 256;

typedef double v4d __attribute__((ext_vector_type(4)));

template <bool A_CPLX, bool B_CPLX>
__global__ __launch_bounds__(NT) void gemm_kernel(GemmArgs g) {
  __shared__ double As_re[BK][BM + 1];
  __shared__ double As_im[A_CPLX ? BK : 1][BM + 1];
  __shared__ double Bs_re[BK][BN + 1];
  __shared__ double Bs_im[B_CPLX ? BK : 1][BN + 1];

  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * BM;
  const int j0 = blockIdx.y * BN;
  const int z = blockIdx.z;
  int K = g.K;
  if (g.kdev) {
    int kd = *g.kdev;
    K = kd < K ? kd : K;
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

  const int wave = tid >> 6, lane = tid & 63;
  const int wr = (wave & 1) * 16, wc = (wave >> 1) * 16;
  const int lr = lane & 15, lk = lane >> 4;

  v4d acc_re = {0, 0, 0, 0};
  v4d acc_im = {0, 0, 0, 0};

  // global -> register loads of one K-tile (2 opA and 2 opB elements per thread)
  double ra_re[2], ra_im[2], rb_re[2], rb_im[2];
  auto load_tile = [&](int k0) {
    if (g.opA == OP_N) {
      const int i = tid & 31;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int kk = (tid >> 5) + 8 * p;
        const int gi = i0 + i, gk = k0 + kk;
        double re = 0.0, im = 0.0;
        if (gi < g.M && gk < K) {
          const int col = g.amap ? g.amap[gk] : gk;
          const long long off = gi + (long long)col * g.lda;
          if (A_CPLX) { cplx v = Ac[off]; re = v.x; im = v.y; } else { re = Ar[off]; }
        }
        ra_re[p] = re;
        ra_im[p] = im;
      }
    } else {
      const int kk = tid & 15;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int i = (tid >> 4) + 16 * p;
        const int gi = i0 + i, gk = k0 + kk;
        double re = 0.0, im = 0.0;
        if (gi < g.M && gk < K) {
          const int row = g.amap ? g.amap[gk] : gk;
          const long long off = row + (long long)gi * g.lda;
          if (A_CPLX) { cplx v = Ac[off]; re = v.x; im = (g.opA == OP_C) ? -v.y : v.y; } else { re = Ar[off]; }
        }
        ra_re[p] = re;
        ra_im[p] = im;
      }
    }
    if (g.opB == OP_N) {
      const int kk = tid & 15;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int j = (tid >> 4) + 16 * p;
        const int gk = k0 + kk, gj = j0 + j;
        double re = 0.0, im = 0.0;
        if (gk < K && gj < g.N) {
          const int row = g.bmap ? g.bmap[gk] : gk;
          const long long off = row + (long long)gj * g.ldb;
          if (g.b_cplx) { cplx v = Bc[off]; re = v.x; im = v.y; } else { re = Br[off]; }
          if (g.bs) { cplx sc = g.bs[gk]; double r2 = re * sc.x - im * sc.y; im = re * sc.y + im * sc.x; re = r2; }
        }
        rb_re[p] = re;
        rb_im[p] = im;
      }
    } else {
      const int j = tid & 31;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int kk = (tid >> 5) + 8 * p;
        const int gk = k0 + kk, gj = j0 + j;
        double re = 0.0, im = 0.0;
        if (gk < K && gj < g.N) {
          const int col = g.bmap ? g.bmap[gk] : gk;
          const long long off = gj + (long long)col * g.ldb;
          if (g.b_cplx) { cplx v = Bc[off]; re = v.x; im = (g.opB == OP_C) ? -v.y : v.y; } else { re = Br[off]; }
          if (g.bs) { cplx sc = g.bs[gk]; double r2 = re * sc.x - im * sc.y; im = re * sc.y + im * sc.x; re = r2; }
        }
        rb_re[p] = re;
        rb_im[p] = im;
      }
    }
  };
  auto store_tile = [&]() {
    if (g.opA == OP_N) {
      const int i = tid & 31;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int kk = (tid >> 5) + 8 * p;
        As_re[kk][i] = ra_re[p];
        if (A_CPLX) As_im[kk][i] = ra_im[p];
      }
    } else {
      const int kk = tid & 15;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int i = (tid >> 4) + 16 * p;
        As_re[kk][i] = ra_re[p];
        if (A_CPLX) As_im[kk][i] = ra_im[p];
      }
    }
    if (g.opB == OP_N) {
      const int kk = tid & 15;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int j = (tid >> 4) + 16 * p;
        Bs_re[kk][j] = rb_re[p];
        if (B_CPLX) Bs_im[kk][j] = rb_im[p];
      }
    } else {
      const int j = tid & 31;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int kk = (tid >> 5) + 8 * p;
        Bs_re[kk][j] = rb_re[p];
        if (B_CPLX) Bs_im[kk][j] = rb_im[p];
      }
    }
  };

  if (K > 0) load_tile(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    store_tile();
    __syncthreads();
    if (k0 + BK < K) load_tile(k0 + BK);   // next tile's loads overlap this tile's MFMAs
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const double a_re = As_re[kk + lk][wr + lr];
      const double b_re = Bs_re[kk + lk][wc + lr];
      acc_re = __builtin_amdgcn_mfma_f64_16x16x4f64(a_re, b_re, acc_re, 0, 0, 0);
      if (B_CPLX) {
        const double b_im = Bs_im[kk + lk][wc + lr];
        acc_im = __builtin_amdgcn_mfma_f64_16x16x4f64(a_re, b_im, acc_im, 0, 0, 0);
        if (A_CPLX) {
          const double a_im = As_im[kk + lk][wr + lr];
          acc_re = __builtin_amdgcn_mfma_f64_16x16x4f64(-a_im, b_im, acc_re, 0, 0, 0);
          acc_im = __builtin_amdgcn_mfma_f64_16x16x4f64(a_im, b_re, acc_im, 0, 0, 0);
        }
      } else if (A_CPLX) {
        const double a_im = As_im[kk + lk][wr + lr];
        acc_im = __builtin_amdgcn_mfma_f64_16x16x4f64(a_im, b_re, acc_im, 0, 0, 0);
      }
    }
    __syncthreads();
  }

  // ---- epilogue: v_mfma_f64_16x16x4f64 leaves D(i, j) in lane j + 16*(i%4), register i/4,
  //      so lane holds rows wr + 4*r + lane/16 (r = 0..3) of column wc + lane%16 ----
  const int gj = j0 + wc + lr;
  if (gj >= g.N) return;
  cplx csj = g.cs ? g.cs[gj] : make_double2(1.0, 0.0);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int gi = i0 + wr + 4 * r + lk;
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

}  // namespace

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
  return g;
}

int gemm_launch(const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return ST_NONE;
  if (g.batch > 65535) { alfq_set_error("gemm_launch: batch too large"); return ST_GENERIC; }
  dim3 grid((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, g.batch > 1 ? g.batch : 1);
  const bool bc = g.b_cplx || g.bs != nullptr;
  if (g.a_cplx && bc) hipLaunchKernelGGL((gemm_kernel<true, true>), grid, dim3(NT), 0, s, g);
  else if (g.a_cplx) hipLaunchKernelGGL((gemm_kernel<true, false>), grid, dim3(NT), 0, s, g);
  else if (bc) hipLaunchKernelGGL((gemm_kernel<false, true>), grid, dim3(NT), 0, s, g);
  else hipLaunchKernelGGL((gemm_kernel<false, false>), grid, dim3(NT), 0, s, g);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}
