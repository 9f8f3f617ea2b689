// Error-free split products on the gfx950 int8 matrix cores: the hopping wraps of G at
// FP64-class accuracy (ALFQ_PREC_INT8_SPLIT).
//
// Split (oz_split_kernel): a workgroup takes 16 rows (or columns) of one chain's operand,
// finds each one's largest |re|, |im| over the contraction index, and writes OZ_S truncated
// base-128 digits of every entry scaled by 2^-e into int8 planes whose contraction index is
// contiguous (zero-padded to a multiple of 64), with the exponent e.  The digits represent
// the entry exactly up to 2^(-7 OZ_S) of its row's range.
//
// Product (oz_gemm_kernel): a workgroup of 8 waves owns a 64 x 64 complex output tile;
// waves 0-3 the real plane, 4-7 the imaginary one, each a 32 x 32 block.  K advances in
// stages of 32 through LDS holding every digit plane of both operands for the stage, the
// next stage's planes loaded into registers while the current one is multiplied.  For each k-step of 32 a wave keeps the OZ_S digit fragments of
// its B plane(s) in registers and, for each digit s of A, issues v_mfma_i32_32x32x32_i8 into the int32 accumulator of
// s + t for every t with s + t <= OZ_S + 1: the sums of one s + t are exact.  At the end
// the OZ_S sums are weighted by 2^(-7 (s + t)) and added in FP64 (smallest first), the two
// planes meet in LDS, and every thread applies the exponents, alpha, the row and column
// scales and beta to its outputs, written column-major and coalesced.
#include "alfq_oz.h"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int OZ_T = 64;             // output tile (rows and columns)
constexpr int OZ_NT = 512;           // threads per product workgroup

// ---- split ---------------------------------------------------------------------------
template <bool BY_ROW, bool CPLX>
__global__ __launch_bounds__(256) void oz_split_kernel(const void* __restrict__ X, int ld, int nidx, int nk,
                                                       const cplx* __restrict__ kscale, long long xs,
                                                       int8_t* __restrict__ out, int* __restrict__ ex, int rows,
                                                       int planes, long long kp, long long os) {
  constexpr int KL = 16;                   // contraction entries per lane and chunk: one 16-byte store
  const long long z = blockIdx.z;
  X = (const char*)X + z * xs;
  if (kscale) kscale = (const cplx*)((const char*)kscale + z * xs);
  out += z * os;
  ex = (int*)((char*)ex + z * os);
  const int t = threadIdx.x;
  // 16 lanes per row (column); consecutive lanes read consecutive addresses of the source
  const int il = BY_ROW ? (t & 15) : (t >> 4);
  const int kl = BY_ROW ? (t >> 4) : (t & 15);
  const int idx = blockIdx.x * 16 + il;
  const bool live = idx < nidx;
  __shared__ double red[16][17];

  auto load = [&](int k, double& re, double& im) {
    re = 0.0;
    im = 0.0;
    if (!live || k >= nk) return;
    const long long off = BY_ROW ? (idx + (long long)k * ld) : (k + (long long)idx * ld);
    if (CPLX) {
      cplx v = ((const cplx*)X)[off];
      if (kscale) v = cmul(v, kscale[k]);
      re = v.x;
      im = v.y;
    } else {
      re = ((const double*)X)[off];
    }
  };

  double m = 0.0;
  for (int k0 = 0; k0 < nk; k0 += 16 * KL) {
#pragma unroll
    for (int u = 0; u < KL; ++u) {
      double re, im;
      load(k0 + KL * kl + u, re, im);
      m = fmax(m, fmax(fabs(re), fabs(im)));
    }
  }
  red[il][kl] = m;
  __syncthreads();
  if (kl == 0) {
    double mm = 0.0;
    for (int q = 0; q < 16; ++q) mm = fmax(mm, red[il][q]);
    red[il][16] = mm;
  }
  __syncthreads();
  m = red[il][16];
  int e = 0;
  if (m > 0.0) frexp(m, &e);               // m 2^-e in [1/2, 1)
  if (live && kl == 0) ex[idx] = e;
  if (!live) return;

  for (long long k0 = 0; k0 < kp; k0 += 16 * KL) {
    const int k = (int)k0 + KL * kl;
    if (k >= kp) break;                    // kp is a multiple of 64 and of KL
    double vr[KL], vi[KL];
#pragma unroll
    for (int u = 0; u < KL; ++u) {
      load(k + u, vr[u], vi[u]);
      vr[u] = ldexp(vr[u], -e);
      vi[u] = ldexp(vi[u], -e);
    }
#pragma unroll
    for (int p = 0; p < (CPLX ? 2 : 1); ++p) {
      double* r = p ? vi : vr;
#pragma unroll
      for (int s = 0; s < OZ_S; ++s) {
        v4i w = {0, 0, 0, 0}, wn = {0, 0, 0, 0};
#pragma unroll
        for (int u = 0; u < KL; ++u) {
          r[u] *= 128.0;
          const double d = trunc(r[u]);
          r[u] -= d;
          const int di = (int)d;            // |di| <= 127
          w[u >> 2] |= (int)((unsigned)(di & 0xff) << (8 * (u & 3)));
          wn[u >> 2] |= (int)((unsigned)((-di) & 0xff) << (8 * (u & 3)));
        }
        *(v4i*)(out + ((long long)(p * OZ_S + s) * rows + idx) * kp + k) = w;
        if (p == 1 && planes == 3) *(v4i*)(out + ((long long)(2 * OZ_S + s) * rows + idx) * kp + k) = wn;
      }
    }
  }
}

// ---- product -------------------------------------------------------------------------
// Variant V: 0 real A x complex B, 1 complex A x real B, 2 A with the negated imaginary
// plane x complex B, 3 complex A x B with the negated imaginary plane.  Output plane q is
// the sum over its pairs (A plane, B plane).
template <int V> struct OzVar;
template <> struct OzVar<0> { static constexpr int PA = 1, PB = 2, NP = 1; };
template <> struct OzVar<1> { static constexpr int PA = 2, PB = 1, NP = 1; };
template <> struct OzVar<2> { static constexpr int PA = 3, PB = 2, NP = 2; };
template <> struct OzVar<3> { static constexpr int PA = 2, PB = 3, NP = 2; };

template <int V>
__device__ constexpr int pair_a(int q, int p) {
  return V == 0 ? 0 : V == 1 ? q : V == 2 ? (q == 0 ? (p == 0 ? 0 : 2) : (p == 0 ? 0 : 1)) : (p == 0 ? 0 : 1);
}
template <int V>
__device__ constexpr int pair_b(int q, int p) {
  return V == 0 ? q : V == 1 ? 0 : V == 2 ? (q == 0 ? p : (p == 0 ? 1 : 0)) : (q == 0 ? (p == 0 ? 0 : 2) : (p == 0 ? 1 : 0));
}

struct OzGemmArgs {
  int M, N, K;
  const int8_t* A; const int* ea; long long akp; int arows; long long as;
  const int8_t* B; const int* eb; long long bkp; int brows; long long bs;
  cplx* C; int ldc; cplx alpha, beta; const cplx* rs; const cplx* cs; long long cst;
};

// byte offset of 16-byte segment sg of LDS row `row` (rows of KC bytes; the segment index
// is swizzled with the row so that 32 lanes reading one segment of 32 rows spread over banks)
template <int KC>
__device__ __forceinline__ int lds_off(int tile, int row, int sg) {
  return (tile * OZ_T + row) * KC + 16 * (sg ^ ((row >> 2) & (KC / 16 - 1)));
}

template <int V, int Q, int KC>
__device__ __forceinline__ void oz_kstep(const int8_t* lds, int bi, int bj, int lane, int kk, v16i (&acc)[OZ_S]) {
  typedef OzVar<V> T;
  const int r = lane & 31, sg = 2 * kk + (lane >> 5);
#pragma unroll
  for (int p = 0; p < T::NP; ++p) {
    const int ap = pair_a<V>(Q, p), bp = pair_b<V>(Q, p);
    v4i bf[OZ_S];
#pragma unroll
    for (int t = 0; t < OZ_S; ++t)
      bf[t] = *(const v4i*)(lds + lds_off<KC>(T::PA * OZ_S + bp * OZ_S + t, 32 * bj + r, sg));
#pragma unroll
    for (int s = 0; s < OZ_S; ++s) {
      const v4i af = *(const v4i*)(lds + lds_off<KC>(ap * OZ_S + s, 32 * bi + r, sg));
#pragma unroll
      for (int t = 0; t + s < OZ_S; ++t) acc[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af, bf[t], acc[s + t], 0, 0, 0);
    }
  }
}

template <int V>
__global__ __launch_bounds__(OZ_NT, 1) void oz_gemm_kernel(OzGemmArgs g) {
  typedef OzVar<V> T;
  constexpr int NTILE = (T::PA + T::PB) * OZ_S;
  // contraction bytes per LDS stage: 32 (one k-step), so that the next stage's segments fit in
  // registers beside the accumulators
  constexpr int KC = 32;
  constexpr int SEGS = KC / 16;
  constexpr int STAGE = NTILE * OZ_T * KC;
  constexpr int RES = 2 * OZ_T * (OZ_T + 1) * (int)sizeof(double);
  __shared__ __align__(16) int8_t lds[STAGE > RES ? STAGE : RES];

  int bx, by, bzi;
  xcd_order(bx, by, bzi);
  const long long z = bzi;
  const int8_t* A = g.A + z * g.as;
  const int8_t* B = g.B + z * g.bs;
  const int* ea = (const int*)((const char*)g.ea + z * g.as);
  const int* eb = (const int*)((const char*)g.eb + z * g.bs);
  cplx* C = (cplx*)((char*)g.C + z * g.cst);
  const cplx* rs = g.rs ? (const cplx*)((const char*)g.rs + z * g.cst) : nullptr;
  const cplx* cs = g.cs ? (const cplx*)((const char*)g.cs + z * g.cst) : nullptr;

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int q = wave >> 2, bi = wave & 1, bj = (wave >> 1) & 1;
  const int i0 = bx * OZ_T, j0 = by * OZ_T;

  v16i acc[OZ_S];
#pragma unroll
  for (int s = 0; s < OZ_S; ++s)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[s][r] = 0;

  // a stage: every digit plane of A's 64 rows and B's 64 columns for k0 .. k0 + KC - 1, NPRE
  // 16-byte segments per thread; the next stage's segments are loaded into registers while
  // the current one is multiplied, and stored to LDS after the barrier that ends it
  constexpr int NSEG = NTILE * OZ_T * SEGS, NPRE = (NSEG + OZ_NT - 1) / OZ_NT;
  v4i pre[NPRE];
  auto gload = [&](int k0) {
#pragma unroll
    for (int p = 0; p < NPRE; ++p) {
      const int seg = tid + p * OZ_NT;
      const int tile = seg / (OZ_T * SEGS), rem = seg % (OZ_T * SEGS), row = rem / SEGS, sg = rem % SEGS;
      v4i v = {0, 0, 0, 0};
      if (seg < NSEG) {
        if (tile < T::PA * OZ_S) {
          const int gi = i0 + row;
          if (gi < g.M) v = *(const v4i*)(A + ((long long)tile * g.arows + gi) * g.akp + k0 + 16 * sg);
        } else {
          const int gj = j0 + row;
          if (gj < g.N)
            v = *(const v4i*)(B + ((long long)(tile - T::PA * OZ_S) * g.brows + gj) * g.bkp + k0 + 16 * sg);
        }
      }
      pre[p] = v;
    }
  };
  auto lstore = [&]() {
#pragma unroll
    for (int p = 0; p < NPRE; ++p) {
      const int seg = tid + p * OZ_NT;
      const int tile = seg / (OZ_T * SEGS), rem = seg % (OZ_T * SEGS), row = rem / SEGS, sg = rem % SEGS;
      if (seg < NSEG) *(v4i*)(lds + lds_off<KC>(tile, row, sg)) = pre[p];
    }
  };
  if (g.K > 0) gload(0);
  for (int k0 = 0; k0 < g.K; k0 += KC) {
    lstore();
    __syncthreads();
    if (k0 + KC < g.K) gload(k0 + KC);
#pragma unroll
    for (int kk = 0; kk < KC / 32; ++kk) {
      if (q == 0) oz_kstep<V, 0, KC>(lds, bi, bj, lane, kk, acc);
      else oz_kstep<V, 1, KC>(lds, bi, bj, lane, kk, acc);
    }
    __syncthreads();
  }

  // weighted FP64 sum of the digit-pair sums: D(i, j) of a 32 x 32 i32 block sits in lane
  // j + 32 ((i / 4) % 2), register 4 (i / 8) + i % 4
  double* res = (double*)lds;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    double v = 0.0;
#pragma unroll
    for (int gq = OZ_S - 1; gq >= 0; --gq) v += ldexp((double)acc[gq][r], -7 * (gq + 2));
    const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3), j = lane & 31;
    res[(q * OZ_T + 32 * bi + i) * (OZ_T + 1) + 32 * bj + j] = v;
  }
  __syncthreads();
  for (int o = tid; o < OZ_T * OZ_T; o += OZ_NT) {
    const int i = o & (OZ_T - 1), j = o >> 6;
    const int gi = i0 + i, gj = j0 + j;
    if (gi >= g.M || gj >= g.N) continue;
    const int e = ea[gi] + eb[gj];
    cplx v = make_double2(ldexp(res[i * (OZ_T + 1) + j], e), ldexp(res[(OZ_T + i) * (OZ_T + 1) + j], e));
    v = cmul(g.alpha, v);
    if (rs) v = cmul(rs[gi], v);
    if (cs) v = cmul(v, cs[gj]);
    const long long off = gi + (long long)gj * g.ldc;
    if (g.beta.x != 0.0 || g.beta.y != 0.0) v = cadd(v, cmul(g.beta, C[off]));
    C[off] = v;
  }
}

}  // namespace

int oz_split(const void* X, bool x_cplx, int ld, int nidx, int nk, bool by_row, const cplx* kscale, long long xs,
             const OzOperand& out, int nchain, hipStream_t s) {
  if (nidx <= 0 || nk < 0 || nidx > out.rows || oz_pitch(nk) > out.kp || (!x_cplx && kscale) ||
      (x_cplx ? (out.planes != 2 && out.planes != 3) : out.planes != 1)) {
    alfq_set_error("oz_split: operand does not fit its digit planes");
    return ST_GENERIC;
  }
  const dim3 grid((unsigned)((nidx + 15) / 16), 1, (unsigned)nchain);
  if (by_row) {
    if (x_cplx) hipLaunchKernelGGL((oz_split_kernel<true, true>), grid, dim3(256), 0, s, X, ld, nidx, nk, kscale, xs,
                                   out.d, out.e, out.rows, out.planes, out.kp, out.stride);
    else hipLaunchKernelGGL((oz_split_kernel<true, false>), grid, dim3(256), 0, s, X, ld, nidx, nk, kscale, xs,
                            out.d, out.e, out.rows, out.planes, out.kp, out.stride);
  } else {
    if (x_cplx) hipLaunchKernelGGL((oz_split_kernel<false, true>), grid, dim3(256), 0, s, X, ld, nidx, nk, kscale, xs,
                                   out.d, out.e, out.rows, out.planes, out.kp, out.stride);
    else hipLaunchKernelGGL((oz_split_kernel<false, false>), grid, dim3(256), 0, s, X, ld, nidx, nk, kscale, xs,
                            out.d, out.e, out.rows, out.planes, out.kp, out.stride);
  }
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

int oz_gemm(int M, int N, int K, const OzOperand& A, const OzOperand& B, cplx* C, int ldc, cplx alpha, cplx beta,
            const cplx* rs, const cplx* cs, long long cstride, int nchain, hipStream_t s) {
  if (M <= 0 || N <= 0) return ST_NONE;
  if (M > A.rows || N > B.rows || oz_pitch(K) > A.kp || oz_pitch(K) > B.kp || nchain < 1 || nchain > 65535) {
    alfq_set_error("oz_gemm: operands do not match the product");
    return ST_GENERIC;
  }
  OzGemmArgs g{M, N, K, A.d, A.e, A.kp, A.rows, A.stride, B.d, B.e, B.kp, B.rows, B.stride,
               C, ldc, alpha, beta, rs, cs, cstride};
  const dim3 grid((unsigned)((M + OZ_T - 1) / OZ_T), (unsigned)((N + OZ_T - 1) / OZ_T), (unsigned)nchain);
  if (A.planes == 1 && B.planes == 2) hipLaunchKernelGGL(oz_gemm_kernel<0>, grid, dim3(OZ_NT), 0, s, g);
  else if (A.planes == 2 && B.planes == 1) hipLaunchKernelGGL(oz_gemm_kernel<1>, grid, dim3(OZ_NT), 0, s, g);
  else if (A.planes == 3 && B.planes == 2) hipLaunchKernelGGL(oz_gemm_kernel<2>, grid, dim3(OZ_NT), 0, s, g);
  else if (A.planes == 2 && B.planes == 3) hipLaunchKernelGGL(oz_gemm_kernel<3>, grid, dim3(OZ_NT), 0, s, g);
  else {
    alfq_set_error("oz_gemm: unsupported operand planes");
    return ST_GENERIC;
  }
  HIP_OK(hipGetLastError());
  return ST_NONE;
}
