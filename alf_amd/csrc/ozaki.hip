// Error-free split products on the gfx950 int8 matrix cores: the hopping wraps of G at
// FP64-class accuracy (ALFQ_PREC_INT8_SPLIT).
//
// Split (oz_split_kernel): a workgroup takes 16 rows (or 8 columns) of one chain's operand,
// finds each one's largest |re|, |im| over the contraction index, and writes OZ_S rounded
// base-128 digits of every entry scaled by 2^-e into int8 planes laid out by k-block of 32
// (every digit of a block contiguous; zero-padded to a multiple of 64), with the exponent e.
// The digits represent the entry exactly up to 2^(-7 OZ_S) of its row's range.  A source row
// of at most 256 entries is read once.
//
// Product (oz_gemm_kernel): each wave owns a 32 x 32 block of one output plane (real or
// imaginary); a real x complex product runs 4 waves on a 64 x 32 or 32 x 64 tile and two
// workgroups per CU, so that one multiplies while the other waits for its next stage.  K
// advances in stages of 32 through LDS holding every digit plane of both operands for the
// stage (whole 256-byte row blocks from global memory, stored conflict-free), the next
// stage loaded into registers while the current one is multiplied.  A wave keeps the OZ_S
// digit fragments of its B plane(s) in registers and, for each digit s of A, issues
// v_mfma_i32_32x32x32_i8 into the int32 accumulator of s + t for every t with
// s + t <= OZ_S - 1 (0-based): the sums of one s + t are exact.  At the end the OZ_S sums
// are weighted by 2^(-7 (s + t + 2)) and added in FP64 (smallest first), the two planes
// meet in LDS, and every thread applies the exponents, alpha, the row and column scales and
// beta to its outputs, written column-major and coalesced.
#include "alfq_oz.h"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// Output tile per variant: each wave owns one 32 x 32 block of one output plane.  The real x
// complex products of the sweep take 4 waves and two workgroups per CU (one computes while the
// other waits for its stage), with the complex operand on the tile's short side so that a
// stage fits the registers that prefetch it; complex x complex keeps 64 x 64 and 8 waves.
template <int V> struct OzTile;
template <> struct OzTile<0> { static constexpr int TM = 64, TN = 32, NT = 256, WPC = 2; };
template <> struct OzTile<1> { static constexpr int TM = 32, TN = 64, NT = 256, WPC = 2; };
template <> struct OzTile<2> { static constexpr int TM = 64, TN = 64, NT = 512, WPC = 1; };
template <> struct OzTile<3> { static constexpr int TM = 64, TN = 64, NT = 512, WPC = 1; };

// ---- split ----------------------------------------------------------------------------
// A workgroup of 256 threads splits RPW rows (columns), 256 / RPW lanes per row, each lane KL
// consecutive contraction entries per chunk: a row of up to 256 entries is one chunk, read
// once.  By row (entries of a row strided in memory): 16 rows of 16 lanes x 16 entries, so
// 16 lanes read 256 contiguous bytes of a column; by column (contiguous): 8 columns of 32
// lanes x 8 entries (measured 129 against 150 us for a batch of 256 at n = 256).
template <bool BY_ROW> struct OzSplitGeom { static constexpr int RPW = BY_ROW ? 16 : 8, KL = BY_ROW ? 16 : 8; };
// Digit d of a scaled entry x (|x| < 1): d = clamp(rint(128 x), -127, 127), x <- 128 x - d.
// Rounding to nearest leaves remainders in [-1/2, 1/2] (in [1/2, 1) only after a saturated
// digit: an entry within 2^-8 of 2^e saturates the first digit, one within about 2^-15 the
// second too), so the digit sums are unbiased, and after OZ_S digits |x| < 1 as with
// truncation.
__device__ __forceinline__ int pack4(int a, int b, int c, int d) {
  return (a & 255) | ((b & 255) << 8) | ((c & 255) << 16) | (d << 24);
}

template <bool BY_ROW, bool CPLX, bool NEG>
__global__ __launch_bounds__(256) void oz_split_kernel(const void* __restrict__ X, int ld, int nidx, int nk,
                                                       const cplx* __restrict__ kscale, long long xs,
                                                       int8_t* __restrict__ out, int* __restrict__ ex, int rows,
                                                       long long kp, long long os) {
  constexpr int OZ_SPLIT_RPW = OzSplitGeom<BY_ROW>::RPW;
  constexpr int KL = OzSplitGeom<BY_ROW>::KL;  // contraction entries per lane and chunk: one KL-byte store
  constexpr int LPR = 256 / OZ_SPLIT_RPW;  // lanes per row (column)
  constexpr int CH = LPR * KL;             // contraction entries per chunk of the lanes of an index
  typedef int vst __attribute__((ext_vector_type(KL / 4)));
  const long long z = blockIdx.z;
  X = (const char*)X + z * xs;
  if (kscale) kscale = (const cplx*)((const char*)kscale + z * xs);
  out += z * os;
  ex = (int*)((char*)ex + z * os);
  const int t = threadIdx.x;
  // LPR lanes per row (column); consecutive lanes read consecutive addresses of the source
  const int il = BY_ROW ? (t % OZ_SPLIT_RPW) : (t / LPR);
  const int kl = BY_ROW ? (t / OZ_SPLIT_RPW) : (t % LPR);
  const int idx = blockIdx.x * OZ_SPLIT_RPW + il;
  const bool live = idx < nidx;
  __shared__ double red[OZ_SPLIT_RPW][LPR + 1];

  double vr[KL], vi[KL];
  auto load_chunk = [&](long long k0) {
#pragma unroll
    for (int u = 0; u < KL; ++u) {
      const long long k = k0 + KL * kl + u;
      vr[u] = 0.0;
      vi[u] = 0.0;
      if (!live || k >= nk) continue;
      const long long off = BY_ROW ? (idx + k * ld) : (k + (long long)idx * ld);
      if (CPLX) {
        cplx v = ((const cplx*)X)[off];
        if (kscale) v = cmul(v, kscale[k]);
        vr[u] = v.x;
        vi[u] = v.y;
      } else {
        vr[u] = ((const double*)X)[off];
      }
    }
  };
  // a contraction length of one chunk is read once: its entries stay in registers between the
  // range and the digits
  const bool one = kp <= CH;
  double m = 0.0;
  for (long long k0 = 0; k0 < nk; k0 += CH) {
    load_chunk(k0);
#pragma unroll
    for (int u = 0; u < KL; ++u) m = fmax(m, fmax(fabs(vr[u]), fabs(vi[u])));
  }
  red[il][kl] = m;
  __syncthreads();
  if (kl == 0) {
    double mm = 0.0;
    for (int q = 0; q < LPR; ++q) mm = fmax(mm, red[il][q]);
    red[il][LPR] = mm;
  }
  __syncthreads();
  m = red[il][LPR];
  int e = 0;
  if (m > 0.0) frexp(m, &e);               // m 2^-e in [1/2, 1)
  if (live && kl == 0) ex[idx] = e;
  if (!live) return;

  for (long long k0 = 0; k0 < kp; k0 += CH) {
    const long long k = k0 + KL * kl;
    if (k >= kp) break;                    // kp is a multiple of 64 and of KL
    if (!one) load_chunk(k0);
    // digit s of entries k .. k + KL - 1 of this row: bytes (k % 32) .. of the digit's
    // 32-byte segment in k-block k / 32
    const long long kb = (k >> 5) * (32 * OZ_S) + (k & 31);
#pragma unroll
    for (int p = 0; p < (CPLX ? 2 : 1); ++p) {
      double* r = p ? vi : vr;
      int8_t* o = out + (long long)(p * rows + idx) * kp * OZ_S + kb;
      int8_t* on = out + (long long)(2 * rows + idx) * kp * OZ_S + kb;
#pragma unroll
      for (int u = 0; u < KL; ++u) r[u] = ldexp(r[u], -e);
#pragma unroll
      for (int s = 0; s < OZ_S; ++s) {
        int di[KL];
#pragma unroll
        for (int u = 0; u < KL; ++u) {
          const double tv = r[u] * 128.0;
          const double d = fmin(fmax(rint(tv), -127.0), 127.0);
          r[u] = tv - d;
          di[u] = (int)d;
        }
        vst w;
#pragma unroll
        for (int q = 0; q < KL / 4; ++q) w[q] = pack4(di[4 * q], di[4 * q + 1], di[4 * q + 2], di[4 * q + 3]);
        *(vst*)(o + 32 * s) = w;
        if (NEG && p == 1) {
#pragma unroll
          for (int q = 0; q < KL / 4; ++q) w[q] = pack4(-di[4 * q], -di[4 * q + 1], -di[4 * q + 2], -di[4 * q + 3]);
          *(vst*)(on + 32 * s) = w;
        }
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

// LDS holds a stage as the global layout does: one 256-byte row (OZ_S digits x 32 bytes) per
// operand row and plane, A's rows first.  16-byte segment j of row r sits at slot
// j ^ (r % 16), so the 8 lanes of one LDS cycle hit 8 different slots both when a stage is
// stored (consecutive lanes: consecutive segments of one row) and when a fragment is read
// (consecutive lanes: one segment of consecutive rows).
constexpr int OZ_ROW = OZ_S * 32;
__device__ __forceinline__ int lds_seg(int row, int j) { return row * OZ_ROW + 16 * (j ^ (row & 15)); }

template <int V, int Q>
__device__ __forceinline__ void oz_kstep(const int8_t* lds, int bi, int bj, int lane, v16i (&acc)[OZ_S]) {
  typedef OzVar<V> T;
  typedef OzTile<V> W;
  const int r = lane & 31, h = lane >> 5;
#pragma unroll
  for (int p = 0; p < T::NP; ++p) {
    const int ap = pair_a<V>(Q, p), bp = pair_b<V>(Q, p);
    const int arow = ap * W::TM + 32 * bi + r, brow = T::PA * W::TM + bp * W::TN + 32 * bj + r;
    v4i bf[OZ_S];
#pragma unroll
    for (int t = 0; t < OZ_S; ++t) bf[t] = *(const v4i*)(lds + lds_seg(brow, 2 * t + h));
#pragma unroll
    for (int s = 0; s < OZ_S; ++s) {
      const v4i af = *(const v4i*)(lds + lds_seg(arow, 2 * s + h));
#pragma unroll
      for (int t = 0; t + s < OZ_S; ++t) acc[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af, bf[t], acc[s + t], 0, 0, 0);
    }
  }
}

template <int V>
__global__ __launch_bounds__(OzTile<V>::NT, OzTile<V>::WPC) void oz_gemm_kernel(OzGemmArgs g) {
  typedef OzVar<V> T;
  typedef OzTile<V> W;
  constexpr int OZ_TM = W::TM, OZ_TN = W::TN, OZ_NT = W::NT;
  // contraction entries per LDS stage: 32 (one k-step), so that the next stage's segments fit
  // in registers beside the accumulators
  constexpr int KC = 32;
  constexpr int STAGE = (T::PA * OZ_TM + T::PB * OZ_TN) * OZ_ROW;
  constexpr int RES = 2 * OZ_TM * (OZ_TN + 1) * (int)sizeof(double);
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
  constexpr int NBI = OZ_TM / 32, NBJ = OZ_TN / 32;
  static_assert(2 * NBI * NBJ * 64 == OZ_NT, "one wave per 32 x 32 block of each plane");
  const int q = wave / (NBI * NBJ), bi = wave % NBI, bj = (wave / NBI) % NBJ;
  const int i0 = bx * OZ_TM, j0 = by * OZ_TN;

  v16i acc[OZ_S];
#pragma unroll
  for (int s = 0; s < OZ_S; ++s)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[s][r] = 0;

  // a stage of one row of one plane is OZ_ROW contiguous bytes (every digit of its k-block);
  // LDS row lr = (plane, row) of A, then of B; thread segment seg = 16 lr + j, so
  // consecutive threads read consecutive 16-byte segments, whole cache lines
  constexpr int RSEG = OZ_ROW / 16;
  constexpr int NAR = T::PA * OZ_TM, NR = NAR + T::PB * OZ_TN;
  constexpr int NSEG = NR * RSEG, NPRE = (NSEG + OZ_NT - 1) / OZ_NT;
  v4i pre[NPRE];
  auto gload = [&](int k0) {
#pragma unroll
    for (int p = 0; p < NPRE; ++p) {
      const int seg = tid + p * OZ_NT;
      const int lr = seg / RSEG, j = seg % RSEG;
      v4i v = {0, 0, 0, 0};
      if (lr < NAR) {
        const int pl = lr / OZ_TM, gi = i0 + lr % OZ_TM;
        if (gi < g.M) v = *(const v4i*)(A + ((long long)pl * g.arows + gi) * g.akp * OZ_S + (long long)k0 * OZ_S + 16 * j);
      } else if (lr < NR) {
        const int pl = (lr - NAR) / OZ_TN, gj = j0 + (lr - NAR) % OZ_TN;
        if (gj < g.N) v = *(const v4i*)(B + ((long long)pl * g.brows + gj) * g.bkp * OZ_S + (long long)k0 * OZ_S + 16 * j);
      }
      pre[p] = v;
    }
  };
  auto lstore = [&]() {
#pragma unroll
    for (int p = 0; p < NPRE; ++p) {
      const int seg = tid + p * OZ_NT;
      if (seg < NSEG) *(v4i*)(lds + lds_seg(seg / RSEG, seg % RSEG)) = pre[p];
    }
  };
  if (g.K > 0) gload(0);
  for (int k0 = 0; k0 < g.K; k0 += KC) {
    lstore();
    __syncthreads();
    if (k0 + KC < g.K) gload(k0 + KC);
    if (q == 0) oz_kstep<V, 0>(lds, bi, bj, lane, acc);
    else oz_kstep<V, 1>(lds, bi, bj, lane, acc);
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
    res[(q * OZ_TM + 32 * bi + i) * (OZ_TN + 1) + 32 * bj + j] = v;
  }
  __syncthreads();
  for (int o = tid; o < OZ_TM * OZ_TN; o += OZ_NT) {
    const int i = o & (OZ_TM - 1), j = o / OZ_TM;
    const int gi = i0 + i, gj = j0 + j;
    if (gi >= g.M || gj >= g.N) continue;
    const int e = ea[gi] + eb[gj];
    cplx v = make_double2(ldexp(res[i * (OZ_TN + 1) + j], e), ldexp(res[(OZ_TM + i) * (OZ_TN + 1) + j], e));
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
  const int rpw = by_row ? OzSplitGeom<true>::RPW : OzSplitGeom<false>::RPW;
  const dim3 grid((unsigned)((nidx + rpw - 1) / rpw), 1, (unsigned)nchain);
  auto launch = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, X, ld, nidx, nk, kscale, xs, out.d, out.e, out.rows, out.kp,
                       out.stride);
  };
  const bool neg = out.planes == 3;
  if (by_row) {
    if (!x_cplx) launch(oz_split_kernel<true, false, false>);
    else if (neg) launch(oz_split_kernel<true, true, true>);
    else launch(oz_split_kernel<true, true, false>);
  } else {
    if (!x_cplx) launch(oz_split_kernel<false, false, false>);
    else if (neg) launch(oz_split_kernel<false, true, true>);
    else launch(oz_split_kernel<false, true, false>);
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
  auto launch = [&](auto kern, int tm, int tn, int nt) {
    const dim3 grid((unsigned)((M + tm - 1) / tm), (unsigned)((N + tn - 1) / tn), (unsigned)nchain);
    hipLaunchKernelGGL(kern, grid, dim3(nt), 0, s, g);
  };
  if (A.planes == 1 && B.planes == 2) launch(oz_gemm_kernel<0>, OzTile<0>::TM, OzTile<0>::TN, OzTile<0>::NT);
  else if (A.planes == 2 && B.planes == 1) launch(oz_gemm_kernel<1>, OzTile<1>::TM, OzTile<1>::TN, OzTile<1>::NT);
  else if (A.planes == 3 && B.planes == 2) launch(oz_gemm_kernel<2>, OzTile<2>::TM, OzTile<2>::TN, OzTile<2>::NT);
  else if (A.planes == 2 && B.planes == 3) launch(oz_gemm_kernel<3>, OzTile<3>::TM, OzTile<3>::TN, OzTile<3>::NT);
  else {
    alfq_set_error("oz_gemm: unsupported operand planes");
    return ST_GENERIC;
  }
  HIP_OK(hipGetLastError());
  return ST_NONE;
}
