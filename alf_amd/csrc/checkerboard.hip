// Checkerboard hopping exponential: products of 2 x 2 bond exponentials applied to n x n
// matrices (Prog/Hop_mod.F90:143-250, OpTTypes_mod.F90:150-304, ZDSLSYMM/ZSLHEMM of
// Libraries/Modules/Mat_subroutines_mod.F90:518-964 for N = 2).
//
// A bond acts on two rows (left product) or two columns (right product).  Each workgroup
// stages a strip of the matrix in LDS (whole columns for the left product, whole rows for
// the right one), applies the families one after another with a barrier between them (the
// bonds of a family are on disjoint sites and commute) and writes the strip back with the
// diagonal scalings of the slice fused into the load and the store.  One pass over HBM per
// application: 2 * 16 * n^2 bytes per chain.
#include "alfq_cb.h"

namespace {

constexpr int CB_NT = 256;
constexpr long long CB_LDS_MAX = 65536;

__device__ __forceinline__ cplx mul2(cplx a, cplx x, cplx b, cplx y) {   // a x + b y
  return cadd(cmul(a, x), cmul(b, y));
}

__global__ __launch_bounds__(CB_NT) void cb_left_kernel(cplx* out, const cplx* in, int n, int w,
                                                        const cplx* pre_rs, const cplx* post_rs, const int2* sites,
                                                        const cplx* mats, long long mstride, CbFams fam, int reverse,
                                                        long long cs) {
  extern __shared__ cplx S[];     // S[c * n + r]: column c of the strip
  out = chain_ptr(out, cs);
  in = chain_ptr(in, cs);
  pre_rs = chain_ptr(pre_rs, cs);
  post_rs = chain_ptr(post_rs, cs);
  mats = (const cplx*)((const char*)mats + (long long)blockIdx.z * mstride);
  const int j0 = blockIdx.x * w;
  const int nc = min(w, n - j0);
  const int tot = nc * n;
  const cplx* src = in + (long long)j0 * n;
  for (int e = threadIdx.x; e < tot; e += CB_NT) {
    cplx x = src[e];
    if (pre_rs) x = cmul(x, pre_rs[e % n]);
    S[e] = x;
  }
  __syncthreads();
  for (int q = 0; q < fam.nfam; ++q) {
    const int f = reverse ? fam.nfam - 1 - q : q;
    const int b0 = fam.off[f];
    const int nb = fam.off[f + 1] - b0;
    for (int e = threadIdx.x; e < nb * nc; e += CB_NT) {
      const int c = e / nb;
      const int b = b0 + (e - c * nb);
      const int2 p = sites[b];
      const cplx* m = mats + 4 * b;
      cplx* col = S + c * n;
      const cplx x1 = col[p.x], x2 = col[p.y];
      col[p.x] = mul2(m[0], x1, m[2], x2);
      col[p.y] = mul2(m[1], x1, m[3], x2);
    }
    __syncthreads();
  }
  cplx* dst = out + (long long)j0 * n;
  for (int e = threadIdx.x; e < tot; e += CB_NT) {
    cplx x = S[e];
    if (post_rs) x = cmul(x, post_rs[e % n]);
    dst[e] = x;
  }
}

__global__ __launch_bounds__(CB_NT) void cb_right_kernel(cplx* out, const cplx* in, int n, int h,
                                                         const cplx* pre_rs, const cplx* pre_cs, const cplx* post_rs,
                                                         const cplx* post_cs, const int2* sites, const cplx* mats,
                                                         long long mstride, CbFams fam, int reverse, long long cs) {
  extern __shared__ cplx S[];     // S[c * h + r]: rows i0 .. i0 + nr - 1 of column c
  out = chain_ptr(out, cs);
  in = chain_ptr(in, cs);
  pre_rs = chain_ptr(pre_rs, cs);
  pre_cs = chain_ptr(pre_cs, cs);
  post_rs = chain_ptr(post_rs, cs);
  post_cs = chain_ptr(post_cs, cs);
  mats = (const cplx*)((const char*)mats + (long long)blockIdx.z * mstride);
  const int i0 = blockIdx.x * h;
  const int nr = min(h, n - i0);
  const int tot = nr * n;
  for (int e = threadIdx.x; e < tot; e += CB_NT) {
    const int c = e / nr;
    const int r = e - c * nr;
    cplx x = in[(long long)c * n + i0 + r];
    if (pre_rs) x = cmul(x, pre_rs[i0 + r]);
    if (pre_cs) x = cmul(x, pre_cs[c]);
    S[c * h + r] = x;
  }
  __syncthreads();
  for (int q = 0; q < fam.nfam; ++q) {
    const int f = reverse ? fam.nfam - 1 - q : q;
    const int b0 = fam.off[f];
    const int nb = fam.off[f + 1] - b0;
    for (int e = threadIdx.x; e < nb * nr; e += CB_NT) {
      const int bb = e / nr;
      const int r = e - bb * nr;
      const int b = b0 + bb;
      const int2 p = sites[b];
      const cplx* m = mats + 4 * b;
      const cplx x1 = S[p.x * h + r], x2 = S[p.y * h + r];
      // (x1 x2) [m11 m12; m21 m22]
      S[p.x * h + r] = mul2(x1, m[0], x2, m[1]);
      S[p.y * h + r] = mul2(x1, m[2], x2, m[3]);
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < tot; e += CB_NT) {
    const int c = e / nr;
    const int r = e - c * nr;
    cplx x = S[c * h + r];
    if (post_rs) x = cmul(x, post_rs[i0 + r]);
    if (post_cs) x = cmul(x, post_cs[c]);
    out[(long long)c * n + i0 + r] = x;
  }
}

// Strip width: as wide as 64 KB of LDS allows (at most 16 columns or rows), narrowed while
// the grid would have fewer than 1024 workgroups.
int pick_strip(int n, int nchain) {
  int s = 16;
  while (s > 1 && (long long)s * n * (long long)sizeof(cplx) > CB_LDS_MAX) s >>= 1;
  while (s > 1 && (long long)((n + s - 1) / s) * nchain < 1024) s >>= 1;
  return s;
}

int check_ops(int n, const CbOps& ops, int nchain) {
  if (n < 1 || (long long)n * (long long)sizeof(cplx) > CB_LDS_MAX) {
    alfq_set_error("checkerboard: Ndim outside 1..4096");
    return ST_GENERIC;
  }
  if (!ops.sites || !ops.mats || ops.fam.nfam < 1 || ops.fam.nfam > CB_MAXFAM || nchain < 1) {
    alfq_set_error("checkerboard: operators not set");
    return ST_GENERIC;
  }
  return ST_NONE;
}

}  // namespace

int cb_left(cplx* out, const cplx* in, int n, const cplx* pre_rs, const cplx* post_rs, const CbOps& ops,
            int reverse, int nchain, long long cs, hipStream_t s) {
  int st;
  if ((st = check_ops(n, ops, nchain))) return st;
  const int w = pick_strip(n, nchain);
  const dim3 grid((n + w - 1) / w, 1, nchain);
  const size_t lds = (size_t)w * n * sizeof(cplx);
  hipLaunchKernelGGL(cb_left_kernel, grid, dim3(CB_NT), lds, s, out, in, n, w, pre_rs, post_rs, ops.sites, ops.mats,
                     ops.mstride, ops.fam, reverse, cs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

int cb_right(cplx* out, const cplx* in, int n, const cplx* pre_rs, const cplx* pre_cs, const cplx* post_rs,
             const cplx* post_cs, const CbOps& ops, int reverse, int nchain, long long cs, hipStream_t s) {
  int st;
  if ((st = check_ops(n, ops, nchain))) return st;
  const int h = pick_strip(n, nchain);
  const dim3 grid((n + h - 1) / h, 1, nchain);
  const size_t lds = (size_t)h * n * sizeof(cplx);
  hipLaunchKernelGGL(cb_right_kernel, grid, dim3(CB_NT), lds, s, out, in, n, h, pre_rs, pre_cs, post_rs, post_cs,
                     ops.sites, ops.mats, ops.mstride, ops.fam, reverse, cs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}
