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
#include <algorithm>

#include "alfq_cb.h"

namespace {

constexpr int CB_NT = 256;
constexpr int CB_U = 8;          // elements a thread loads back to back
constexpr int CB_PRE = 4;        // families whose bond data a thread loads before the strip
constexpr size_t CB_LDS_MAX = 65536;

// One bond: its sites and its 2 x 2 matrix, column-major (m11, m21, m12, m22); real
// matrices (RealExpOpT) keep only the real parts.
template <bool REAL>
struct Bond;
template <>
struct Bond<false> {
  int2 p;
  cplx m0, m1, m2, m3;
};
template <>
struct Bond<true> {
  int2 p;
  double m0, m1, m2, m3;
};

__device__ __forceinline__ cplx mulm(cplx m, cplx x) { return cmul(m, x); }
__device__ __forceinline__ cplx mulm(double m, cplx x) { return cscale(x, m); }

// Thread tid's bond in family f when the family has at most CB_NT bonds: bonds are spread
// over the threads, and the groups = CB_NT / nb copies of a bond split the strip's other
// dimension (columns for the left product, rows for the right one) between them.
__device__ __forceinline__ bool my_bond(const CbFams& fam, int f, int tid, int& b, int& g, int& groups) {
  const int b0 = fam.off[f];
  const int nb = fam.off[f + 1] - b0;
  if (nb <= 0 || nb > CB_NT) return false;
  groups = CB_NT / nb;
  g = tid / nb;
  b = b0 + (tid - g * nb);
  return g < groups;
}

template <bool REAL>
__device__ __forceinline__ Bond<REAL> load_bond(const int2* sites, const cplx* mats, int b) {
  Bond<REAL> o;
  o.p = sites[b];
  if constexpr (REAL) {
    o.m0 = mats[4 * b].x;
    o.m1 = mats[4 * b + 1].x;
    o.m2 = mats[4 * b + 2].x;
    o.m3 = mats[4 * b + 3].x;
  } else {
    o.m0 = mats[4 * b];
    o.m1 = mats[4 * b + 1];
    o.m2 = mats[4 * b + 2];
    o.m3 = mats[4 * b + 3];
  }
  return o;
}

// LDS slot of element (site p, k) of a strip: LEFT strips hold whole columns (k = column
// of the strip, S[k n + p]); right strips hold s rows of every column (k = row of the
// strip), column p at S[p s ..], rotated by p so that the bonds of a wave, which read the
// same row k of different columns, spread over the banks.
template <bool LEFT>
__device__ __forceinline__ int slot(int p, int k, int ld) {
  return LEFT ? k * ld + p : p * ld + ((k + p) & (ld - 1));
}

// One bond on the strip: LEFT acts on rows p.x, p.y of each column k, !LEFT on columns
// p.x, p.y of each row k.
template <bool LEFT, bool REAL>
__device__ __forceinline__ void apply_bond(cplx* S, int ld, const Bond<REAL>& o, int k0, int kend, int kstep) {
  for (int k = k0; k < kend; k += kstep) {
    const int a1 = slot<LEFT>(o.p.x, k, ld);
    const int a2 = slot<LEFT>(o.p.y, k, ld);
    const cplx x1 = S[a1], x2 = S[a2];
    if (LEFT) {       // (m11 x1 + m12 x2, m21 x1 + m22 x2)
      S[a1] = cadd(mulm(o.m0, x1), mulm(o.m2, x2));
      S[a2] = cadd(mulm(o.m1, x1), mulm(o.m3, x2));
    } else {          // (x1 m11 + x2 m21, x1 m12 + x2 m22)
      S[a1] = cadd(mulm(o.m0, x1), mulm(o.m1, x2));
      S[a2] = cadd(mulm(o.m2, x1), mulm(o.m3, x2));
    }
  }
}

template <bool LEFT, bool REAL>
__device__ __forceinline__ void apply_family(cplx* S, int ld, int ext, const int2* sites, const cplx* mats,
                                             const CbFams& fam, int f, int tid, const Bond<REAL>* pre) {
  int b, g, groups;
  if (my_bond(fam, f, tid, b, g, groups)) {
    const Bond<REAL> o = pre ? *pre : load_bond<REAL>(sites, mats, b);
    apply_bond<LEFT, REAL>(S, ld, o, g, ext, groups);
  } else if (fam.off[f + 1] - fam.off[f] > CB_NT) {
    for (int bb = fam.off[f] + tid; bb < fam.off[f + 1]; bb += CB_NT)
      apply_bond<LEFT, REAL>(S, ld, load_bond<REAL>(sites, mats, bb), 0, ext, 1);
  }
}

// The families in application order; the first CB_PRE have their bond data in registers,
// loaded once per workgroup.
template <bool LEFT, bool REAL>
__device__ __forceinline__ void apply_all(cplx* S, int ld, int ext, const int2* sites, const cplx* mats,
                                          const CbFams& fam, int reverse, int tid,
                                          const Bond<REAL> (&pre)[CB_PRE]) {
#pragma unroll
  for (int q = 0; q < CB_PRE; ++q) {
    if (q < fam.nfam) {
      const int f = reverse ? fam.nfam - 1 - q : q;
      apply_family<LEFT, REAL>(S, ld, ext, sites, mats, fam, f, tid, &pre[q]);
      __syncthreads();
    }
  }
  for (int q = CB_PRE; q < fam.nfam; ++q) {
    const int f = reverse ? fam.nfam - 1 - q : q;
    apply_family<LEFT, REAL>(S, ld, ext, sites, mats, fam, f, tid, (const Bond<REAL>*)nullptr);
    __syncthreads();
  }
}

template <bool REAL>
__device__ __forceinline__ void preload(Bond<REAL> (&pre)[CB_PRE], const int2* sites, const cplx* mats,
                                        const CbFams& fam, int reverse, int tid) {
#pragma unroll
  for (int q = 0; q < CB_PRE; ++q) {
    int b = 0, g, groups;
    if (q < fam.nfam) {
      const int f = reverse ? fam.nfam - 1 - q : q;
      if (!my_bond(fam, f, tid, b, g, groups)) b = -1;
    } else {
      b = -1;
    }
    if (b >= 0) pre[q] = load_bond<REAL>(sites, mats, b);
  }
}

// Element e of strip st: its global offset in the n x n column-major matrix, its LDS
// slot, and its row and column.  LEFT: e runs down the strip's columns; right: down the
// strip's ext rows of every column.
template <bool LEFT>
__device__ __forceinline__ void locate(int e, int o0, int ext, int n, int ld, long long& g, int& a, int& row,
                                       int& col, bool need_rc) {
  if (LEFT) {
    g = (long long)o0 * n + e;
    a = e;
    if (need_rc) {
      const int c = e / n;
      row = e - c * n;
      col = o0 + c;
    }
  } else {
    col = (ext == ld) ? (e >> (31 - __clz(ld))) : e / ext;     // ld = s is a power of two
    const int r = e - col * ext;
    row = o0 + r;
    g = (long long)col * n + row;
    a = slot<false>(col, r, ld);
  }
}

// out = diag(qr) (diag(pr) in diag(pc)) x Prod x diag(qc) for the right product, or
// diag(qr) Prod diag(pr) in diag(pc) ... for the left product (only the row scales are
// used there).  Each workgroup handles spw consecutive strips of one chain and prefetches
// strip i + 1 into registers while strip i is in LDS; the scale vectors sit in LDS after
// the strip.
template <bool LEFT, bool REAL>
__global__ __launch_bounds__(CB_NT) void cb_kernel(cplx* out, const cplx* in, int n, int s, int spw,
                                                   const cplx* pr, const cplx* pc, const cplx* qr, const cplx* qc,
                                                   const int2* sites, const cplx* mats, long long mstride, CbFams fam,
                                                   int reverse, long long cs) {
  extern __shared__ cplx S[];
  const int tid = threadIdx.x;
  out = chain_ptr(out, cs);
  in = chain_ptr(in, cs);
  pr = chain_ptr(pr, cs);
  pc = chain_ptr(pc, cs);
  qr = chain_ptr(qr, cs);
  qc = chain_ptr(qc, cs);
  mats = (const cplx*)((const char*)mats + (long long)blockIdx.z * mstride);
  const int ld = LEFT ? n : s;
  cplx* V = S + (long long)s * n;
  cplx* Vpr = nullptr;
  cplx* Vpc = nullptr;
  cplx* Vqr = nullptr;
  cplx* Vqc = nullptr;
  if (pr) { Vpr = V; V += n; }
  if (pc) { Vpc = V; V += n; }
  if (qr) { Vqr = V; V += n; }
  if (qc) { Vqc = V; V += n; }
  for (int i = tid; i < n; i += CB_NT) {
    if (pr) Vpr[i] = pr[i];
    if (pc) Vpc[i] = pc[i];
    if (qr) Vqr[i] = qr[i];
    if (qc) Vqc[i] = qc[i];
  }
  const bool pre_scaled = pr || pc;
  const bool post_scaled = qr || qc;
  Bond<REAL> pre[CB_PRE];
  preload<REAL>(pre, sites, mats, fam, reverse, tid);
  const int nstrips = (n + s - 1) / s;
  const int first = blockIdx.x * spw;
  const int last = min(nstrips, first + spw);
  cplx v[CB_U];
  {
    const int o0 = first * s, ext = min(s, n - o0), tot = ext * n;
#pragma unroll
    for (int u = 0; u < CB_U; ++u) {
      const int e = u * CB_NT + tid;
      if (e < tot) {
        long long g; int a, row, col;
        locate<LEFT>(e, o0, ext, n, ld, g, a, row, col, false);
        v[u] = in[g];
      }
    }
  }
  __syncthreads();
  for (int st = first; st < last; ++st) {
    const int o0 = st * s, ext = min(s, n - o0), tot = ext * n;
#pragma unroll
    for (int u = 0; u < CB_U; ++u) {
      const int e = u * CB_NT + tid;
      if (e < tot) {
        long long g; int a, row, col;
        locate<LEFT>(e, o0, ext, n, ld, g, a, row, col, pre_scaled);
        cplx x = v[u];
        if (Vpr) x = cmul(x, Vpr[row]);
        if (Vpc) x = cmul(x, Vpc[col]);
        S[a] = x;
      }
    }
    __syncthreads();
    if (st + 1 < last) {
      const int o1 = (st + 1) * s, ext1 = min(s, n - o1), tot1 = ext1 * n;
#pragma unroll
      for (int u = 0; u < CB_U; ++u) {
        const int e = u * CB_NT + tid;
        if (e < tot1) {
          long long g; int a, row, col;
          locate<LEFT>(e, o1, ext1, n, ld, g, a, row, col, false);
          v[u] = in[g];
        }
      }
    }
    apply_all<LEFT, REAL>(S, ld, ext, sites, mats, fam, reverse, tid, pre);
#pragma unroll
    for (int u = 0; u < CB_U; ++u) {
      const int e = u * CB_NT + tid;
      if (e < tot) {
        long long g; int a, row, col;
        locate<LEFT>(e, o0, ext, n, ld, g, a, row, col, post_scaled);
        cplx x = S[a];
        if (Vqr) x = cmul(x, Vqr[row]);
        if (Vqc) x = cmul(x, Vqc[col]);
        out[g] = x;
      }
    }
    __syncthreads();
  }
}

// Strip width s (columns for the left product, rows for the right one): a power of two,
// at most 16, with s n <= CB_NT CB_U elements (one register load per thread and element),
// narrowed while the chains' strips number fewer than 1024; spw strips per workgroup such
// that about 2048 workgroups cover the batch.
void plan(int n, int nchain, int& s, int& spw) {
  s = 16;
  while (s > 1 && (long long)s * n > (long long)CB_NT * CB_U) s >>= 1;
  while (s > 1 && (long long)((n + s - 1) / s) * nchain < 1024) s >>= 1;
  const long long strips = (long long)((n + s - 1) / s) * nchain;
  spw = (int)std::max(1LL, std::min(8LL, strips / 2048));
}

int check_ops(int n, const CbOps& ops, int nchain) {
  if (n < 1 || n > 1024) {
    alfq_set_error("checkerboard: Ndim outside 1..1024");
    return ST_GENERIC;
  }
  if (!ops.sites || !ops.mats || ops.fam.nfam < 1 || ops.fam.nfam > CB_MAXFAM || nchain < 1) {
    alfq_set_error("checkerboard: operators not set");
    return ST_GENERIC;
  }
  return ST_NONE;
}

template <bool LEFT>
int cb_launch(cplx* out, const cplx* in, int n, const cplx* pr, const cplx* pc, const cplx* qr, const cplx* qc,
              const CbOps& ops, int reverse, int nchain, long long cs, hipStream_t st) {
  int e;
  if ((e = check_ops(n, ops, nchain))) return e;
  int s, spw;
  plan(n, nchain, s, spw);
  const int nstrips = (n + s - 1) / s;
  const int nvec = (pr != nullptr) + (pc != nullptr) + (qr != nullptr) + (qc != nullptr);
  const size_t lds = ((size_t)s * n + (size_t)nvec * n) * sizeof(cplx);
  if (lds > (size_t)CB_LDS_MAX) {
    alfq_set_error("checkerboard: strip and scales exceed 64 KB of LDS");
    return ST_GENERIC;
  }
  const dim3 grid((nstrips + spw - 1) / spw, 1, nchain);
  if (ops.is_real)
    hipLaunchKernelGGL((cb_kernel<LEFT, true>), grid, dim3(CB_NT), lds, st, out, in, n, s, spw, pr, pc, qr, qc,
                       ops.sites, ops.mats, ops.mstride, ops.fam, reverse, cs);
  else
    hipLaunchKernelGGL((cb_kernel<LEFT, false>), grid, dim3(CB_NT), lds, st, out, in, n, s, spw, pr, pc, qr, qc,
                       ops.sites, ops.mats, ops.mstride, ops.fam, reverse, cs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

}  // namespace

int cb_left(cplx* out, const cplx* in, int n, const cplx* pre_rs, const cplx* post_rs, const CbOps& ops,
            int reverse, int nchain, long long cs, hipStream_t s) {
  return cb_launch<true>(out, in, n, pre_rs, nullptr, post_rs, nullptr, ops, reverse, nchain, cs, s);
}

int cb_right(cplx* out, const cplx* in, int n, const cplx* pre_rs, const cplx* pre_cs, const cplx* post_rs,
             const cplx* post_cs, const CbOps& ops, int reverse, int nchain, long long cs, hipStream_t s) {
  return cb_launch<false>(out, in, n, pre_rs, pre_cs, post_rs, post_cs, ops, reverse, nchain, cs, s);
}
