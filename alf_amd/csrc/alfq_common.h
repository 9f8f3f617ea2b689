// Common types and helpers for the gfx950 DQMC sweep engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>

typedef double2 cplx;   // interleaved (re, im), the layout of Fortran complex(8)

__host__ __device__ inline cplx cmake(double r, double i) { return make_double2(r, i); }
__host__ __device__ inline cplx cadd(cplx a, cplx b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline cplx csub(cplx a, cplx b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline cplx cmul(cplx a, cplx b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline cplx cconj(cplx a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline cplx cscale(cplx a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double cabs2(cplx a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ inline double cabs_(cplx a) { return hypot(a.x, a.y); }
__host__ __device__ inline cplx cdiv(cplx a, cplx b) {
  // Smith's algorithm keeps the intermediate products in range
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, d = b.x + r * b.y;
    return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  } else {
    double r = b.x / b.y, d = b.y + r * b.x;
    return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
  }
}
__host__ __device__ inline cplx cinv(cplx b) { return cdiv(make_double2(1.0, 0.0), b); }
__device__ inline cplx cexp_(cplx z) {
  double e = exp(z.x);
  double s, c;
  sincos(z.y, &s, &c);
  return make_double2(e * c, e * s);
}
__host__ __device__ inline cplx cfma(cplx a, cplx b, cplx c) {  // a*b + c
  return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}

// Batches of chains: every per-chain device buffer of chain z lives at the same offset of
// chain z's arena, so a kernel finds its chain's data as (chain 0's pointer) + z * stride
// bytes, with z = blockIdx.z.  Null pointers (absent optional operands) stay null.
template <class T>
__device__ __forceinline__ T* chain_ptr(T* p, long long stride) {
  return p ? (T*)((char*)p + (long long)blockIdx.z * stride) : p;
}
template <class T>
__device__ __forceinline__ T* chain_ptr_at(T* p, long long stride, int z) {
  return p ? (T*)((char*)p + (long long)z * stride) : p;
}

// XCD-aware order of a batched grid (tiles in x, y; chains or problems in z).  The device
// hands workgroups to its 8 XCDs in turn by linear id, so in grid order the tiles of one
// chain spread over all 8 L2s, each fetching the chain's operands.  When gridDim.z is a
// multiple of 8 the linear ids are renumbered (a bijection) so that every tile of one z lands
// on the same XCD: id = 8 i + x runs z = x + 8 (i / tiles), tile i % tiles.  Other grids keep
// their order.
__device__ __forceinline__ void xcd_order(int& bx, int& by, int& bz) {
  bx = blockIdx.x;
  by = blockIdx.y;
  bz = blockIdx.z;
  if ((gridDim.z & 7) != 0) return;
  const long long P = (long long)gridDim.x * gridDim.y;
  const long long L = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
  const long long i = L >> 3, t = i % P;
  bz = (int)((L & 7) + 8 * (i / P));
  bx = (int)(t % gridDim.x);
  by = (int)(t / gridDim.x);
}

// Two-pass layout of a chain's arena: the first pass (base == nullptr) only sizes it, the
// second assigns pointers at the same 256-byte aligned offsets.  With cap > 0 the layout
// spans regions of cap bytes, region r at base + r * rstride: a buffer that does not fit
// the rest of a region starts the next one (so buffers in separate allocations can keep
// the chain stride of the main arena).
struct ArenaLayout {
  char* base = nullptr;
  size_t off = 0;
  size_t cap = 0;
  long long rstride = 0;
  int region = 0;
  bool overflow = false;     // a buffer larger than cap was requested
  template <class T>
  T* take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    const size_t bytes = count * sizeof(T);
    if (cap > 0 && off + bytes > cap && off > 0) {
      ++region;
      off = 0;
    }
    if (cap > 0 && bytes > cap) overflow = true;
    T* p = base ? (T*)(base + (long long)region * rstride + off) : nullptr;
    off += bytes;
    return p;
  }
};

// Status codes: the values of the reference's runtime_error_mod enum
// (Libraries/Modules/runtime_error_mod.F90:56-68), returned by every C entry point.
enum alfq_status_internal {
  ST_NONE = 0,
  ST_GENERIC = 1,
  ST_FILE_NOT_FOUND = 2,
  ST_UNSTABLE_MATRIX = 3,
  ST_MISSING_OBS = 4,
  ST_FIELDS = 5,
  ST_HAMILTONIAN = 6,
};

void alfq_set_error(const std::string& msg);

// Synchronous copies, sets, allocations and frees order themselves against the other
// streams' work, which invalidates a HIP-graph capture running in another host thread.
// Entry points that make such calls and the sweep captures hold this process-wide lock
// (recursive: a failing create releases its context through alfq_ctx_destroy).
std::recursive_mutex& null_stream_mutex();
#define NULL_STREAM_LOCK std::lock_guard<std::recursive_mutex> null_stream_lock_(null_stream_mutex())

#define HIP_OK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      alfq_set_error(std::string(#expr) + ": " + hipGetErrorString(e_));          \
      return ST_GENERIC;                                                          \
    }                                                                             \
  } while (0)

// ---- GEMM (gemm.hip) ---------------------------------------------------------
// C(M x N) = alpha * opA(A) * opB(B) + beta * C, all column-major FP64.
// A/B are real (double) or complex (double2).  Optional per-row scale of the result
// (rs, complex), per-column scale (cs, complex), per-row scale of opB (bs, complex),
// row gather of opB (bmap: opB(k, j) = B'(bmap[k], j)) and column gather of opA
// (amap: opA(i, k) = A'(i, amap[k])).  K may be read from device memory (kdev).
enum GemmOp { OP_N = 0, OP_T = 1, OP_C = 2 };

struct GemmArgs {
  int M, N, K;
  const int* kdev;          // if non-null, K = *kdev (clamped to K)
  const void* A; int lda; int opA; bool a_cplx;
  const void* B; int ldb; int opB; bool b_cplx;
  void* C; int ldc;
  cplx alpha, beta;
  const cplx* rs;           // rows of the product
  const cplx* cs;           // columns of the product
  const cplx* bs;           // rows of opB
  const double* rs_real;    // real per-row scale (alternative to rs)
  const int* amap;          // column gather of opA
  const int* bmap;          // row gather of opB
  int batch;                // number of independent problems per chain
  long long strideA, strideB, strideC;  // element strides between problems
  int nchain;               // chains of the batch (grid z = batch * nchain)
  long long cstride;        // byte stride between the chains' arenas (applies to every pointer)
  int prec;                 // 0: FP64 MFMA; 1: operands rounded to FP32, FP32 MFMA and accumulation
  int policy;               // tiling policy (GEMM_POLICIES): the grid sizes at which the launcher
                            // switches to the direct-load and to the 64 x 64 kernels
  // batches of uneven Ltrot: for a chain z with lz[z] <= slice, operand hop_side (1: A, 2: B),
  // the hopping exponential, is replaced by ident (an identity of the same type, shared): the
  // slice is the identity for that chain
  const int* lz;
  int slice;
  int hop_side;
  const void* ident;
  // direct-load kernel only (gemm_direct_path): with amap set and a complex opA = A (OP_N),
  // also write yhat(r, j) = delta(r, amap[j]) - A(amap[j], r) for r < M, j < K (ld M), the
  // rows amap of A read before any later kernel changes A
  cplx* yhat;
};

GemmArgs gemm_args(int M, int N, int K, const void* A, int lda, int opA, bool a_cplx,
                   const void* B, int ldb, int opB, bool b_cplx, void* C, int ldc,
                   cplx alpha, cplx beta);
int gemm_launch(const GemmArgs& g, hipStream_t s);
// Tiling policies a context may pick (its autotuner measures them): 0 the estimate-based
// default, 1 the 64 x 64 kernel from one tile per CU, 2 never the 64 x 64 kernel, 3 the
// direct-load kernel up to 512 tiles of 32 x 32.
constexpr int GEMM_POLICIES = 4;
// true when gemm_launch runs g on the direct-load kernel (the only one that honours yhat)
bool gemm_direct_path(const GemmArgs& g);
// G += X Y^T (all n x n / n x m, column-major, ld n) with the rank m read from *mdev and
// clamped to kmax <= 32, for the nchain chains of a batch (pointers advance by cs bytes).
int flush_launch(cplx* G, const cplx* X, const cplx* Y, const int* mdev, int n, int kmax, int nchain,
                 long long cs, hipStream_t s);
// the round's X product and flush in one launch for a few chains: Go = Gi + (Gi(:, A) W) Yhat^T
// with Yhat(c, j) = delta(c, amap[j]) - Gi(amap[j], c), rank m = *mdev <= kmax <= 32; W is
// m x m with leading dimension 32; Gi and Go are different buffers.
int round_flush_launch(const cplx* Gi, cplx* Go, const cplx* W, const int* amap, const int* mdev, int n, int kmax,
                       int nchain, long long cs, hipStream_t s);
