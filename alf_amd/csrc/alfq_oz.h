// Error-free split products on the int8 matrix cores (ozaki.hip).
//
// A double matrix is written as 2^e_i (sum_{s=1..S} 2^(-7 s) D_s) with one exponent per
// row (or column) and S int8 digit matrices D_s in [-127, 127] (base-128 digits of the row
// scaled into (-1, 1), each rounded to nearest and saturated at +-127); the remainder is
// below 2^(-7 S) of the row's largest entry.
// A product is the sum over digit pairs with s + t <= S + 1: each pair is an exact integer
// product (v_mfma_i32_32x32x32_i8, exact int32 sums), the pairs of one s + t are added in
// int32 and the S sums are combined in FP64 with their power-of-two weights.
#pragma once
#include "alfq_common.h"

constexpr int OZ_S = 8;              // digits per operand: 56 bits of each row's range

// Digit planes of one operand: planes x rows x kp x OZ_S bytes and one exponent per row.  Row
// r of plane p starts at (p rows + r) kp OZ_S and holds its k-blocks of 32 entries in order,
// each block as OZ_S segments of 32 bytes (digit s of entries 32 b .. 32 b + 31 at
// (32 b) OZ_S + 32 s); zero beyond the contraction length.  A product stage of 32 entries of
// one row is then OZ_S x 32 contiguous bytes.  Plane 0 is the real part, 1 the imaginary
// part, 2 its negation.
struct OzOperand {
  int8_t* d = nullptr;
  int* e = nullptr;
  int rows = 0, planes = 1;
  long long kp = 0;                  // row pitch: contraction length rounded up to 64
  long long stride = 0;              // bytes between the chains of a batch (d and e)
};

inline long long oz_pitch(int k) { return ((long long)k + 63) / 64 * 64; }
inline size_t oz_bytes(int rows, int k, int planes) { return (size_t)planes * OZ_S * rows * oz_pitch(k); }

// Digits of X for the contraction index k.  by_row = 0: X is K x N (column-major, ld), one
// exponent per column j (the right operand of a product); by_row = 1: X is M x K, one
// exponent per row i (the left operand).  X is complex (planes 2, or 3 with the negated
// imaginary part) or real (planes 1).  kscale (complex, per k, may be null) multiplies X's
// contraction index before the split.  xs: chain stride of X and kscale.
int oz_split(const void* X, bool x_cplx, int ld, int nidx, int nk, bool by_row, const cplx* kscale, long long xs,
             const OzOperand& out, int nchain, hipStream_t s);

// C = alpha (A B) + beta C with row scale rs and column scale cs of the product (complex,
// may be null), from the digit planes of A (by row, M x K) and B (by column, K x N): real x
// complex (A 1 plane, B 2), complex x real (2, 1), complex x complex (3, 2 or 2, 3).
// C, rs and cs advance by cstride per chain, the operands by their own strides.
int oz_gemm(int M, int N, int K, const OzOperand& A, const OzOperand& B, cplx* C, int ldc, cplx alpha, cplx beta,
            const cplx* rs, const cplx* cs, long long cstride, int nchain, hipStream_t s);
