// Checkerboard hopping exponential on the device: the product of 2 x 2 bond
// exponentials (Prog/Hop_mod.F90:143-250 with OpTTypes_mod.F90:150-304) applied to an
// n x n column-major matrix without forming the dense product.
#pragma once

#include "alfq_common.h"

#define CB_MAXFAM 16

// Bonds b = off[f] .. off[f + 1] - 1 form family f; a family's bonds sit on disjoint sites,
// so they commute and are applied together.
struct CbFams {
  int nfam;
  int off[CB_MAXFAM + 1];
};

// Shared by the chains: the two sites of every bond (0-based).  Per chain: the bond
// matrices, 4 complex per bond in column-major order (m11, m21, m12, m22), chain z's at
// (chain 0's pointer) + z * mstride bytes.
struct CbOps {
  const int2* sites;
  const cplx* mats;
  long long mstride;
  CbFams fam;
  int nbond;
  bool is_real;             // every matrix of every chain has zero imaginary parts
};

// out = diag(post_rs) * Prod * diag(pre_rs) * in, where Prod applies the families in
// ascending (reverse = 0) or descending (reverse = 1) order, each bond from the left
// (ZSLHEMM('L', ...) per bond).  Optional scales may be null.  in == out is allowed.
int cb_left(cplx* out, const cplx* in, int n, const cplx* pre_rs, const cplx* post_rs, const CbOps& ops,
            int reverse, int nchain, long long cs, hipStream_t s);

// out = diag(post_rs) * (diag(pre_rs) * in * diag(pre_cs)) * Prod * diag(post_cs), each bond
// applied from the right (ZSLHEMM('R', ...)).  in == out is allowed.
int cb_right(cplx* out, const cplx* in, int n, const cplx* pre_rs, const cplx* pre_cs, const cplx* post_rs,
             const cplx* post_cs, const CbOps& ops, int reverse, int nchain, long long cs, hipStream_t s);
