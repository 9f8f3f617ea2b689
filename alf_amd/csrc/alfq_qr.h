// Stabilisation building blocks (qr.hip).
#pragma once
#include "alfq_common.h"

struct QRWork {
  int n = 0, nb = 0, nblk = 0;
  int nchain = 1;         // chains of the batch; every buffer below is per chain
  long long cs = 0;       // byte stride between the chains' arenas
  int policy = 0;         // GEMM tiling policy of the context
  bool lds_panel = false; // panel factorisation in LDS (n <= 2048)
  cplx* A = nullptr;      // n x n factorisation workspace
  cplx* Vw = nullptr;     // explicit Householder panels (n x n)
  cplx* Y1 = nullptr;     // V T^H per panel
  cplx* Y2 = nullptr;     // V T per panel
  cplx* Rt = nullptr;     // normalised R (upper)
  cplx* Rinv = nullptr;   // Rt^-1
  cplx* Tmp = nullptr;    // n x n scratch
  cplx* T = nullptr;      // nblk x (nb x nb)
  cplx* W = nullptr;      // nb x n
  cplx* tau = nullptr;
  cplx* pdet = nullptr;   // det of each panel's reflector product
  double* norms = nullptr;
  int* perm = nullptr;    // column i of A P is column perm[i] of A
  int* pinv = nullptr;
  cplx* scal = nullptr;   // 8 n-vectors of scale factors
};

// places the workspace of one chain in its arena (sizes it when ar.base is null)
void qr_work_place(QRWork& w, int n, ArenaLayout& ar);
int launch_fill(cplx* X, long long count, cplx val, int nchain, long long cs, hipStream_t s);
int launch_identity(cplx* X, int n, int nchain, long long cs, hipStream_t s);

// decompose_UDV_state: (U, D, V) of the propagated state -> new (U, D, V) with
// U D V = Usrc Dsrc Vsrc (side R) or U D V^H = Usrc Dsrc Vsrc^H (side L).
int udv_decompose_dev(QRWork& w, const cplx* Usrc, const cplx* Dsrc, const cplx* Vsrc, int side_l,
                      cplx* Uout, cplx* Dout, cplx* Vout, cplx* phase_out, cplx* detU_out, hipStream_t s);

// QR of TPUP (nvar = 1) or TPUP^H (nvar = 2): Q into Qout, D into DUP, normalised R into
// w.Rt, permutation in w.perm/w.pinv, CGR determinant phase into z_out.
int qr_factor_cgr(QRWork& w, const cplx* TPUP, int nvar, const cplx* detUr, const cplx* detUl, cplx* Qout,
                  cplx* DUP, cplx* z_out, hipStream_t s);

// Rinv = Rt^-1 for the upper-triangular Rt
int tri_inverse_dev(QRWork& w, const cplx* Rt, cplx* Rinv, hipStream_t s);
