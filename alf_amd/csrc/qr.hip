// Stabilisation kernels: norm-presorted blocked Householder QR (the UDV
// decomposition and the QR of CGR's middle matrix), Q formation, the
// scale/phase bookkeeping of QDRP_decompose, and a blocked upper-triangular
// inverse for CGR's solve.
//
// Pivoting: columns are ordered once by descending norm and the QR itself is
// unpivoted -- the pivoting strategy of the reference's STABLOG decompose
// (Prog/udv_state_mod.F90:503-540) and of UDV_Wrap_Pivot
// (Prog/UDV_WRAP_mod.F90:125-208), applied in linear scale.  The Green's
// function and its phase do not depend on the pivot sequence.
//
// Blocked QR: a single-workgroup panel kernel factors nb columns (zgeqr2 /
// zlarfg conventions: v(1) = 1, tau = (beta - alpha)/beta, H = I - tau v v^H)
// and forms the compact-WY factor T (zlarft, forward columnwise); the trailing
// matrix and the explicit Q are updated with the MFMA GEMM.
#include <algorithm>
#include <mutex>

#include "alfq_qr.h"

namespace {

constexpr int PT = 256;   // threads of the single-workgroup kernels
constexpr int NBMAX = 32; // largest panel width

// A double moved between the lanes of a 16-lane row by the DPP control CTRL (both halves).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over each 16-lane row, in every lane of the row: quad_perm [1,0,3,2], quad_perm
// [2,3,0,1], row_half_mirror, row_mirror (register moves, no LDS round trip).  Every lane
// adds the same two values, so all lanes hold the bit-identical sum.
__device__ __forceinline__ double row16_sum(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  v += dpp_f64<0x140>(v);
  return v;
}

// gfx950 lane swaps (register moves, no LDS): with both operands x,
// permlane16_swap gives [0] = rows (0, 0, 2, 2) and [1] = rows (1, 1, 3, 3) of x,
// permlane32_swap gives [0] = (low half, low half) and [1] = (high half, high half).
__device__ __forceinline__ void swap16_f64(double v, double& even, double& odd) {
  const long long b = __double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)(b & 0xffffffffLL), (unsigned)(b & 0xffffffffLL), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
  even = __longlong_as_double(((long long)hi[0] << 32) | (long long)lo[0]);
  odd = __longlong_as_double(((long long)hi[1] << 32) | (long long)lo[1]);
}
__device__ __forceinline__ void swap32_f64(double v, double& low, double& high) {
  const long long b = __double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)(b & 0xffffffffLL), (unsigned)(b & 0xffffffffLL), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
  low = __longlong_as_double(((long long)hi[0] << 32) | (long long)lo[0]);
  high = __longlong_as_double(((long long)hi[1] << 32) | (long long)lo[1]);
}

// Sum over each 32-lane half of the wave, in every lane of the half (bit-identical).
__device__ __forceinline__ double half_sum(double v) {
  v = row16_sum(v);
  double e, o;
  swap16_f64(v, e, o);
  return e + o;
}

__device__ inline double wave_sum(double v) {
  v = half_sum(v);
  double lo, hi;
  swap32_f64(v, lo, hi);
  return lo + hi;
}

// block-wide sum of a double; all threads get the result.  red: >= PT/64 doubles of LDS.
__device__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
  return s;
}

// ---- column norms of the matrix to be factorised -------------------------------
// src(i, j) = S(i, j) * D(j)   or, with ctrans, conj(S(j, i)).
__global__ void colnorm_kernel(const cplx* __restrict__ S, int n, int lds, const cplx* __restrict__ D,
                               int ctrans, double* __restrict__ norms, long long cs) {
  S = chain_ptr(S, cs);
  D = chain_ptr(D, cs);
  norms = chain_ptr(norms, cs);
  const int j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= n) return;
  double s = 0.0;
  for (int i = lane; i < n; i += 64) {
    cplx v = ctrans ? S[j + (long long)i * lds] : S[i + (long long)j * lds];
    s += cabs2(v);
  }
  s = wave_sum(s);
  if (lane == 0) {
    double nr = sqrt(s);
    if (D) nr *= cabs_(D[j]);
    norms[j] = nr;
  }
}

// rank sort, descending, ties broken by index: perm[rank] = j, pinv[j] = rank
__global__ void sortperm_kernel(const double* __restrict__ norms, int n, int* __restrict__ perm,
                                int* __restrict__ pinv, long long cs) {
  norms = chain_ptr(norms, cs);
  perm = chain_ptr(perm, cs);
  pinv = chain_ptr(pinv, cs);
  // four lanes per column j, each counting every fourth i; the lanes' counts are summed with
  // two exchanges inside the wave, so every lane runs the same number of rounds
  const int q = threadIdx.x & 3, g = threadIdx.x >> 2, G = blockDim.x >> 2;
  const int rounds = (n + G - 1) / G;
  for (int it = 0; it < rounds; ++it) {
    const int j0 = it * G + g;
    const int j = j0 < n ? j0 : n - 1;
    const double x = norms[j];
    int r = 0;
    for (int i = q; i < n; i += 4) {
      const double y = norms[i];
      r += (y > x) || (y == x && i < j);
    }
    r += __shfl_xor(r, 1, 64);
    r += __shfl_xor(r, 2, 64);
    if (q == 0 && j0 < n) {
      perm[r] = j;
      pinv[j] = r;
    }
  }
}

// A(:, i) = src(:, perm[i])
__global__ void gather_kernel(const cplx* __restrict__ S, int n, int lds, const cplx* __restrict__ D,
                              int ctrans, const int* __restrict__ perm, cplx* __restrict__ A, long long cs) {
  S = chain_ptr(S, cs);
  D = chain_ptr(D, cs);
  perm = chain_ptr(perm, cs);
  A = chain_ptr(A, cs);
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int r = (int)(idx % n), i = (int)(idx / n);
  const int j = perm[i];
  cplx v;
  if (ctrans) {
    v = cconj(S[j + (long long)r * lds]);
  } else {
    v = S[r + (long long)j * lds];
    if (D) v = cmul(v, D[j]);
  }
  A[idx] = v;
}

__global__ void fill_kernel(cplx* __restrict__ X, long long count, cplx val, long long cs) {
  X = chain_ptr(X, cs);
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < count) X[idx] = val;
}

__global__ void identity_kernel(cplx* __restrict__ X, int n, long long cs) {
  X = chain_ptr(X, cs);
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  X[idx] = make_double2(r == c ? 1.0 : 0.0, 0.0);
}

// ---- panel factorisation (one workgroup) ----------------------------------------
// Factors A(k:n-1, k:k+nb-1) in place.  Writes tau(k..), the explicit V panel into
// Vw(k:n-1, k:k+nb-1) (unit diagonal, zeros above), T (nb x nb, upper), Y1 = V T^H and,
// if Y2w is given, Y2 = V T into the same positions of Y1w/Y2w, and
// det(H_k...H_{k+nb-1}) into pdet[b].
__global__ __launch_bounds__(PT) void panel_kernel(cplx* __restrict__ A, int n, int k, int nb,
                                                   cplx* __restrict__ tau, cplx* __restrict__ Vw,
                                                   cplx* __restrict__ Y1w, cplx* __restrict__ Y2w,
                                                   cplx* __restrict__ Tw, cplx* __restrict__ pdet, long long cs) {
  A = chain_ptr(A, cs);
  tau = chain_ptr(tau, cs);
  Vw = chain_ptr(Vw, cs);
  Y1w = chain_ptr(Y1w, cs);
  Y2w = chain_ptr(Y2w, cs);
  Tw = chain_ptr(Tw, cs);
  pdet = chain_ptr(pdet, cs);
  __shared__ double red[PT / 64];
  __shared__ cplx s_tau[NBMAX];
  __shared__ cplx s_T[NBMAX * NBMAX];
  __shared__ cplx s_w[NBMAX];
  __shared__ double s_vn2[NBMAX];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int nw = blockDim.x >> 6;
  const int ncol = (k + nb <= n) ? nb : n - k;

  for (int j = 0; j < ncol; ++j) {
    const int c = k + j;
    cplx* col = A + (long long)c * n;
    // reflector for col(c..n-1)
    double xs = 0.0;
    for (int r = c + 1 + tid; r < n; r += blockDim.x) xs += cabs2(col[r]);
    const double xnorm2 = block_sum(xs, red);
    const cplx alpha = col[c];
    cplx t = make_double2(0.0, 0.0);
    cplx scal = make_double2(1.0, 0.0);
    double beta = alpha.x;
    if (!(xnorm2 == 0.0 && alpha.y == 0.0)) {
      beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2), alpha.x);
      t = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
      scal = cinv(make_double2(alpha.x - beta, alpha.y));
    }
    __syncthreads();
    for (int r = c + 1 + tid; r < n; r += blockDim.x) col[r] = cmul(col[r], scal);
    if (tid == 0) {
      col[c] = make_double2(beta, 0.0);
      s_tau[j] = t;
      tau[c] = t;
    }
    __syncthreads();
    // apply H^H = I - conj(tau) v v^H to the remaining panel columns, one wave per column
    const cplx tc = cconj(t);
    for (int q = c + 1 + wave; q < k + ncol; q += nw) {
      cplx* aq = A + (long long)q * n;
      double wr = 0.0, wi = 0.0;
      for (int r = c + lane; r < n; r += 64) {
        const cplx v = (r == c) ? make_double2(1.0, 0.0) : col[r];
        const cplx a = aq[r];
        wr += v.x * a.x + v.y * a.y;     // conj(v) * a
        wi += v.x * a.y - v.y * a.x;
      }
      wr = wave_sum(wr);
      wi = wave_sum(wi);
      const cplx f = cmul(tc, make_double2(wr, wi));
      for (int r = c + lane; r < n; r += 64) {
        const cplx v = (r == c) ? make_double2(1.0, 0.0) : col[r];
        aq[r] = csub(aq[r], cmul(v, f));
      }
    }
    __syncthreads();
  }

  // explicit V panel and |v_j|^2
  for (int j = 0; j < ncol; ++j) {
    const int c = k + j;
    double s = 0.0;
    for (int r = k + tid; r < n; r += blockDim.x) {
      cplx v;
      if (r < c) v = make_double2(0.0, 0.0);
      else if (r == c) v = make_double2(1.0, 0.0);
      else v = A[r + (long long)c * n];
      Vw[r + (long long)c * n] = v;
      s += cabs2(v);
    }
    s = block_sum(s, red);
    if (tid == 0) s_vn2[j] = s;
  }
  __syncthreads();
  // T (zlarft forward columnwise): T(j,j) = tau_j, T(0:j, j) = -tau_j T(0:j,0:j) V(:,0:j)^H v_j
  for (int idx = tid; idx < NBMAX * NBMAX; idx += blockDim.x) s_T[idx] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int j = 0; j < ncol; ++j) {
    const int c = k + j;
    // s_w[i] = V(:, i)^H v_j for i < j, one wave per i
    for (int i = wave; i < j; i += nw) {
      const int ci = k + i;
      double wr = 0.0, wi = 0.0;
      for (int r = c + lane; r < n; r += 64) {   // v_j is zero above row c
        const cplx a = Vw[r + (long long)ci * n];
        const cplx b = Vw[r + (long long)c * n];
        wr += a.x * b.x + a.y * b.y;
        wi += a.x * b.y - a.y * b.x;
      }
      wr = wave_sum(wr);
      wi = wave_sum(wi);
      if (lane == 0) s_w[i] = make_double2(wr, wi);
    }
    __syncthreads();
    if (tid < j) {
      // T(tid, j) = -tau_j * sum_{l=tid..j-1} T(tid, l) w_l   (T(0:j,0:j) is upper)
      cplx acc = make_double2(0.0, 0.0);
      for (int l = tid; l < j; ++l) acc = cfma(s_T[tid + NBMAX * l], s_w[l], acc);
      s_T[tid + NBMAX * j] = cmul(make_double2(-s_tau[j].x, -s_tau[j].y), acc);
    }
    if (tid == 0) s_T[j + NBMAX * j] = s_tau[j];
    __syncthreads();
  }
  // store T, Y1 = V T^H, Y2 = V T  (rows k..n-1, columns k..k+ncol-1)
  const int b = k / nb;
  for (int idx = tid; idx < ncol * ncol; idx += blockDim.x) {
    const int i = idx % ncol, j = idx / ncol;
    Tw[(long long)b * nb * nb + i + j * nb] = s_T[i + NBMAX * j];
  }
  for (long long idx = tid; idx < (long long)(n - k) * ncol; idx += blockDim.x) {
    const int r = k + (int)(idx % (n - k));
    const int j = (int)(idx / (n - k));
    cplx y1 = make_double2(0.0, 0.0), y2 = make_double2(0.0, 0.0);
    for (int l = 0; l < ncol; ++l) {
      const cplx v = Vw[r + (long long)(k + l) * n];
      // Y1(r, j) = sum_l V(r, l) conj(T(j, l));  Y2(r, j) = sum_l V(r, l) T(l, j)
      y1 = cfma(v, cconj(s_T[j + NBMAX * l]), y1);
      y2 = cfma(v, s_T[l + NBMAX * j], y2);
    }
    Y1w[r + (long long)(k + j) * n] = y1;
    if (Y2w) Y2w[r + (long long)(k + j) * n] = y2;
  }
  if (tid == 0) {
    cplx d = make_double2(1.0, 0.0);
    for (int j = 0; j < ncol; ++j) {
      const cplx h = csub(make_double2(1.0, 0.0), cscale(s_tau[j], s_vn2[j]));
      d = cmul(d, h);
    }
    pdet[b] = d;
  }
}

// ---- panel factorisation with the panel resident in LDS -------------------------
// Same contract as panel_kernel.  The rows x ncol panel sits in dynamic LDS; for each
// column the reflector is formed from the norm left by the previous step, and the
// waves apply H^H to the remaining columns (wave w owns columns j+1+w, j+5+w, ...)
// with the reflector's scale folded in, so one barrier separates two columns.  The
// Gram matrix V^H V, and Y1 = V T^H, Y2 = V T, are MFMA products out of LDS.
typedef double v4d_t __attribute__((ext_vector_type(4)));

template <class FA, class FB>
__device__ __forceinline__ void cmfma16(int K, FA A, FB B, v4d_t& re, v4d_t& im) {
  const int lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += 4) {
    const cplx a = A(lr, k0 + lk);
    const cplx b = B(k0 + lk, lr);
    re = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, re, 0, 0, 0);
    re = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.y, re, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, im, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, im, 0, 0, 0);
  }
}

constexpr int PB = 32;          // widest LDS panel
constexpr int PLT = 1024;       // threads of the LDS panel kernel (16 waves)
constexpr int PLW = PLT / 64;   // its waves
constexpr int QPW = PB / PLW;   // panel columns per wave

__global__ __launch_bounds__(PLT) void panel_lds_kernel(cplx* __restrict__ A, int n, int k, int nb,
                                                       cplx* __restrict__ tau, cplx* __restrict__ Vw,
                                                       cplx* __restrict__ Y1w, cplx* __restrict__ Y2w,
                                                       cplx* __restrict__ Tw, cplx* __restrict__ pdet, long long cs) {
  A = chain_ptr(A, cs);
  tau = chain_ptr(tau, cs);
  Vw = chain_ptr(Vw, cs);
  Y1w = chain_ptr(Y1w, cs);
  Y2w = chain_ptr(Y2w, cs);
  Tw = chain_ptr(Tw, cs);
  pdet = chain_ptr(pdet, cs);
  extern __shared__ cplx sp[];                 // rows x ncol, column-major, ld = rows
  __shared__ cplx s_TG[PB * PB];               // upper incl. diagonal: T; strict lower: Gram
  __shared__ cplx s_tau[PB], s_sc[PB];
  __shared__ double s_beta[PB], s_norm[PB + 1], s_vn2[PB];
  __shared__ double red[PLW];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int rows = n - k;
  const int ncol = (k + nb <= n) ? nb : n - k;

  for (long long idx = tid; idx < (long long)rows * ncol; idx += blockDim.x) {
    const int r = (int)(idx % rows), c = (int)(idx / rows);
    sp[idx] = A[(k + r) + (long long)(k + c) * n];
  }
  __syncthreads();
  {
    double xs = 0.0;
    for (int r = 1 + tid; r < rows; r += blockDim.x) xs += cabs2(sp[r]);
    xs = block_sum(xs, red);
    if (tid == 0) s_norm[0] = xs;
  }
  __syncthreads();

  for (int j = 0; j < ncol; ++j) {
    const cplx* xj = sp + (long long)j * rows;
    const cplx alpha = xj[j];
    const double xn2 = s_norm[j];
    cplx t = make_double2(0.0, 0.0), sc = make_double2(1.0, 0.0);
    double beta = alpha.x;
    if (!(xn2 == 0.0 && alpha.y == 0.0)) {
      beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2), alpha.x);
      t = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
      sc = cinv(make_double2(alpha.x - beta, alpha.y));
    }
    if (tid == 0) {
      s_tau[j] = t;
      s_sc[j] = sc;
      s_beta[j] = beta;
      tau[k + j] = t;
    }
    // dots d_q = sum_{r>j} conj(x_r) A(r, q) for this wave's columns
    cplx d[QPW];
#pragma unroll
    for (int i = 0; i < QPW; ++i) d[i] = make_double2(0.0, 0.0);
    for (int r = j + 1 + lane; r < rows; r += 64) {
      const cplx xr = xj[r];
#pragma unroll
      for (int i = 0; i < QPW; ++i) {
        const int q = j + 1 + wave + PLW * i;
        if (q < ncol) {
          const cplx a = sp[r + (long long)q * rows];
          d[i].x += xr.x * a.x + xr.y * a.y;
          d[i].y += xr.x * a.y - xr.y * a.x;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < QPW; ++i) {
      d[i].x = wave_sum(d[i].x);
      d[i].y = wave_sum(d[i].y);
    }
    const cplx tc = cconj(t), scc = cconj(sc);
#pragma unroll
    for (int i = 0; i < QPW; ++i) {
      const int q = j + 1 + wave + PLW * i;
      if (q < ncol) {
        cplx* aq = sp + (long long)q * rows;
        const cplx w = cadd(aq[j], cmul(scc, d[i]));
        const cplx f = cmul(tc, w);
        const cplx scf = cmul(sc, f);
        for (int r = j + lane; r < rows; r += 64) {
          if (r == j) aq[j] = csub(aq[j], f);
          else aq[r] = csub(aq[r], cmul(xj[r], scf));
        }
        if (q == j + 1) {
          double xs = 0.0;
          for (int r = j + 2 + lane; r < rows; r += 64) xs += cabs2(aq[r]);
          xs = wave_sum(xs);
          if (lane == 0) s_norm[j + 1] = xs;
        }
      }
    }
    __syncthreads();
  }
  // finalise the reflectors in place: R on and above the diagonal, scaled v below
  for (long long idx = tid; idx < (long long)rows * ncol; idx += blockDim.x) {
    const int r = (int)(idx % rows), c = (int)(idx / rows);
    if (r == c) sp[idx] = make_double2(s_beta[c], 0.0);
    else if (r > c) sp[idx] = cmul(sp[idx], s_sc[c]);
  }
  __syncthreads();
  for (long long idx = tid; idx < (long long)rows * ncol; idx += blockDim.x) {
    const int r = (int)(idx % rows), c = (int)(idx / rows);
    const cplx a = sp[idx];
    A[(k + r) + (long long)(k + c) * n] = a;
    cplx v = make_double2(0.0, 0.0);
    if (r == c) v = make_double2(1.0, 0.0);
    else if (r > c) v = a;
    Vw[(k + r) + (long long)(k + c) * n] = v;
  }
  auto Vget = [&](int r, int c) -> cplx {
    if (r >= rows || c >= ncol || r < c) return make_double2(0.0, 0.0);
    if (r == c) return make_double2(1.0, 0.0);
    return sp[r + (long long)c * rows];
  };
  // Gram = V^H V, one 16 x 16 block per wave
  {
    const int bi = wave & 1, bj = wave >> 1;
    v4d_t re = {0, 0, 0, 0}, im = {0, 0, 0, 0};
    if (16 * bi < ncol && 16 * bj < ncol) {
      cmfma16(
          rows, [&](int i, int kk) { return cconj(Vget(kk, 16 * bi + i)); },
          [&](int kk, int jj) { return Vget(kk, 16 * bj + jj); }, re, im);
      const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
      for (int r4 = 0; r4 < 4; ++r4) {
        const int gi = 16 * bi + 4 * r4 + lk, gj = 16 * bj + lr;
        if (gi < ncol && gj < ncol) {
          if (gi > gj) s_TG[gi + PB * gj] = make_double2(re[r4], im[r4]);
          if (gi == gj) s_vn2[gi] = re[r4];
        }
      }
    }
  }
  __syncthreads();
  // T recurrence (wave 0; lane i owns row i of T)
  if (wave == 0 && lane < ncol) {
    const int i = lane;
    s_TG[i + PB * i] = s_tau[i];
    for (int j = i + 1; j < ncol; ++j) {
      cplx acc = make_double2(0.0, 0.0);
      for (int l = i; l < j; ++l) {
        const cplx g = cconj(s_TG[j + PB * l]);       // G(l, j) = conj(G(j, l))
        acc = cfma(s_TG[i + PB * l], g, acc);
      }
      s_TG[i + PB * j] = cmul(make_double2(-s_tau[j].x, -s_tau[j].y), acc);
    }
  }
  __syncthreads();
  auto Tget = [&](int i, int j) -> cplx {
    return (i < ncol && j < ncol && i <= j) ? s_TG[i + PB * j] : make_double2(0.0, 0.0);
  };
  const int b = k / nb;
  for (int idx = tid; idx < ncol * ncol; idx += blockDim.x) {
    const int i = idx % ncol, j = idx / ncol;
    Tw[(long long)b * nb * nb + i + j * nb] = Tget(i, j);
  }
  // Y1 = V T^H, Y2 = V T
  {
    const int tr = (rows + 15) / 16, tc = (ncol + 15) / 16;
    const int lr = lane & 15, lk = lane >> 4;
    for (int tI = wave; tI < tr * tc; tI += PLW) {
      const int r0 = 16 * (tI % tr), c0 = 16 * (tI / tr);
      v4d_t re1 = {0, 0, 0, 0}, im1 = {0, 0, 0, 0}, re2 = {0, 0, 0, 0}, im2 = {0, 0, 0, 0};
      const int Kc = (ncol + 3) & ~3;
      cmfma16(
          Kc, [&](int i, int kk) { return Vget(r0 + i, kk); },
          [&](int kk, int jj) { return cconj(Tget(c0 + jj, kk)); }, re1, im1);
      if (Y2w)
        cmfma16(
            Kc, [&](int i, int kk) { return Vget(r0 + i, kk); }, [&](int kk, int jj) { return Tget(kk, c0 + jj); },
            re2, im2);
#pragma unroll
      for (int r4 = 0; r4 < 4; ++r4) {
        const int gi = r0 + 4 * r4 + lk, gj = c0 + lr;
        if (gi < rows && gj < ncol) {
          Y1w[(k + gi) + (long long)(k + gj) * n] = make_double2(re1[r4], im1[r4]);
          if (Y2w) Y2w[(k + gi) + (long long)(k + gj) * n] = make_double2(re2[r4], im2[r4]);
        }
      }
    }
  }
  if (tid == 0) {
    cplx dd = make_double2(1.0, 0.0);
    for (int j = 0; j < ncol; ++j) dd = cmul(dd, csub(make_double2(1.0, 0.0), cscale(s_tau[j], s_vn2[j])));
    pdet[b] = dd;
  }
}

// ---- panel factorisation with the panel resident in registers (rows <= 256) -----
// Same contract as panel_kernel except that Y2 = V T is not formed here (qr_form_q
// forms it with the GEMM).  1024 threads: the 32 lanes of half-wave c hold column c,
// rows l, l + 32, ..., l + 224 (l = lane in the half-wave).  Column j's reflector is
// formed by its own half-wave and handed to the others through a double-buffered LDS
// vector, one barrier per column; columns right of j apply H^H in registers, columns
// left of j (already final) use the step to accumulate the Gram entry v_c^H v_j.
// T follows from T^-1 = striu(V^H V) + diag(1/tau) (the columnwise zlarft recurrence
// T(0:j, j) = -tau_j T(0:j, 0:j) V^H v_j, inverted), by a fan-in solve per column,
// and Y1 = V T^H is an MFMA product out of LDS.
constexpr int RPL = 8;          // rows per lane: 256 rows
constexpr int RP = 32 * RPL;

__device__ __forceinline__ double hsum32(double v) { return half_sum(v); }

__global__ __launch_bounds__(1024) void panel_reg_kernel(cplx* __restrict__ A, int n, int k, int nb,
                                                         cplx* __restrict__ tau, cplx* __restrict__ Vw,
                                                         cplx* __restrict__ Y1w, cplx* __restrict__ Tw,
                                                         cplx* __restrict__ pdet, long long cs) {
  A = chain_ptr(A, cs);
  tau = chain_ptr(tau, cs);
  Vw = chain_ptr(Vw, cs);
  Y1w = chain_ptr(Y1w, cs);
  Tw = chain_ptr(Tw, cs);
  pdet = chain_ptr(pdet, cs);
  extern __shared__ cplx smem[];      // loop: vbuf[2][RP]; afterwards: V[RP x 32], column-major
  __shared__ cplx s_S[32 * 32];       // Gram above the diagonal, then T
  __shared__ cplx s_tau[32];
  __shared__ double s_vn2[32];
  const int tid = threadIdx.x;
  const int l = tid & 31, c = tid >> 5;
  const int rows = n - k;
  const int ncol = (k + nb <= n) ? nb : n - k;

  cplx a[RPL];
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int r = l + 32 * t;
    a[t] = (c < ncol && r < rows) ? A[(k + r) + (long long)(k + c) * n] : make_double2(0.0, 0.0);
  }
  for (int j = 0; j < ncol; ++j) {
    cplx* vb = smem + (j & 1) * RP;
    if (c == j) {
      // zlarfg on rows j.. of this column (alpha = row j: lane j, slot 0)
      const cplx alpha = make_double2(__shfl(a[0].x, j, 32), __shfl(a[0].y, j, 32));
      double xs = 0.0;
#pragma unroll
      for (int t = 0; t < RPL; ++t)
        if (l + 32 * t > j) xs += cabs2(a[t]);
      xs = hsum32(xs);
      cplx tj = make_double2(0.0, 0.0), sc = make_double2(1.0, 0.0);
      double beta = alpha.x;
      if (!(xs == 0.0 && alpha.y == 0.0)) {
        beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xs), alpha.x);
        const double rb = 1.0 / beta;
        tj = make_double2((beta - alpha.x) * rb, -alpha.y * rb);
        // 1 / (alpha - beta) by Smith's scaling with one reciprocal
        const double zx = alpha.x - beta, zy = alpha.y;
        if (fabs(zx) >= fabs(zy)) {
          const double r = zy / zx, id = 1.0 / (zx + r * zy);
          sc = make_double2(id, -r * id);
        } else {
          const double r = zx / zy, id = 1.0 / (zy + r * zx);
          sc = make_double2(r * id, -id);
        }
      }
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int r = l + 32 * t;
        cplx v = make_double2(r == j ? 1.0 : 0.0, 0.0);
        if (r > j) v = cmul(a[t], sc);
        vb[r] = v;
        if (r == j) a[t] = make_double2(beta, 0.0);
        else if (r > j) a[t] = v;
      }
      if (l == 0) {
        s_tau[j] = tj;
        tau[k + j] = tj;
      }
    }
    __syncthreads();
    if (c < ncol && c != j) {
      cplx vr[RPL];
#pragma unroll
      for (int t = 0; t < RPL; ++t) vr[t] = vb[l + 32 * t];
      if (c > j) {
        // a <- a - conj(tau) v (v^H a)
        double dx = 0.0, dy = 0.0;
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          dx += vr[t].x * a[t].x + vr[t].y * a[t].y;
          dy += vr[t].x * a[t].y - vr[t].y * a[t].x;
        }
        dx = hsum32(dx);
        dy = hsum32(dy);
        const cplx f = cmul(cconj(s_tau[j]), make_double2(dx, dy));
#pragma unroll
        for (int t = 0; t < RPL; ++t) a[t] = csub(a[t], cmul(f, vr[t]));
      } else {
        // Gram entry v_c^H v_j; column c holds R above its diagonal and v below
        double gx = 0.0, gy = 0.0;
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const int r = l + 32 * t;
          const cplx vc = r > c ? a[t] : make_double2(r == c ? 1.0 : 0.0, 0.0);
          gx += vc.x * vr[t].x + vc.y * vr[t].y;
          gy += vc.x * vr[t].y - vc.y * vr[t].x;
        }
        gx = hsum32(gx);
        gy = hsum32(gy);
        if (l == 0) s_S[c + 32 * j] = make_double2(gx, gy);
      }
    }
  }
  __syncthreads();
  // T column c by a fan-in solve of T^-1 T = 1 from the bottom row up:
  //   T(i, c) = tau_i (delta_ic - sum_{m > i} S(i, m) T(m, c))
  cplx tcol = make_double2(0.0, 0.0);
  if (c < ncol) {
    const cplx tl_own = l <= c ? s_tau[l] : make_double2(0.0, 0.0);
    cplx acc = make_double2(0.0, 0.0);
    for (int i = c; i >= 0; --i) {
      const cplx mine = cmul(tl_own, make_double2((l == c ? 1.0 : 0.0) - acc.x, -acc.y));
      const cplx ti = make_double2(__shfl(mine.x, i, 32), __shfl(mine.y, i, 32));
      if (l == i) tcol = ti;
      if (l < i) acc = cfma(s_S[l + 32 * i], ti, acc);
    }
  }
  // V into LDS (the reflector buffers are free), the panel and V to global memory
  double vn = 0.0;
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int r = l + 32 * t;
    cplx vc = make_double2(0.0, 0.0);
    if (c < ncol && r < rows) {
      vc = r > c ? a[t] : make_double2(r == c ? 1.0 : 0.0, 0.0);
      A[(k + r) + (long long)(k + c) * n] = a[t];
      Vw[(k + r) + (long long)(k + c) * n] = vc;
    }
    smem[r + RP * c] = vc;
    vn += cabs2(vc);
  }
  vn = hsum32(vn);
  __syncthreads();   // all reads of the Gram entries are done
  if (l == 0 && c < ncol) s_vn2[c] = vn;
  s_S[l + 32 * c] = (c < ncol && l <= c) ? tcol : make_double2(0.0, 0.0);
  const int b = k / nb;
  if (c < ncol && l < ncol) Tw[(long long)b * nb * nb + l + c * nb] = (l <= c) ? tcol : make_double2(0.0, 0.0);
  __syncthreads();
  // Y1 = V T^H: Y1(r, j) = sum_{m >= j} V(r, m) conj(T(j, m)), 16 x 16 MFMA tiles
  {
    const int wave = tid >> 6, lane = tid & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int tr = (rows + 15) / 16, tcn = (ncol + 15) / 16;
    const int kc = (ncol + 3) & ~3;
    for (int ti = wave; ti < tr * tcn; ti += 16) {
      const int r0 = 16 * (ti % tr), c0 = 16 * (ti / tr);
      v4d_t re = {0, 0, 0, 0}, im = {0, 0, 0, 0};
      for (int k0 = c0; k0 < kc; k0 += 4) {
        const cplx av = smem[(r0 + lr) + RP * (k0 + lk)];
        const cplx tv = s_S[(c0 + lr) + 32 * (k0 + lk)];
        const cplx bv = cconj(tv);
        re = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.x, re, 0, 0, 0);
        re = __builtin_amdgcn_mfma_f64_16x16x4f64(-av.y, bv.y, re, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.y, im, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f64_16x16x4f64(av.y, bv.x, im, 0, 0, 0);
      }
#pragma unroll
      for (int r4 = 0; r4 < 4; ++r4) {
        const int gi = r0 + 4 * r4 + lk, gj = c0 + lr;
        if (gi < rows && gj < ncol) Y1w[(k + gi) + (long long)(k + gj) * n] = make_double2(re[r4], im[r4]);
      }
    }
  }
  if (tid == 0) {
    cplx dd = make_double2(1.0, 0.0);
    for (int j = 0; j < ncol; ++j) dd = cmul(dd, csub(make_double2(1.0, 0.0), cscale(s_tau[j], s_vn2[j])));
    pdet[b] = dd;
  }
}

__device__ cplx block_cprod(cplx v, cplx* sh) {
  // product over the block of one complex per thread; result valid in all threads
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = cmul(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  cplx r = sh[0];
  __syncthreads();
  return r;
}

// ---- QDRP finishing --------------------------------------------------------------
// D(i) = |R_ii|, Rt = R with row i divided by D(i) (zeros below the diagonal).
// mode 0/1: UDV decompose, side R/L: phase = prod Rt_ii * sign(perm) (conj for L),
//           Rt row 0 /= phase, phase_out = phase, detU_out = det(Q) * phase.
// mode 2/3: CGR with NVAR = 1/2: z_out = conj(detRHS)/|detRHS| * sign(perm) *
//           prod phase(Rt_ii) * prod phase(det H_j)   (conj of the last two for NVAR = 2).
__global__ __launch_bounds__(PT) void qr_finish_kernel(const cplx* __restrict__ A, int n, int nblk,
                                                       const int* __restrict__ perm,
                                                       const cplx* __restrict__ pdet, int mode,
                                                       const cplx* __restrict__ detUr,
                                                       const cplx* __restrict__ detUl,
                                                       cplx* __restrict__ Dout, cplx* __restrict__ Rt,
                                                       cplx* __restrict__ phase_out,
                                                       cplx* __restrict__ detU_out, long long cs) {
  A = chain_ptr(A, cs);
  perm = chain_ptr(perm, cs);
  pdet = chain_ptr(pdet, cs);
  detUr = chain_ptr(detUr, cs);
  detUl = chain_ptr(detUl, cs);
  Dout = chain_ptr(Dout, cs);
  Rt = chain_ptr(Rt, cs);
  phase_out = chain_ptr(phase_out, cs);
  detU_out = chain_ptr(detU_out, cs);
  __shared__ cplx sh[PT];
  const int tid = threadIdx.x;
  // product of normalised diagonal entries
  cplx pd = make_double2(1.0, 0.0);
  for (int i = tid; i < n; i += blockDim.x) {
    const cplx a = A[i + (long long)i * n];
    const double x = cabs_(a);
    Dout[i] = make_double2(x, 0.0);
    cplx u = cscale(a, 1.0 / x);
    if (mode == 3) u = cconj(u);
    pd = cmul(pd, u);
  }
  pd = block_cprod(pd, sh);
  // reflector determinants
  cplx qd = make_double2(1.0, 0.0);
  for (int b = tid; b < nblk; b += blockDim.x) qd = cmul(qd, pdet[b]);
  qd = block_cprod(qd, sh);
  // permutation parity: the parity of its inversions, counted by all threads from a copy
  // of perm in LDS (n <= 4 PT; larger n walk the cycles in thread 0)
  __shared__ int s_par;
  int* sp = (int*)sh;
  const bool lds_perm = n <= 4 * PT;
  if (tid == 0) s_par = 0;
  if (lds_perm)
    for (int i = tid; i < n; i += blockDim.x) sp[i] = perm[i];
  __syncthreads();
  if (lds_perm) {
    int odd = 0;
    for (int i = tid; i < n; i += blockDim.x) {
      const int pi = sp[i];
      for (int j = i + 1; j < n; ++j) odd ^= (pi > sp[j]) ? 1 : 0;
    }
    if (odd) atomicXor(&s_par, 1);
    __syncthreads();
  }
  if (tid == 0) {
    int sgn = s_par ? -1 : 1;
    if (!lds_perm) {
      // cycle decomposition; sh as a visited bitmap (n <= 16384)
      sgn = 1;
      unsigned* vis = (unsigned*)sh;
      const int words = (n + 31) / 32;
      for (int w = 0; w < words; ++w) vis[w] = 0u;
      for (int i = 0; i < n; ++i) {
        if (vis[i >> 5] & (1u << (i & 31))) continue;
        int L = 0, nx = i;
        while (!(vis[nx >> 5] & (1u << (nx & 31)))) {
          vis[nx >> 5] |= 1u << (nx & 31);
          ++L;
          nx = perm[nx];
        }
        if ((L & 1) == 0) sgn = -sgn;
      }
    }
    cplx ph;
    if (mode <= 1) {
      ph = cscale(pd, (double)sgn);
      if (mode == 1) ph = cconj(ph);
      *phase_out = ph;
      *detU_out = cmul(qd, ph);
    } else {
      const cplx dr = *detUr, dl = *detUl;
      cplx drhs = cmul(cconj(dr), dl);              // det(UR^H UL)
      drhs = cscale(cconj(drhs), 1.0 / cabs_(drhs));
      cplx qph = cscale(qd, 1.0 / cabs_(qd));
      if (mode == 3) qph = cconj(qph);
      ph = cmul(cmul(drhs, cscale(pd, (double)sgn)), qph);
      *phase_out = ph;
    }
  }
}

// Rt = R with row i divided by |R_ii| (zero below the diagonal); for the UDV decompose
// (mode 0/1) row 0 is further divided by the decompose phase.
__global__ void qr_rt_kernel(const cplx* __restrict__ A, int n, int mode, const cplx* __restrict__ phase,
                             cplx* __restrict__ Rt, long long cs) {
  A = chain_ptr(A, cs);
  phase = chain_ptr(phase, cs);
  Rt = chain_ptr(Rt, cs);
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  cplx v = make_double2(0.0, 0.0);
  if (i <= j) {
    v = cscale(A[idx], 1.0 / cabs_(A[i + (long long)i * n]));
    if (i == 0 && mode <= 1) v = cmul(v, cinv(*phase));
  }
  Rt[idx] = v;
}

// X(:, 0) *= s
__global__ void scale_col0_kernel(cplx* __restrict__ X, int n, const cplx* __restrict__ s, long long cs) {
  X = chain_ptr(X, cs);
  s = chain_ptr(s, cs);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) X[i] = cmul(X[i], *s);
}

// Invert the nb x nb diagonal blocks of the upper-triangular Rt into Rinv (one block per WG).
__global__ __launch_bounds__(PT) void trinv_diag_kernel(const cplx* __restrict__ Rt, int n, int nb,
                                                        cplx* __restrict__ Rinv, long long cs) {
  Rt = chain_ptr(Rt, cs);
  Rinv = chain_ptr(Rinv, cs);
  __shared__ cplx s[NBMAX * NBMAX];
  __shared__ cplx x[NBMAX * NBMAX];
  const int b = blockIdx.x;
  const int k = b * nb;
  const int m = (k + nb <= n) ? nb : n - k;
  for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
    const int i = idx % m, j = idx / m;
    s[i + NBMAX * j] = Rt[(k + i) + (long long)(k + j) * n];
    x[i + NBMAX * j] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  // column j of the inverse by back substitution, eight lanes per column (blockDim.x = 8
  // NBMAX): lane q adds the terms l = i + 1 + q, i + 9 + q, ... of row i, the eight partial
  // sums meet by three exchanges inside the lane group, and lane 0 of the group stores x(i, j),
  // which the group reads back in later rows (LDS operations of one wave complete in order)
  {
    static_assert(PT == 8 * NBMAX, "the kernel is launched with eight lanes per column");
    const int j = threadIdx.x >> 3, q = threadIdx.x & 7;
    if (j < m) {
      for (int i = j; i >= 0; --i) {
        cplx acc = make_double2(0.0, 0.0);
        for (int l = i + 1 + q; l <= j; l += 8) acc = cfma(s[i + NBMAX * l], x[l + NBMAX * j], acc);
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
          acc.x += __shfl_xor(acc.x, o, 8);
          acc.y += __shfl_xor(acc.y, o, 8);
        }
        if (q == 0) x[i + NBMAX * j] = cdiv(make_double2((i == j ? 1.0 : 0.0) - acc.x, -acc.y), s[i + NBMAX * i]);
      }
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
    const int i = idx % m, j = idx / m;
    Rinv[(k + i) + (long long)(k + j) * n] = x[i + NBMAX * j];
  }
}

inline unsigned nblocks(long long count, int per) { return (unsigned)((count + per - 1) / per); }

}  // namespace

void qr_work_place(QRWork& w, int n, ArenaLayout& ar) {
  w.n = n;
  // panel width: the widest whose rows x nb panel fits the LDS kernel's 128 KiB
  if (n <= 256) w.nb = 32;
  else if (n <= 512) w.nb = 16;
  else if (n <= 1024) w.nb = 8;
  else if (n <= 2048) w.nb = 4;
  else w.nb = 32;
  w.lds_panel = n <= 2048;
  w.nblk = (n + w.nb - 1) / w.nb;
  const size_t nn = (size_t)n * n;
  w.A = ar.take<cplx>(nn);
  w.Vw = ar.take<cplx>(nn);   // the panels never write above their diagonal block: zero from creation
  w.Y1 = ar.take<cplx>(nn);
  w.Y2 = ar.take<cplx>(nn);
  w.Rt = ar.take<cplx>(nn);
  w.Rinv = ar.take<cplx>(nn);
  w.Tmp = ar.take<cplx>(nn);
  w.T = ar.take<cplx>((size_t)w.nblk * w.nb * w.nb);
  w.W = ar.take<cplx>((size_t)w.nb * n);
  w.tau = ar.take<cplx>(n);
  w.pdet = ar.take<cplx>(w.nblk);
  w.norms = ar.take<double>(n);
  w.perm = ar.take<int>(n);
  w.pinv = ar.take<int>(n);
  w.scal = ar.take<cplx>(8 * (size_t)n);
}

static int gemm_w(const QRWork& w, GemmArgs g, hipStream_t s) {
  g.nchain = w.nchain;
  g.cstride = w.cs;
  g.policy = w.policy;
  return gemm_launch(g, s);
}

int launch_fill(cplx* X, long long count, cplx val, int nchain, long long cs, hipStream_t s) {
  if (count <= 0) return ST_NONE;
  hipLaunchKernelGGL(fill_kernel, dim3(nblocks(count, 256), 1, nchain), dim3(256), 0, s, X, count, val, cs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

int launch_identity(cplx* X, int n, int nchain, long long cs, hipStream_t s) {
  hipLaunchKernelGGL(identity_kernel, dim3(nblocks((long long)n * n, 256), 1, nchain), dim3(256), 0, s, X, n, cs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

// Factor the (presorted) source into w.A = Q R; leaves V/T/Y panels in w.
static int qr_factor(QRWork& w, const cplx* S, int lds, const cplx* D, int ctrans, hipStream_t s) {
  const int n = w.n, nb = w.nb;
  int st;
  hipLaunchKernelGGL(colnorm_kernel, dim3(nblocks(n, 4), 1, w.nchain), dim3(256), 0, s, S, n, lds, D, ctrans,
                     w.norms, w.cs);
  hipLaunchKernelGGL(sortperm_kernel, dim3(1, 1, w.nchain), dim3(1024), 0, s, w.norms, n, w.perm, w.pinv, w.cs);
  hipLaunchKernelGGL(gather_kernel, dim3(nblocks((long long)n * n, 256), 1, w.nchain), dim3(256), 0, s, S, n, lds, D,
                     ctrans, w.perm, w.A, w.cs);
  HIP_OK(hipGetLastError());
  const cplx one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0), mone = make_double2(-1.0, 0.0);
  for (int k = 0; k < n; k += nb) {
    // the register panel takes any panel of at most RP rows (the kernel works on rows k..n-1
    // with leading dimension n), so the later panels of a larger matrix use it too
    if (n - k <= RP && nb <= 32) {
      static std::once_flag once_r;
      static hipError_t attr_r = hipSuccess;
      std::call_once(once_r, [] {
        attr_r = hipFuncSetAttribute((const void*)panel_reg_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(RP * 32 * sizeof(cplx)));
      });
      HIP_OK(attr_r);
      hipLaunchKernelGGL(panel_reg_kernel, dim3(1, 1, w.nchain), dim3(1024), RP * 32 * sizeof(cplx), s, w.A, n, k, nb,
                         w.tau, w.Vw, w.Y1, w.T, w.pdet, w.cs);
    } else if (w.lds_panel) {
      const size_t lds = (size_t)(n - k) * std::min(nb, n - k) * sizeof(cplx);
      static std::once_flag once;
      static hipError_t attr_err = hipSuccess;
      std::call_once(once, [] {
        attr_err = hipFuncSetAttribute((const void*)panel_lds_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       128 * 1024);
      });
      HIP_OK(attr_err);
      hipLaunchKernelGGL(panel_lds_kernel, dim3(1, 1, w.nchain), dim3(PLT), lds, s, w.A, n, k, nb, w.tau, w.Vw, w.Y1,
                         (cplx*)nullptr, w.T, w.pdet, w.cs);
    } else {
      hipLaunchKernelGGL(panel_kernel, dim3(1, 1, w.nchain), dim3(PT), 0, s, w.A, n, k, nb, w.tau, w.Vw, w.Y1,
                         (cplx*)nullptr, w.T, w.pdet, w.cs);
    }
    HIP_OK(hipGetLastError());
    const int kc = k + nb;
    if (kc >= n) break;
    const int rows = n - k, ntr = n - kc;
    // W = V^H A(k:, kc:)
    GemmArgs g1 = gemm_args(nb, ntr, rows, w.Vw + k + (long long)k * n, n, OP_C, true,
                            w.A + k + (long long)kc * n, n, OP_N, true, w.W, nb, one, zero);
    if ((st = gemm_w(w, g1, s))) return st;
    // A(k:, kc:) -= Y1 W
    GemmArgs g2 = gemm_args(rows, ntr, nb, w.Y1 + k + (long long)k * n, n, OP_N, true, w.W, nb, OP_N, true,
                            w.A + k + (long long)kc * n, n, mone, one);
    if ((st = gemm_w(w, g2, s))) return st;
  }
  return ST_NONE;
}

// Q = H_0 H_1 ... H_{n-1} formed explicitly into Q (n x n).
static int qr_form_q(QRWork& w, cplx* Q, hipStream_t s) {
  const int n = w.n, nb = w.nb;
  int st;
  if ((st = launch_identity(Q, n, w.nchain, w.cs, s))) return st;
  const cplx one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0), mone = make_double2(-1.0, 0.0);
  // Y2 = V T for every panel (V is zero above each panel's diagonal block)
  if (n % nb == 0) {
    GemmArgs gy = gemm_args(n, nb, nb, w.Vw, n, OP_N, true, w.T, nb, OP_N, true, w.Y2, n, one, zero);
    gy.batch = w.nblk;
    gy.strideA = (long long)n * nb;
    gy.strideB = (long long)nb * nb;
    gy.strideC = (long long)n * nb;
    if ((st = gemm_w(w, gy, s))) return st;
  } else {
    for (int b = 0; b < w.nblk; ++b) {
      const int k = b * nb;
      const int ncol = (k + nb <= n) ? nb : n - k;
      GemmArgs gy = gemm_args(n - k, ncol, ncol, w.Vw + k + (long long)k * n, n, OP_N, true,
                              w.T + (long long)b * nb * nb, nb, OP_N, true, w.Y2 + k + (long long)k * n, n, one, zero);
      if ((st = gemm_w(w, gy, s))) return st;
    }
  }
  for (int b = w.nblk - 1; b >= 0; --b) {
    const int k = b * nb;
    const int ncol = (k + nb <= n) ? nb : n - k;
    const int rows = n - k;
    // W = V^H Q(k:, k:)
    GemmArgs g1 = gemm_args(ncol, rows, rows, w.Vw + k + (long long)k * n, n, OP_C, true,
                            Q + k + (long long)k * n, n, OP_N, true, w.W, nb, one, zero);
    if ((st = gemm_w(w, g1, s))) return st;
    // Q(k:, k:) -= Y2 W
    GemmArgs g2 = gemm_args(rows, rows, ncol, w.Y2 + k + (long long)k * n, n, OP_N, true, w.W, nb, OP_N, true,
                            Q + k + (long long)k * n, n, mone, one);
    if ((st = gemm_w(w, g2, s))) return st;
  }
  return ST_NONE;
}

int udv_decompose_dev(QRWork& w, const cplx* Usrc, const cplx* Dsrc, const cplx* Vsrc, int side_l,
                      cplx* Uout, cplx* Dout, cplx* Vout, cplx* phase_out, cplx* detU_out, hipStream_t s) {
  const int n = w.n;
  int st;
  if ((st = qr_factor(w, Usrc, n, Dsrc, 0, s))) return st;
  hipLaunchKernelGGL(qr_finish_kernel, dim3(1, 1, w.nchain), dim3(PT), 0, s, w.A, n, w.nblk, w.perm, w.pdet,
                     side_l ? 1 : 0, (const cplx*)nullptr, (const cplx*)nullptr, Dout, w.Rt, phase_out, detU_out, w.cs);
  hipLaunchKernelGGL(qr_rt_kernel, dim3(nblocks((long long)n * n, 256), 1, w.nchain), dim3(256), 0, s, w.A, n,
                     side_l ? 1 : 0, (const cplx*)phase_out, w.Rt, w.cs);
  HIP_OK(hipGetLastError());
  const cplx one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  if (!side_l) {
    // V = Rt * P^T * V   (row i of P^T V is row perm[i] of V)
    GemmArgs g = gemm_args(n, n, n, w.Rt, n, OP_N, true, Vsrc, n, OP_N, true, Vout, n, one, zero);
    g.bmap = w.perm;
    if ((st = gemm_w(w, g, s))) return st;
  } else {
    // V = V P * Rt^H    (column i of V P is column perm[i] of V)
    GemmArgs g = gemm_args(n, n, n, Vsrc, n, OP_N, true, w.Rt, n, OP_C, true, Vout, n, one, zero);
    g.amap = w.perm;
    if ((st = gemm_w(w, g, s))) return st;
  }
  if ((st = qr_form_q(w, Uout, s))) return st;
  hipLaunchKernelGGL(scale_col0_kernel, dim3(nblocks(n, 256), 1, w.nchain), dim3(256), 0, s, Uout, n,
                     (const cplx*)phase_out, w.cs);
  HIP_OK(hipGetLastError());
  return ST_NONE;
}

int qr_factor_cgr(QRWork& w, const cplx* TPUP, int nvar, const cplx* detUr, const cplx* detUl, cplx* Qout,
                  cplx* DUP, cplx* z_out, hipStream_t s) {
  const int n = w.n;
  int st;
  if ((st = qr_factor(w, TPUP, n, nullptr, nvar == 1 ? 0 : 1, s))) return st;
  hipLaunchKernelGGL(qr_finish_kernel, dim3(1, 1, w.nchain), dim3(PT), 0, s, w.A, n, w.nblk, w.perm, w.pdet,
                     nvar == 1 ? 2 : 3, detUr, detUl, DUP, w.Rt, z_out, (cplx*)nullptr, w.cs);
  hipLaunchKernelGGL(qr_rt_kernel, dim3(nblocks((long long)n * n, 256), 1, w.nchain), dim3(256), 0, s, w.A, n,
                     nvar == 1 ? 2 : 3, (const cplx*)z_out, w.Rt, w.cs);
  HIP_OK(hipGetLastError());
  if ((st = qr_form_q(w, Qout, s))) return st;
  return ST_NONE;
}

int tri_inverse_dev(QRWork& w, const cplx* Rt, cplx* Rinv, hipStream_t s) {
  const int n = w.n, nb = w.nb;
  int st;
  if ((st = launch_fill(Rinv, (long long)n * n, make_double2(0.0, 0.0), w.nchain, w.cs, s))) return st;
  hipLaunchKernelGGL(trinv_diag_kernel, dim3(w.nblk, 1, w.nchain), dim3(PT), 0, s, Rt, n, nb, Rinv, w.cs);
  HIP_OK(hipGetLastError());
  const cplx one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0), mone = make_double2(-1.0, 0.0);
  for (int b = 1; b < w.nblk; ++b) {
    const int k = b * nb;
    const int m = (k + nb <= n) ? nb : n - k;
    // T1 = Rt(0:k, k:k+m) Rinv(k:k+m, k:k+m)
    GemmArgs g1 = gemm_args(k, m, m, Rt + (long long)k * n, n, OP_N, true, Rinv + k + (long long)k * n, n, OP_N,
                            true, w.Tmp, n, one, zero);
    if ((st = gemm_w(w, g1, s))) return st;
    // Rinv(0:k, k:k+m) = -Rinv(0:k, 0:k) T1
    GemmArgs g2 = gemm_args(k, m, k, Rinv, n, OP_N, true, w.Tmp, n, OP_N, true, Rinv + (long long)k * n, n, mone,
                            zero);
    if ((st = gemm_w(w, g2, s))) return st;
  }
  return ST_NONE;
}
