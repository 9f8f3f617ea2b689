// Metropolis chain kernel interface (chain.hip).
#pragma once
#include "alfq_common.h"

#include <utility>
#include <vector>

#define CHAIN_KMAX 64   // largest delayed-update depth
#define CHAIN_W_LD 32   // leading dimension of the block kernel's W

// Per-chain device state touched by the sequential Metropolis kernel.
struct ChainState {
  cplx phase;                     // Phase of the weight (upgrade_mod.F90: PHASE)
  long long nc, acc;              // Control_upgrade counters (control_mod.F90:164-169)
  unsigned long long rng_pos;     // global index of the next uniform
  int pos;                        // next visit within the current slice
  int m;                          // pending delayed updates in X/Y
  int status;                     // 1: random stream exhausted
  int pad;
  unsigned long long rng_slice_base;  // rng_pos at the start of the current slice
  int slice_acc;                  // accepted flips in the current slice
  int max_slice_acc;              // largest slice_acc seen since the host last cleared it
  unsigned long long u_base;      // global index of u[0] in the uniform buffer
  long long u_len;                // uniforms held by the buffer
  unsigned long long seed;        // Philox key of this chain
  unsigned stream;                // Philox stream of this chain
  int pad2;
};

struct ChainArgs {
  cplx* G; int n;
  cplx* X; cplx* Y; int Kmax;
  int8_t* f; int n_op; int slice; int dir;
  const int* op_site; const int* op_type;
  const cplx* op_myexp;           // n_op x 25: exp(Z1 E) for (old, new) field values
  const cplx* op_expalpha;        // n_op x 25: exp(Z1 alpha)
  const double* gama_ratio;       // n_op x 25: Gama(new)/Gama(old)
  int n_sun;
  const double* u;                // uniforms u[q - st->u_base], q in [st->u_base, st->u_base + st->u_len)
  const int* uoff;                // uniforms consumed before visit v of the slice (n_op + 1 entries)
  int first_round;                // start of the slice: visit 0, rng base = current position
  int final_round;                // last scheduled round: flush in place instead of stopping at K
  ChainState* st;
  // block kernel outputs for the flush G += (G(:, amap) W) Y^T: W (32 x 32, ld 32), site list
  cplx* W; int* amap;
  // nonzero: the round's X = G(:, amap) W product writes Yhat (GemmArgs::yhat), so the block
  // kernel leaves Y alone except in its own in-place flush
  int yhat_external;
  // batches: byte stride between the chains' arenas; G, X, Y, f, the coupling tables
  // (op_myexp, op_expalpha, gama_ratio), u, st, W and amap are per chain, op_site,
  // op_type and uoff are shared
  long long cs;
  int nchain;
  // batches of uneven Ltrot: chain z has lz[z] slices; slices at or beyond it hold no fields
  // (null: every chain has the batch's Ltrot)
  const int* lz;
};

// Optional per-launch timing of the chain kernel: start/stop events recorded by the
// launch itself (hipExtLaunchKernelGGL) on the launch's stream.
struct KTimer {
  bool on = false;
  size_t used = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
};

int launch_chain(const ChainArgs& a, hipStream_t s, KTimer* timer = nullptr);
// largest delay depth the submatrix block kernel supports (deeper delays use the generic kernel)
int chain_max_delay(int n);
// uniforms of global indices [st->rng_pos, st->rng_pos + count) into u with the chain's
// Philox key and stream (st->seed, st->stream); records the window in st
int launch_philox_fill(double* u, ChainState* st, long long count, int nchain, long long cs, hipStream_t s);
// f, eexp, d, dinv, dconj (and phasefac, out) per chain; site_ptr, site_ops shared
int launch_dvec(const int8_t* f, int n_op, int s0, int ns, int n, const int* site_ptr, const int* site_ops,
                const cplx* eexp, cplx* d, cplx* dinv, cplx* dconj, int nchain, long long cs, hipStream_t s);
int launch_opphase(const int8_t* f, long long count, int n_op, const cplx* phasefac, cplx* out, int nchain,
                   long long cs, hipStream_t s);
