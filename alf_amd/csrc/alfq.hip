// Context, sweep orchestration and the C ABI of the gfx950 DQMC sweep engine.
//
// The control flow restates Prog/main.F90:589-631 (initialisation) and
// 714-887 (sequential sweep); the per-slice work maps onto the kernels as
//   WRAPGRUP:  GEMM exp(-dtau T) G, GEMM (.) exp(dtau T) with the slice's
//              diagonal exponentials in the epilogue, then Metropolis rounds
//              (chain kernel + rank-K flush GEMM)
//   WRAPGRDO:  Metropolis rounds, then GEMM with the new diagonal exponentials
//              and exp(-dtau T) fused, then GEMM exp(dtau T) (.)
//   WRAPUR/UL: Nwrap GEMMs on U, then the presorted blocked QR (qr.hip)
//   CGR:       two GEMMs for the scale-separated middle matrix, its QR,
//              triangular inverse and three GEMMs for G
// Everything stays on the device; the host only enqueues work.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/alfq.h"
#include "alfq_cb.h"
#include "alfq_chain.h"
#include "alfq_obs.h"
#include "alfq_oz.h"
#include "alfq_qr.h"
#include "alfq_tau.h"

static thread_local std::string g_last_error;
std::recursive_mutex& null_stream_mutex() {
  static std::recursive_mutex m;
  return m;
}

void alfq_set_error(const std::string& msg) { g_last_error = msg; }

namespace {

struct DevScalars {
  cplx z_cgr;             // phase returned by CGR
  cplx opph;              // Op_phase product
  double xmaxg, xmeang;   // Control_PrecisionG accumulators
  long long ncg;
  double xmaxp;           // Control_PrecisionP
  int status;             // 1: NaN in G, 3: XMAX > 10
  int pad;
  long long nlog;         // records written to the precision log
};

struct UDVSlot {
  cplx* U = nullptr;
  cplx* V = nullptr;
  cplx* D = nullptr;
  cplx* sc = nullptr;     // [0] = decompose phase, [1] = det(U)
  int refs = 0;
  char side = 'R';
};

constexpr int LOG_CAP = 1 << 16;

__global__ void cgr_scales_kernel(const cplx* __restrict__ DR, const cplx* __restrict__ DL, int n,
                                  cplx* __restrict__ out, long long cs) {
  // out: [0] 1/DR+  [1] 1/DL+  [2] DR-  [3] DL-   (cgr1_mod.F90:238-268 split at 1.0)
  DR = chain_ptr(DR, cs);
  DL = chain_ptr(DL, cs);
  out = chain_ptr(out, cs);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dr = DR[i].x, dl = DL[i].x;
  out[i] = make_double2(dr <= 1.0 ? 1.0 : 1.0 / dr, 0.0);
  out[n + i] = make_double2(dl <= 1.0 ? 1.0 : 1.0 / dl, 0.0);
  out[2 * n + i] = make_double2(dr <= 1.0 ? dr : 1.0, 0.0);
  out[3 * n + i] = make_double2(dl <= 1.0 ? dl : 1.0, 0.0);
}

__global__ void recip_real_kernel(const cplx* __restrict__ D, int n, cplx* __restrict__ out, long long cs) {
  D = chain_ptr(D, cs);
  out = chain_ptr(out, cs);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_double2(1.0 / D[i].x, 0.0);
}

// Control_PrecisionG on (fresh, wrapped), then Op_phase/Z**N_SUN/Control_PrecisionP and
// PHASE = Z (main.F90:744-753).  With check == 0 only PHASE = Z.
__global__ __launch_bounds__(1024) void stabilise_kernel(const cplx* __restrict__ Gf, const cplx* __restrict__ Gw,
                                                         int n, int check, int n_sun, DevScalars* sc,
                                                         ChainState* st, double* __restrict__ log, long long cs) {
  __shared__ double smax[1024];
  Gf = chain_ptr(Gf, cs);
  Gw = chain_ptr(Gw, cs);
  sc = chain_ptr(sc, cs);
  st = chain_ptr(st, cs);
  log = chain_ptr(log, cs);
  __shared__ double ssum[1024];
  __shared__ int snan;
  const int tid = threadIdx.x;
  if (tid == 0) snan = 0;
  __syncthreads();
  double mx = 0.0, sm = 0.0;
  int nan = 0;
  if (check) {
    // |a - b| as sqrt(dx^2 + dy^2): entries of G are O(1), and an overflow gives inf, which
    // fails the XMAX check as the value itself would; four entries per step keep loads ahead
    const int nn = n * n;
    int i = tid;
    for (; i + 3 * (int)blockDim.x < nn; i += 4 * blockDim.x) {
      cplx a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = Gf[i + u * blockDim.x];
        b[u] = Gw[i + u * blockDim.x];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (a[u].x != a[u].x || a[u].y != a[u].y || b[u].x != b[u].x || b[u].y != b[u].y) nan = 1;
        const double dx = a[u].x - b[u].x, dy = a[u].y - b[u].y;
        const double d = sqrt(dx * dx + dy * dy);
        mx = fmax(mx, d);
        sm += d;
      }
    }
    for (; i < nn; i += blockDim.x) {
      const cplx a = Gf[i], b = Gw[i];
      if (a.x != a.x || a.y != a.y || b.x != b.x || b.y != b.y) nan = 1;
      const double dx = a.x - b.x, dy = a.y - b.y;
      const double d = sqrt(dx * dx + dy * dy);
      mx = fmax(mx, d);
      sm += d;
    }
  }
  smax[tid] = mx;
  ssum[tid] = sm;
  if (nan) atomicOr(&snan, 1);
  __syncthreads();
  for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
    if (tid < s) {
      smax[tid] = fmax(smax[tid], smax[tid + s]);
      ssum[tid] += ssum[tid + s];
    }
    __syncthreads();
  }
  if (tid == 0) {
    cplx z1 = cmul(sc->z_cgr, sc->opph);
    cplx z = z1;
    for (int p = 1; p < n_sun; ++p) z = cmul(z, z1);
    if (check) {
      if (snan) {
        sc->status = 1;
      } else {
        const double xmax = smax[0];
        const double xmean = ssum[0] / ((double)n * (double)n);
        sc->ncg += 1;
        if (xmax > 10.0) sc->status = 3;
        sc->xmaxg = fmax(sc->xmaxg, xmax);
        sc->xmeang += xmean;
        const double dp = hypot(z.x - st->phase.x, z.y - st->phase.y);
        sc->xmaxp = fmax(sc->xmaxp, dp);
        if (sc->nlog < LOG_CAP) {
          double* r = log + 3 * sc->nlog;
          r[0] = xmax;
          r[1] = xmean;
          r[2] = dp;
        }
        sc->nlog += 1;
      }
    }
    st->phase = z;
  }
}

__global__ void chain_reset_kernel(ChainState* st, long long cs) {
  st = chain_ptr(st, cs);
  st->pos = 0;
  st->m = 0;
  st->rng_slice_base = st->rng_pos;
}

inline unsigned nblk(long long count, int per) { return (unsigned)((count + per - 1) / per); }

}  // namespace

struct alfq_ctx {
  int n = 0, L = 0, W = 0, nstm = 0, n_op = 0, n_sun = 0;
  bool t_real = true;
  int K = 32;
  int gemm_policy = 0;                   // GEMM tiling policy (alfq_set_gemm_policy)
  int prec = ALFQ_PREC_FP64;             // arithmetic of the hopping wraps of G
  // ALFQ_PREC_INT8_SPLIT: digit planes of exp(-dtau T) and exp(dtau T) as left operands (by
  // row) and right operands (by column), and of the wrapped matrix, oz_cs bytes per chain
  char* oz_buf = nullptr;
  long long oz_cs = 0;
  OzOperand oz_e_row, oz_e_col, oz_ei_row, oz_ei_col, oz_x;
  bool symm = false;                     // Symm = .T.: measurements on Hop_mod_Symm(G)
  // equal-time measurements
  ObsDev obs;
  long long obs_cs = 0;                  // bytes per chain of the measurement blocks
  int* obs_tables = nullptr;             // site_of, jm, oi, oj (shared)
  std::vector<int> obs_key;              // host copy of the shared tables (batch consistency)
  bool obs_on = false;
  int lobs_st = 1, lobs_en = 0;
  // checkerboard Op_T: bond sites (shared), bond exponentials per chain
  CbOps cb{};
  CbOps cb_inv{};
  CbOps cb_half{};                       // Symm: bond half steps exp(g O_b / 2) and inverses
  CbOps cb_half_inv{};
  bool cb_symm = false;
  bool cb_on = false;
  char* cb_buf = nullptr;
  int2* cb_sites = nullptr;
  std::vector<int> cb_key;
  std::vector<char> cb_set;
  std::vector<char> cb_real;
  // time-displaced Green's functions and measurements (Ltau = 1): the buffers sit in regions
  // of cs bytes per chain, so they keep the chain stride of the arena; the sums per chain at
  // acc + z * acc_cs
  struct TauDev {
    char* mem = nullptr;
    int regions = 0;
    cplx* b[5] = {};                     // G(tau,0), G(0,tau), G(tau,tau), G(0,0), scratch
    cplx *Xv = nullptr, *Qv = nullptr, *F1 = nullptr, *F2 = nullptr;
    cplx *H2 = nullptr, *Q2 = nullptr, *Y2 = nullptr, *O2 = nullptr;
    cplx *DUPv = nullptr, *DUPvinv = nullptr, *DUP2 = nullptr, *DUP2inv = nullptr, *zs = nullptr;
    int* flag = nullptr;
    TauCtl* ctl = nullptr;
    double* part = nullptr;              // partial precision sums of CGR2_2's extraction
    QRWork w2;                           // QR of the 2n x 2n matrix of CGR2_2
    char* acc = nullptr;
    long long acc_cs = 0, off_den0 = 0, off_latt = 0;
  } tau;
  bool tau_on = false;
  int var_tr = -1;                       // TAU_M's right chain
  hipStream_t s = nullptr;
  // batch: B chains in lockstep; chain z's buffers sit at (chain 0's pointer) + z * cs bytes
  int B = 1;
  long long cs = 0;
  char* arena = nullptr;
  int cur = 0;                           // chain addressed by the per-chain accessors
  std::vector<int> stab;                 // Stab_nt(0:NSTM)
  // model
  void* expT = nullptr;
  void* expTinv = nullptr;
  void* expTh = nullptr;                 // Symm: exp(-dtau T / 2), exp(dtau T / 2)
  void* expThinv = nullptr;
  int* op_site = nullptr;
  int* op_type = nullptr;
  int* site_ptr = nullptr;
  int* site_ops = nullptr;
  cplx* eexp = nullptr;
  cplx* myexp = nullptr;
  cplx* expalpha = nullptr;
  cplx* phasefac = nullptr;
  double* gama_ratio = nullptr;
  int* uoff_up = nullptr;                // uniforms consumed before visit v (ops ascending)
  int* uoff_dn = nullptr;                // the same for the descending visit order
  std::vector<int> h_op_type;
  std::vector<int> h_op_site;
  // batches of uneven Ltrot: chain z has h_lz[z] <= L slices (L is the batch's largest); its
  // later slices hold no fields and apply the identity (lz on the device, ident an identity of
  // exp(-dtau T)'s type); empty when every chain has L slices
  std::vector<int> h_lz;
  int* lz = nullptr;
  void* ident = nullptr;
  // state
  int8_t* f = nullptr;
  cplx* G = nullptr;
  cplx* Gtmp = nullptr;
  cplx* H = nullptr;
  cplx* X = nullptr;
  cplx* Y = nullptr;
  cplx* Wk = nullptr;       // block kernel: W = C D^T of the last round (32 x 32)
  int* amapk = nullptr;     // block kernel: sites of the last round's accepted flips
  ChainState* st = nullptr;
  DevScalars* sc = nullptr;
  double* log = nullptr;
  // UDV pool
  std::vector<UDVSlot> pool;
  int var_r = -1, var_l = -1;
  std::vector<int> var_st;               // udvst(1..NSTM) -> slot
  // scratch
  cplx* dv_d = nullptr;
  cplx* dv_dinv = nullptr;
  cplx* dv_dconj = nullptr;
  cplx* U0 = nullptr;
  cplx* U1 = nullptr;
  cplx* TP = nullptr;
  cplx* Qc = nullptr;
  cplx* DUP = nullptr;
  cplx* DUPinv = nullptr;
  cplx* Xc = nullptr;
  cplx* Yc = nullptr;
  cplx* cscal = nullptr;
  QRWork qr;
  // random numbers
  int rng_mode = 0;                      // 0: none, 1: tape, 2: philox
  double* u = nullptr;                   // uniforms in use: the arena's Philox buffer or a tape
  double* u_arena = nullptr;             // the arena's Philox buffer (u_cap uniforms per chain)
  double* u_tape = nullptr;              // caller-supplied tape (single chains)
  long long u_cap = 0;
  long long u_arena_cap = 0;
  long long u_len = 0;
  unsigned long long u_base = 0;
  unsigned long long rng_next = 0;       // host mirror of st->rng_pos
  unsigned long long seed = 0;
  unsigned stream_id = 0;
  long long per_visit_sum = 0;           // uniforms per pass over all ops
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  KTimer chain_timer;
  int rounds = 1 << 30;                  // chain rounds per slice (adapted per sweep)
  bool round_fusion = true;              // one launch for a round's X product and flush (few chains)
  int* stat_dev = nullptr;               // 3 words per chain read after every sweep
  int* stat_host = nullptr;              // their pinned host copy
  bool uniforms_ready = false;           // a whole sweep's uniforms are already on the device
  // sweeps replayed as HIP graphs, keyed by the host state they start from
  struct HostState {
    cplx* G; cplx* Gtmp; double* u;
    int var_r, var_l, rounds, K, rng_mode, prec, gemm_policy;
    bool obs_on, cb_on, tau_on, round_fusion;
    std::vector<int> var_st, refs;
    bool operator==(const HostState& o) const {
      return G == o.G && Gtmp == o.Gtmp && u == o.u && var_r == o.var_r && var_l == o.var_l && rounds == o.rounds &&
             K == o.K && rng_mode == o.rng_mode && prec == o.prec && gemm_policy == o.gemm_policy &&
             obs_on == o.obs_on && cb_on == o.cb_on &&
             tau_on == o.tau_on && round_fusion == o.round_fusion &&
             var_st == o.var_st && refs == o.refs;
    }
  };
  struct SweepGraph {
    HostState start, end;
    unsigned long long rng_advance;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<SweepGraph> graphs;
  bool graphs_enabled = true;
  bool capturing = false;
  int eager_sweeps = 0;
};

namespace {

// GEMM over every chain of the context
int gemm_c(alfq_ctx* c, GemmArgs g) {
  g.nchain = c->B;
  g.cstride = c->cs;
  g.policy = c->gemm_policy;
  return gemm_launch(g, c->s);
}

int slot_alloc(alfq_ctx* c) {
  // the slots are placed in the chains' arenas at creation; a sweep needs at most NSTM + 4
  for (size_t i = 0; i < c->pool.size(); ++i)
    if (c->pool[i].refs == 0) { c->pool[i].refs = 1; return (int)i; }
  alfq_set_error("slot_alloc: all UDV slots are in use");
  return -1;
}

void slot_release(alfq_ctx* c, int id) {
  if (id >= 0 && c->pool[id].refs > 0) c->pool[id].refs -= 1;
}

int* var_ref(alfq_ctx* c, int var) {
  if (var == ALFQ_UDVR) return &c->var_r;
  if (var == ALFQ_UDVL) return &c->var_l;
  if (var >= 1 && var <= c->nstm) return &c->var_st[var];
  return nullptr;
}

void var_bind(alfq_ctx* c, int* v, int slot) {
  c->pool[slot].refs += 1;
  slot_release(c, *v);
  *v = slot;
}

__global__ void udv_identity_kernel(cplx* U, cplx* V, cplx* D, cplx* sc, int n, long long cs) {
  U = chain_ptr(U, cs);
  V = chain_ptr(V, cs);
  D = chain_ptr(D, cs);
  sc = chain_ptr(sc, cs);
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < (long long)n * n) {
    const int r = (int)(idx % n), cc = (int)(idx / n);
    const cplx v = make_double2(r == cc ? 1.0 : 0.0, 0.0);
    U[idx] = v;
    V[idx] = v;
  }
  if (idx < n) D[idx] = make_double2(1.0, 0.0);
  if (idx < 2) sc[idx] = make_double2(1.0, 0.0);
}

// reset: bind var to a fresh identity slot (U = V = 1, D = 1, det U = 1)
int udv_reset_dev(alfq_ctx* c, int* v, char side) {
  const int id = slot_alloc(c);
  if (id < 0) return ST_GENERIC;
  UDVSlot& sl = c->pool[id];
  sl.side = side;
  hipLaunchKernelGGL(udv_identity_kernel, dim3(nblk((long long)c->n * c->n, 256), 1, c->B), dim3(256), 0, c->s,
                     sl.U, sl.V, sl.D, sl.sc, c->n, c->cs);
  HIP_OK(hipGetLastError());
  var_bind(c, v, id);
  slot_release(c, id);
  return ST_NONE;
}

const cplx ONE = {1.0, 0.0}, ZERO = {0.0, 0.0}, MONE = {-1.0, 0.0};

int fill_uniforms(alfq_ctx* c, long long count) {
  // make uniforms [rng_next, rng_next + count) available on the device
  if (c->rng_mode == 1) return ST_NONE;   // the tape is resident
  if (c->rng_mode != 2) { alfq_set_error("no random stream: call alfq_rng_set_tape or alfq_rng_set_philox"); return ST_GENERIC; }
  if (count > c->u_cap) { alfq_set_error("uniform buffer too small for the request"); return ST_GENERIC; }
  c->u_base = c->rng_next;
  c->u_len = count;
  return launch_philox_fill(c->u, c->st, count, c->B, c->cs, c->s);
}

int metropolis_slice(alfq_ctx* c, int slice, int dir) {
  int st;
  ChainArgs a{};
  a.G = c->G; a.n = c->n; a.X = c->X; a.Y = c->Y; a.Kmax = c->K;
  a.f = c->f; a.n_op = c->n_op; a.slice = slice; a.dir = dir;
  a.op_site = c->op_site; a.op_type = c->op_type;
  a.op_myexp = c->myexp; a.op_expalpha = c->expalpha; a.gama_ratio = c->gama_ratio;
  a.n_sun = c->n_sun;
  a.u = c->u; a.st = c->st;
  a.W = c->Wk; a.amap = c->amapk;
  a.cs = c->cs; a.nchain = c->B;
  a.lz = c->lz;
  a.uoff = dir > 0 ? c->uoff_up : c->uoff_dn;
  const int worst = c->n_op / c->K + 1;
  const bool reg_kernel = chain_max_delay(c->n) >= c->K;
  int rounds = worst;
  if (reg_kernel) {
    // the final round flushes in place if the slice needs more, so any count is correct
    rounds = std::max(1, std::min(c->rounds, worst));
  } else {
    hipLaunchKernelGGL(chain_reset_kernel, dim3(1, 1, c->B), dim3(1), 0, c->s, c->st, c->cs);
    HIP_OK(hipGetLastError());
  }
  // X = G(:, A) W: the delayed-update columns of a round from its sites and W = C D^T.  When
  // it runs on the direct-load kernel (few chains), that kernel also copies the rows A of G
  // into Yhat, spread over its grid, in place of the block kernel's single workgroup.
  GemmArgs gx = gemm_args(c->n, c->K, c->K, c->G, c->n, OP_N, true, c->Wk, CHAIN_W_LD, OP_N, true, c->X, c->n,
                          ONE, ZERO);
  gx.amap = c->amapk;
  gx.kdev = &c->st->m;
  gx.nchain = c->B;
  gx.policy = c->gemm_policy;
  a.yhat_external = reg_kernel && gemm_direct_path(gx) ? 1 : 0;
  // few chains: X product and flush in one launch from G into the other G buffer (Gtmp is
  // free during the rounds), then the two swap.  Up to Ndim 256, where it was measured faster:
  // at Ndim 576 each 16 x 16 tile's gather of 32 rows of G costs more than the launch it saves
  const bool fused = a.yhat_external && c->round_fusion && c->n <= 256;
  if (a.yhat_external && !fused) gx.yhat = c->Y;
  for (int r = 0; r < rounds; ++r) {
    a.first_round = (r == 0);
    a.final_round = (r == rounds - 1);
    a.G = c->G;
    if ((st = launch_chain(a, c->s, &c->chain_timer))) return st;
    if (fused) {
      if ((st = round_flush_launch(c->G, c->Gtmp, c->Wk, c->amapk, &c->st->m, c->n, c->K, c->B, c->cs, c->s)))
        return st;
      std::swap(c->G, c->Gtmp);
      continue;
    }
    if (reg_kernel) {
      if ((st = gemm_c(c, gx))) return st;
    }
    // G += X Y^T with the rank read from the chain state
    const long long ftiles = (long long)((c->n + 63) / 64) * ((c->n + 63) / 64) * c->B;
    if (reg_kernel && ftiles >= 256) {
      // a batch with enough 64 x 64 tiles to fill the device: one memory round trip per tile
      if ((st = flush_launch(c->G, c->X, c->Y, &c->st->m, c->n, c->K, c->B, c->cs, c->s))) return st;
    } else {
      GemmArgs g = gemm_args(c->n, c->n, c->K, c->X, c->n, OP_N, true, c->Y, c->n, OP_T, true, c->G, c->n, ONE, ONE);
      g.kdev = &c->st->m;
      if ((st = gemm_c(c, g))) return st;
    }
  }
  c->rng_next += c->per_visit_sum;
  return ST_NONE;
}

// The words the host reads after a sweep, gathered over the chains into one small buffer:
// [3z] the largest accepted count of a slice of chain z since the last read (then cleared
// when clear_acc), [3z + 1] its random-stream status, [3z + 2] its Green-function status.
__global__ void status_gather_kernel(ChainState* st, const DevScalars* sc, int* out, int clear_acc, long long cs) {
  const int z = blockIdx.z;
  st = chain_ptr(st, cs);
  sc = chain_ptr(sc, cs);
  out[3 * z] = st->max_slice_acc;
  out[3 * z + 1] = st->status;
  out[3 * z + 2] = sc->status;
  if (clear_acc) st->max_slice_acc = 0;
}

int gather_status(alfq_ctx* c, bool clear_acc) {
  hipLaunchKernelGGL(status_gather_kernel, dim3(1, 1, c->B), dim3(1), 0, c->s, c->st, c->sc, c->stat_dev,
                     clear_acc ? 1 : 0, c->cs);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(c->stat_host, c->stat_dev, 3 * sizeof(int) * c->B, hipMemcpyDeviceToHost, c->s));
  HIP_OK(hipStreamSynchronize(c->s));
  return ST_NONE;
}

// Terminate_on_error's conditions (Control_PrecisionG: NaN or XMAX > 10; an exhausted
// random stream) of any chain, from the last gather_status.
int status_error(const alfq_ctx* c) {
  for (int z = 0; z < c->B; ++z) {
    const std::string who = c->B > 1 ? "chain " + std::to_string(z) + ": " : "";
    if (c->stat_host[3 * z + 1] == 1) { alfq_set_error(who + "random stream exhausted"); return ST_GENERIC; }
    if (c->stat_host[3 * z + 2] == 1) {
      alfq_set_error(who + "Green function contains NaN, calculation is being aborted!");
      return ST_GENERIC;
    }
    if (c->stat_host[3 * z + 2] == 3) {
      alfq_set_error(who + "XMAX is exceeding the threshold of 10 for G difference! Try with smaller Nwrap or dtau.");
      return ST_UNSTABLE_MATRIX;
    }
  }
  return ST_NONE;
}

// After a sweep: stop at the first failing sweep (the reference terminates at once), and
// choose the number of chain rounds per slice from the largest accepted count of a slice
// seen in the sweep.
int end_of_sweep(alfq_ctx* c) {
  int st;
  if ((st = gather_status(c, true))) return st;
  if ((st = status_error(c))) return st;
  int mx = 0;
  for (int z = 0; z < c->B; ++z) mx = std::max(mx, c->stat_host[3 * z]);
  const int worst = c->n_op / c->K + 1;
  // enough rounds for the busiest slice seen; a slice that needs more is finished by the
  // last round's in-place flush
  const int want = (mx + c->K - 1) / c->K;
  c->rounds = std::max(1, std::min(worst, want));
  return ST_NONE;
}

// a batch whose checkerboard operators are set for some chains only cannot propagate
int hop_ready(const alfq_ctx* c) {
  if (!c->cb_set.empty() && !c->cb_on) {
    alfq_set_error("checkerboard operators are set for some chains of the batch only");
    return ST_HAMILTONIAN;
  }
  return ST_NONE;
}

// flag |= 1 where a chain's words differ from chain 0's (grid y = chains 1 .. B - 1)
__global__ void chains_differ_kernel(const unsigned long long* __restrict__ x, long long words, long long cs_words,
                                     int* flag) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long z = (long long)blockIdx.y + 1;
  if (i < words && x[z * cs_words + i] != x[i]) atomicOr(flag, 1);
}

// The digit planes of exp(-dtau T) and exp(dtau T).  When every chain of the batch holds the
// same two matrices (one lattice, one dtau: the U and beta of the chains may differ), only
// chain 0's are split and the products read them with chain stride 0, so that the four
// operands (4 x 8 n^2 bytes for real T) stay in each XCD's L2 for the whole batch.
int oz_split_hopping(alfq_ctx* c) {
  const int n = c->n;
  const size_t tbytes = (size_t)n * n * (c->t_real ? sizeof(double) : sizeof(cplx));
  bool shared = c->B > 1 && c->cs % 8 == 0;
  if (shared) {
    int* flag = c->stat_dev;            // a word of the status buffer, read back at once
    HIP_OK(hipMemsetAsync(flag, 0, sizeof(int), c->s));
    const long long words = (long long)(tbytes / 8);
    const dim3 grid((unsigned)((words + 255) / 256), (unsigned)(c->B - 1));
    for (const void* m : {c->expT, c->expTinv}) {
      hipLaunchKernelGGL(chains_differ_kernel, grid, dim3(256), 0, c->s, (const unsigned long long*)m, words,
                         c->cs / 8, flag);
      HIP_OK(hipGetLastError());
    }
    int differ = 1;
    HIP_OK(hipMemcpyAsync(&differ, flag, sizeof(int), hipMemcpyDeviceToHost, c->s));
    HIP_OK(hipStreamSynchronize(c->s));
    shared = differ == 0;
  }
  const int nz = shared ? 1 : c->B;
  OzOperand* ops[4] = {&c->oz_e_row, &c->oz_e_col, &c->oz_ei_row, &c->oz_ei_col};
  for (OzOperand* o : ops) o->stride = shared ? 0 : c->oz_cs;
  int st;
  if ((st = oz_split(c->expT, !c->t_real, n, n, n, true, nullptr, c->cs, c->oz_e_row, nz, c->s)) ||
      (st = oz_split(c->expT, !c->t_real, n, n, n, false, nullptr, c->cs, c->oz_e_col, nz, c->s)) ||
      (st = oz_split(c->expTinv, !c->t_real, n, n, n, true, nullptr, c->cs, c->oz_ei_row, nz, c->s)) ||
      (st = oz_split(c->expTinv, !c->t_real, n, n, n, false, nullptr, c->cs, c->oz_ei_col, nz, c->s)))
    return st;
  return ST_NONE;
}

// ALFQ_PREC_INT8_SPLIT: one digit-plane buffer per chain (oz_cs bytes) holding exp(-dtau T)
// and exp(dtau T) split by row (left operands) and by column (right operands), split once
// here since a chain's hopping is fixed, and the planes of the matrix being wrapped.
int oz_setup(alfq_ctx* c) {
  const int n = c->n, pe = c->t_real ? 1 : 3;
  const long long kp = oz_pitch(n);
  size_t off = 0;
  auto take = [&off](size_t bytes) {
    const size_t o = off;
    off = (off + bytes + 255) & ~(size_t)255;
    return o;
  };
  size_t o_d[5], o_e[5];
  for (int q = 0; q < 5; ++q) {
    o_d[q] = take(oz_bytes(n, n, q < 4 ? pe : 2));
    o_e[q] = take((size_t)n * sizeof(int));
  }
  c->oz_cs = (long long)((off + 4095) & ~(size_t)4095);
  if (hipMalloc(&c->oz_buf, (size_t)c->oz_cs * c->B) != hipSuccess) {
    c->oz_buf = nullptr;
    alfq_set_error("precision: hipMalloc of the digit planes");
    return ST_GENERIC;
  }
  OzOperand* ops[5] = {&c->oz_e_row, &c->oz_e_col, &c->oz_ei_row, &c->oz_ei_col, &c->oz_x};
  for (int q = 0; q < 5; ++q) {
    ops[q]->d = (int8_t*)(c->oz_buf + o_d[q]);
    ops[q]->e = (int*)(c->oz_buf + o_e[q]);
    ops[q]->rows = n;
    ops[q]->planes = q < 4 ? pe : 2;
    ops[q]->kp = kp;
    ops[q]->stride = c->oz_cs;
  }
  int st;
  if ((st = oz_split_hopping(c))) {
    hipStreamSynchronize(c->s);
    hipFree(c->oz_buf);
    c->oz_buf = nullptr;
    return st;
  }
  return ST_NONE;
}

// out = rs (E X) cs with E split by row, X (n x n complex) split by column here
int oz_left(alfq_ctx* c, const OzOperand& E, const cplx* X, cplx* out, const cplx* rs, const cplx* cs) {
  const int n = c->n;
  int st = oz_split(X, true, n, n, n, false, nullptr, c->cs, c->oz_x, c->B, c->s);
  return st ? st : oz_gemm(n, n, n, E, c->oz_x, out, n, ONE, ZERO, rs, cs, c->cs, c->B, c->s);
}

// out = rs (X K E) cs with X (n x n complex) split by row here, its columns scaled by the
// diagonal K (kscale, may be null), and E split by column
int oz_right(alfq_ctx* c, const cplx* X, const cplx* kscale, const OzOperand& E, cplx* out, const cplx* rs,
             const cplx* cs) {
  const int n = c->n;
  int st = oz_split(X, true, n, n, n, true, kscale, c->cs, c->oz_x, c->B, c->s);
  return st ? st : oz_gemm(n, n, n, c->oz_x, E, out, n, ONE, ZERO, rs, cs, c->cs, c->B, c->s);
}

// In a batch of uneven Ltrot, operand `side` (1: A, 2: B) of a hopping product of slice
// `slice` (0-based) becomes the identity for the chains that have no such slice.
void hop_mask(const alfq_ctx* c, GemmArgs& g, int slice, int side) {
  if (!c->lz) return;
  g.lz = c->lz;
  g.slice = slice;
  g.hop_side = side;
  g.ident = c->ident;
}

int wrapgrup_dev(alfq_ctx* c, int ntau) {
  int st;
  if ((st = hop_ready(c))) return st;
  const int slice = ntau;   // fields column of time slice NTAU1 = NTAU + 1
  if ((st = launch_dvec(c->f, c->n_op, slice, 1, c->n, c->site_ptr, c->site_ops, c->eexp, c->dv_d, c->dv_dinv,
                        c->dv_dconj, c->B, c->cs, c->s)))
    return st;
  if (c->cb_on) {
    // H = E_1 ... E_N G (Hop_mod_mmthr, bonds Ncheck..1); G = D (H E_N^-1 ... E_1^-1) D^-1
    // (Hop_mod_mmthl_m1, bonds Ncheck..1, then Op_Wrapup type 1 with the old fields)
    if ((st = cb_left(c->H, c->G, c->n, nullptr, nullptr, c->cb, 1, c->B, c->cs, c->s))) return st;
    if ((st = cb_right(c->Gtmp, c->H, c->n, nullptr, nullptr, c->dv_d, c->dv_dinv, c->cb_inv, 1, c->B, c->cs,
                       c->s)))
      return st;
  } else if (c->prec == ALFQ_PREC_INT8_SPLIT) {
    // H = exp(-dtau T) G;  G = D (H exp(dtau T)) D^-1, both on the int8 split products
    if ((st = oz_left(c, c->oz_e_row, c->G, c->H, nullptr, nullptr))) return st;
    if ((st = oz_right(c, c->H, nullptr, c->oz_ei_col, c->Gtmp, c->dv_d, c->dv_dinv))) return st;
  } else {
    // H = exp(-dtau T) G
    GemmArgs g1 = gemm_args(c->n, c->n, c->n, c->expT, c->n, OP_N, !c->t_real, c->G, c->n, OP_N, true, c->H, c->n,
                            ONE, ZERO);
    g1.prec = c->prec;
    hop_mask(c, g1, slice, 1);
    if ((st = gemm_c(c, g1))) return st;
    // G = D (H exp(dtau T)) D^-1  (Op_Wrapup type 1 of every operator, old fields)
    GemmArgs g2 = gemm_args(c->n, c->n, c->n, c->H, c->n, OP_N, true, c->expTinv, c->n, OP_N, !c->t_real, c->Gtmp,
                            c->n, ONE, ZERO);
    g2.rs = c->dv_d;
    g2.cs = c->dv_dinv;
    g2.prec = c->prec;
    hop_mask(c, g2, slice, 2);
    if ((st = gemm_c(c, g2))) return st;
  }
  std::swap(c->G, c->Gtmp);
  if (!c->uniforms_ready && (st = fill_uniforms(c, c->per_visit_sum))) return st;
  return metropolis_slice(c, slice, +1);
}

int wrapgrdo_dev(alfq_ctx* c, int ntau) {
  int st;
  if ((st = hop_ready(c))) return st;
  const int slice = ntau - 1;
  if (!c->uniforms_ready && (st = fill_uniforms(c, c->per_visit_sum))) return st;
  if ((st = metropolis_slice(c, slice, -1))) return st;
  if ((st = launch_dvec(c->f, c->n_op, slice, 1, c->n, c->site_ptr, c->site_ops, c->eexp, c->dv_d, c->dv_dinv,
                        c->dv_dconj, c->B, c->cs, c->s)))
    return st;
  if (c->cb_on) {
    // H = (D^-1 G D) E_1 ... E_N (Op_Wrapdo, then Hop_mod_mmthl, bonds 1..Ncheck);
    // G = E_N^-1 ... E_1^-1 H (Hop_mod_mmthr_m1, bonds 1..Ncheck)
    if ((st = cb_right(c->H, c->G, c->n, c->dv_dinv, c->dv_d, nullptr, nullptr, c->cb, 0, c->B, c->cs, c->s)))
      return st;
    if ((st = cb_left(c->Gtmp, c->H, c->n, nullptr, nullptr, c->cb_inv, 0, c->B, c->cs, c->s))) return st;
  } else if (c->prec == ALFQ_PREC_INT8_SPLIT) {
    // H = D^-1 (G D) exp(-dtau T);  G = exp(dtau T) H, both on the int8 split products
    if ((st = oz_right(c, c->G, c->dv_d, c->oz_e_col, c->H, c->dv_dinv, nullptr))) return st;
    if ((st = oz_left(c, c->oz_ei_row, c->H, c->Gtmp, nullptr, nullptr))) return st;
  } else {
    // H = D^-1 G D exp(-dtau T)   (Op_Wrapdo type 1 with the new fields, then Hop_mod_mmthl)
    GemmArgs g1 = gemm_args(c->n, c->n, c->n, c->G, c->n, OP_N, true, c->expT, c->n, OP_N, !c->t_real, c->H, c->n,
                            ONE, ZERO);
    g1.rs = c->dv_dinv;
    g1.bs = c->dv_d;
    g1.prec = c->prec;
    hop_mask(c, g1, slice, 2);
    if ((st = gemm_c(c, g1))) return st;
    // G = exp(dtau T) H   (Hop_mod_mmthr_m1)
    GemmArgs g2 = gemm_args(c->n, c->n, c->n, c->expTinv, c->n, OP_N, !c->t_real, c->H, c->n, OP_N, true, c->Gtmp,
                            c->n, ONE, ZERO);
    g2.prec = c->prec;
    hop_mask(c, g2, slice, 1);
    if ((st = gemm_c(c, g2))) return st;
  }
  std::swap(c->G, c->Gtmp);
  return ST_NONE;
}

// B-products of slices and the decomposition into a new slot bound to *var.
int wrap_udv_dev(alfq_ctx* c, int* var, int nt_lo, int nt_hi, bool left) {
  int st;
  const int n = c->n;
  const int ns = nt_hi - nt_lo;          // slices nt_lo+1 .. nt_hi (1-based)
  if (ns < 0) { alfq_set_error("wrap: NTAU1 < NTAU"); return ST_GENERIC; }
  if (ns > c->W) { alfq_set_error("wrap: more than Nwrap slices"); return ST_GENERIC; }
  if ((st = hop_ready(c))) return st;
  UDVSlot src = c->pool[*var];
  const cplx* cur = src.U;
  if (ns > 0) {
    if ((st = launch_dvec(c->f, c->n_op, nt_lo, ns, n, c->site_ptr, c->site_ops, c->eexp, c->dv_d, c->dv_dinv,
                          c->dv_dconj, c->B, c->cs, c->s)))
      return st;
  }
  cplx* bufs[2] = {c->U0, c->U1};
  int which = 0;
  for (int q = 0; q < ns; ++q) {
    // right chain: slices ascending, U <- D exp(-dtau T) U (Hop_mod_mmthR, Op_mmultR 'n')
    // left chain:  slices descending, U <- exp(-dtau T) conj(D) U (Op_mmultR 'c', Hop_mod_mmthlc)
    const int sl = left ? (ns - 1 - q) : q;
    if (c->cb_on) {
      // right: D E_1 ... E_N U (bonds Ncheck..1); left: E_N ... E_1 conj(D) U (Hop_mod_mmthlc)
      st = left ? cb_left(bufs[which], cur, n, c->dv_dconj + (long long)sl * n, nullptr, c->cb, 0, c->B, c->cs, c->s)
                : cb_left(bufs[which], cur, n, nullptr, c->dv_d + (long long)sl * n, c->cb, 1, c->B, c->cs, c->s);
      if (st) return st;
    } else {
      GemmArgs g = gemm_args(n, n, n, c->expT, n, OP_N, !c->t_real, cur, n, OP_N, true, bufs[which], n, ONE, ZERO);
      if (left) g.bs = c->dv_dconj + (long long)sl * n;
      else g.rs = c->dv_d + (long long)sl * n;
      hop_mask(c, g, nt_lo + sl, 1);
      if ((st = gemm_c(c, g))) return st;
    }
    cur = bufs[which];
    which ^= 1;
  }
  const int id = slot_alloc(c);
  if (id < 0) return ST_GENERIC;
  UDVSlot& out = c->pool[id];
  out.side = left ? 'L' : 'R';
  if ((st = udv_decompose_dev(c->qr, cur, src.D, src.V, left ? 1 : 0, out.U, out.D, out.V, out.sc, out.sc + 1,
                              c->s)))
    return st;
  var_bind(c, var, id);
  slot_release(c, id);
  return ST_NONE;
}

int cgr_dev(alfq_ctx* c, int nvar, int slot_r, int slot_l, int check) {
  int st;
  const int n = c->n;
  const UDVSlot& R = c->pool[slot_r];
  const UDVSlot& Lf = c->pool[slot_l];
  cplx* rbinv = c->cscal;
  cplx* lbinv = c->cscal + n;
  cplx* rsm = c->cscal + 2 * n;
  cplx* lsm = c->cscal + 3 * n;
  hipLaunchKernelGGL(cgr_scales_kernel, dim3(nblk(n, 256), 1, c->B), dim3(256), 0, c->s, R.D, Lf.D, n, c->cscal,
                     c->cs);
  HIP_OK(hipGetLastError());
  // TP = DR+^-1 (UR^H UL) DL+^-1
  GemmArgs g1 = gemm_args(n, n, n, R.U, n, OP_C, true, Lf.U, n, OP_N, true, c->TP, n, ONE, ZERO);
  g1.rs = rbinv;
  g1.cs = lbinv;
  if ((st = gemm_c(c, g1))) return st;
  // TP += DR- (VR VL) DL-
  GemmArgs g2 = gemm_args(n, n, n, R.V, n, OP_N, true, Lf.V, n, OP_N, true, c->TP, n, ONE, ONE);
  g2.rs = rsm;
  g2.cs = lsm;
  if ((st = gemm_c(c, g2))) return st;
  if ((st = qr_factor_cgr(c->qr, c->TP, nvar, R.sc + 1, Lf.sc + 1, c->Qc, c->DUP, &c->sc->z_cgr, c->s))) return st;
  hipLaunchKernelGGL(recip_real_kernel, dim3(nblk(n, 256), 1, c->B), dim3(256), 0, c->s, c->DUP, n, c->DUPinv,
                     c->cs);
  HIP_OK(hipGetLastError());
  if ((st = tri_inverse_dev(c->qr, c->qr.Rt, c->qr.Rinv, c->s))) return st;
  if (nvar == 1) {
    // Xc = DUP^-1 Q^H DR+^-1 UR^H
    GemmArgs a = gemm_args(n, n, n, c->Qc, n, OP_C, true, R.U, n, OP_C, true, c->Xc, n, ONE, ZERO);
    a.bs = rbinv;
    a.rs = c->DUPinv;
    if ((st = gemm_c(c, a))) return st;
    // Yc = Rt^-1 Xc
    GemmArgs b = gemm_args(n, n, n, c->qr.Rinv, n, OP_N, true, c->Xc, n, OP_N, true, c->Yc, n, ONE, ZERO);
    if ((st = gemm_c(c, b))) return st;
    // G = UL DL+^-1 P Yc
    GemmArgs d = gemm_args(n, n, n, Lf.U, n, OP_N, true, c->Yc, n, OP_N, true, c->Gtmp, n, ONE, ZERO);
    d.bmap = c->qr.pinv;
    d.bs = lbinv;
    if ((st = gemm_c(c, d))) return st;
  } else {
    // Xc = UL DL+^-1 Q DUP^-1
    GemmArgs a = gemm_args(n, n, n, Lf.U, n, OP_N, true, c->Qc, n, OP_N, true, c->Xc, n, ONE, ZERO);
    a.bs = lbinv;
    a.cs = c->DUPinv;
    if ((st = gemm_c(c, a))) return st;
    // Yc = Xc Rt^-H
    GemmArgs b = gemm_args(n, n, n, c->Xc, n, OP_N, true, c->qr.Rinv, n, OP_C, true, c->Yc, n, ONE, ZERO);
    if ((st = gemm_c(c, b))) return st;
    // G = Yc P^T DR+^-1 UR^H
    GemmArgs d = gemm_args(n, n, n, c->Yc, n, OP_N, true, R.U, n, OP_C, true, c->Gtmp, n, ONE, ZERO);
    d.amap = c->qr.pinv;
    d.bs = rbinv;
    if ((st = gemm_c(c, d))) return st;
  }
  if ((st = launch_opphase(c->f, (long long)c->n_op * c->L, c->n_op, c->phasefac, &c->sc->opph, c->B, c->cs, c->s)))
    return st;
  hipLaunchKernelGGL(stabilise_kernel, dim3(1, 1, c->B), dim3(1024), 0, c->s, c->Gtmp, c->G, n, check, c->n_sun,
                     c->sc, c->st, c->log, c->cs);
  HIP_OK(hipGetLastError());
  std::swap(c->G, c->Gtmp);
  return ST_NONE;
}

int sweep_body(alfq_ctx* c);

// CGR2_2 (cgr2_2_mod.F90:196-428) from the right state R = B(tau, 0) and the left state
// Lf = B(beta, tau)^H into G00, G0T, GT0, GTT, with Control_Precision_tau against the
// propagated values.  With H the scale-split matrix and H^H P = Q D R~ its QR,
// O^-1 = Q D^-1 R~^-H P^T diag(F1, F2) (solve_extended_System, cgr2_2_mod.F90:130-194).
int cgr2_2_dev(alfq_ctx* c, const UDVSlot& R, const UDVSlot& Lf, cplx* G00, cplx* G0T, cplx* GT0, cplx* GTT) {
  int st;
  auto& t = c->tau;
  const int n = c->n, m = 2 * n;
  // V_L^-1 = P R~^-1 D^-1 Q^H from V_L P = Q D R~ (the equal-time CGR's QR workspace)
  if ((st = qr_factor_cgr(c->qr, Lf.V, 1, R.sc + 1, Lf.sc + 1, t.Qv, t.DUPv, t.zs, c->s))) return st;
  hipLaunchKernelGGL(recip_real_kernel, dim3(nblk(n, 256), 1, c->B), dim3(256), 0, c->s, t.DUPv, n, t.DUPvinv,
                     c->cs);
  HIP_OK(hipGetLastError());
  if ((st = tri_inverse_dev(c->qr, c->qr.Rt, c->qr.Rinv, c->s))) return st;
  GemmArgs gx = gemm_args(n, n, n, c->qr.Rinv, n, OP_N, true, t.Qv, n, OP_C, true, t.Xv, n, ONE, ZERO);
  gx.bs = t.DUPvinv;
  if ((st = gemm_c(c, gx))) return st;
  if ((st = cgr22_build(t.Xv, c->qr.pinv, Lf.U, Lf.D, R.U, R.D, R.V, t.H2, t.F1, t.F2, t.flag, n, c->B, c->cs,
                        c->s)))
    return st;
  if ((st = qr_factor_cgr(t.w2, t.H2, 2, R.sc + 1, Lf.sc + 1, t.Q2, t.DUP2, t.zs, c->s))) return st;
  hipLaunchKernelGGL(recip_real_kernel, dim3(nblk(m, 256), 1, c->B), dim3(256), 0, c->s, t.DUP2, m, t.DUP2inv,
                     c->cs);
  HIP_OK(hipGetLastError());
  if ((st = tri_inverse_dev(t.w2, t.w2.Rt, t.w2.Rinv, c->s))) return st;
  // Y = Q D^-1 R~^-H
  GemmArgs gy = gemm_args(m, m, m, t.Q2, m, OP_N, true, t.w2.Rinv, m, OP_C, true, t.Y2, m, ONE, ZERO);
  gy.bs = t.DUP2inv;
  if ((st = gemm_c(c, gy))) return st;
  // O^-1 = (Y P^T) diag(F1, F2), one column block at a time
  GemmArgs g1 = gemm_args(m, n, n, t.Y2, m, OP_N, true, t.F1, n, OP_N, true, t.O2, m, ONE, ZERO);
  g1.amap = t.w2.pinv;
  if ((st = gemm_c(c, g1))) return st;
  GemmArgs g2 = gemm_args(m, n, n, t.Y2, m, OP_N, true, t.F2, n, OP_N, true, t.O2 + (long long)m * n, m, ONE, ZERO);
  g2.amap = t.w2.pinv + n;
  if ((st = gemm_c(c, g2))) return st;
  return cgr22_extract(t.O2, t.flag, c->G, G00, G0T, GT0, GTT, t.part, t.ctl, n, c->B, c->cs, c->s);
}

// TAU_M (tau_m_mod.F90:56-195), finite temperature, after a sweep: G(tau, 0), G(0, tau),
// G(tau, tau), G(0, 0) for tau = 0..Ltrot, measured at every tau.
int tau_m_dev(alfq_ctx* c) {
  int st;
  auto& t = c->tau;
  const int n = c->n, L = c->L;
  if (!t.mem || !c->obs.buf) { alfq_set_error("TAU_M: call alfq_tau_enable first"); return ST_GENERIC; }
  if ((st = hop_ready(c))) return st;
  cplx *GT0 = t.b[0], *G0T = t.b[1], *GTT = t.b[2], *G00 = t.b[3], *T1 = t.b[4];
  auto observe = [&](int nt) {
    return tau_measure(c->obs, GT0, G0T, G00, GTT, c->st, t.acc, t.acc_cs, t.off_den0, t.off_latt, nt, L + 1, c->B,
                       c->cs, c->s);
  };
  if ((st = tau_init(c->G, G00, GT0, GTT, G0T, n, c->B, c->cs, c->s))) return st;
  if ((st = observe(0))) return st;
  if ((st = udv_reset_dev(c, &c->var_tr, 'R'))) return st;
  int nst = 1;
  for (int nt = 0; nt < L; ++nt) {
    const int nt1 = nt + 1;
    if ((st = launch_dvec(c->f, c->n_op, nt, 1, n, c->site_ptr, c->site_ops, c->eexp, c->dv_d, c->dv_dinv,
                          c->dv_dconj, c->B, c->cs, c->s)))
      return st;
    if (c->cb_on) {
      // PROPR: G(tau,0) <- D P G(tau,0); PROPRM1: G(0,tau) <- G(0,tau) P^-1 D^-1;
      // G(tau,tau) <- D P G(tau,tau) P^-1 D^-1
      if ((st = cb_left(T1, GT0, n, nullptr, c->dv_d, c->cb, 1, c->B, c->cs, c->s))) return st;
      std::swap(GT0, T1);
      if ((st = cb_right(T1, G0T, n, nullptr, nullptr, nullptr, c->dv_dinv, c->cb_inv, 1, c->B, c->cs, c->s)))
        return st;
      std::swap(G0T, T1);
      if ((st = cb_left(T1, GTT, n, nullptr, nullptr, c->cb, 1, c->B, c->cs, c->s))) return st;
      if ((st = cb_right(GTT, T1, n, nullptr, nullptr, c->dv_d, c->dv_dinv, c->cb_inv, 1, c->B, c->cs, c->s)))
        return st;
    } else {
      GemmArgs a = gemm_args(n, n, n, c->expT, n, OP_N, !c->t_real, GT0, n, OP_N, true, T1, n, ONE, ZERO);
      a.rs = c->dv_d;
      if ((st = gemm_c(c, a))) return st;
      std::swap(GT0, T1);
      GemmArgs b = gemm_args(n, n, n, G0T, n, OP_N, true, c->expTinv, n, OP_N, !c->t_real, T1, n, ONE, ZERO);
      b.cs = c->dv_dinv;
      if ((st = gemm_c(c, b))) return st;
      std::swap(G0T, T1);
      GemmArgs d1 = gemm_args(n, n, n, c->expT, n, OP_N, !c->t_real, GTT, n, OP_N, true, T1, n, ONE, ZERO);
      if ((st = gemm_c(c, d1))) return st;
      GemmArgs d2 = gemm_args(n, n, n, T1, n, OP_N, true, c->expTinv, n, OP_N, !c->t_real, GTT, n, ONE, ZERO);
      d2.rs = c->dv_d;
      d2.cs = c->dv_dinv;
      if ((st = gemm_c(c, d2))) return st;
    }
    if ((st = observe(nt1))) return st;
    if (nst <= c->nstm && c->stab[nst] == nt1) {
      if ((st = wrap_udv_dev(c, &c->var_tr, c->stab[nst - 1], nt1, false))) return st;
      if ((st = cgr2_2_dev(c, c->pool[c->var_tr], c->pool[c->var_st[nst]], G00, G0T, GT0, GTT))) return st;
      ++nst;
    }
  }
  slot_release(c, c->var_tr);
  c->var_tr = -1;
  return ST_NONE;
}

int init_dev(alfq_ctx* c) {
  int st;
  const auto& s = c->stab;
  if ((st = udv_reset_dev(c, &c->var_st[c->nstm], 'L'))) return st;
  if ((st = udv_reset_dev(c, &c->var_l, 'L'))) return st;
  if ((st = udv_reset_dev(c, &c->var_r, 'R'))) return st;
  for (int nst = c->nstm - 1; nst >= 1; --nst) {
    if ((st = wrap_udv_dev(c, &c->var_l, s[nst], s[nst + 1], true))) return st;
    var_bind(c, &c->var_st[nst], c->var_l);
  }
  if ((st = wrap_udv_dev(c, &c->var_l, 0, s[1], true))) return st;
  return cgr_dev(c, 1, c->var_r, c->var_l, 0);
}

int sweep_dev(alfq_ctx* c) {
  int st;
  const auto& s = c->stab;
  const int L = c->L;
  if ((st = fill_uniforms(c, 2LL * L * c->per_visit_sum))) return st;
  c->uniforms_ready = true;
  st = sweep_body(c);
  c->uniforms_ready = false;
  return st;
}

// Hop_mod_Symm (Hop_mod.F90:279-299): out = exp(-dtau T/2) G exp(dtau T/2), dense or bond by
// bond (for bonds Ncheck..1: E_b^1/2 (.) E_b^-1/2); tmp is scratch.
int hop_symm_dev(alfq_ctx* c, const cplx* G, cplx* tmp, cplx* out) {
  int st;
  const int n = c->n;
  if (c->cb_on) {
    if (!c->cb_symm) {
      alfq_set_error("Symm with a checkerboard Op_T needs the bond half steps (alfq_checkerboard_desc)");
      return ST_HAMILTONIAN;
    }
    if ((st = cb_left(tmp, G, n, nullptr, nullptr, c->cb_half, 1, c->B, c->cs, c->s))) return st;
    return cb_right(out, tmp, n, nullptr, nullptr, nullptr, nullptr, c->cb_half_inv, 1, c->B, c->cs, c->s);
  }
  GemmArgs a = gemm_args(n, n, n, c->expTh, n, OP_N, !c->t_real, G, n, OP_N, true, tmp, n, ONE, ZERO);
  if ((st = gemm_c(c, a))) return st;
  GemmArgs b = gemm_args(n, n, n, tmp, n, OP_N, true, c->expThinv, n, OP_N, !c->t_real, out, n, ONE, ZERO);
  return gemm_c(c, b);
}

// ham%Obser on slice NTAU1 when measurements are on and NTAU1 is in LOBS_ST..LOBS_EN, on
// Hop_mod_Symm(GR) with Symm = .T. (main.F90:757-772, 789-802).  H and Gtmp are free
// between the wraps and the stabilisation, so they hold the transformed G.
int measure_dev(alfq_ctx* c, int ntau1) {
  if (!c->obs_on || ntau1 < c->lobs_st || ntau1 > c->lobs_en) return ST_NONE;
  const cplx* g = c->G;
  if (c->symm) {
    int st;
    if ((st = hop_symm_dev(c, c->G, c->H, c->Gtmp))) return st;
    g = c->Gtmp;
  }
  return obs_measure(c->obs, g, c->st, c->B, c->cs, c->obs_cs, c->s);
}

int sweep_body(alfq_ctx* c) {
  int st;
  const auto& s = c->stab;
  const int L = c->L;
  if ((st = udv_reset_dev(c, &c->var_r, 'R'))) return st;
  int nst = 1;
  for (int ntau = 0; ntau < L; ++ntau) {
    const int ntau1 = ntau + 1;
    if ((st = wrapgrup_dev(c, ntau))) return st;
    if (ntau1 == s[nst]) {
      if ((st = wrap_udv_dev(c, &c->var_r, s[nst - 1], ntau1, false))) return st;
      var_bind(c, &c->var_l, c->var_st[nst]);
      var_bind(c, &c->var_st[nst], c->var_r);
      const int nvar = (ntau1 > L / 2) ? 2 : 1;
      if ((st = cgr_dev(c, nvar, c->var_r, c->var_l, 1))) return st;
      ++nst;
    }
    if ((st = measure_dev(c, ntau1))) return st;          // main.F90:757-772
  }
  if ((st = udv_reset_dev(c, &c->var_l, 'L'))) return st;
  nst = c->nstm - 1;
  for (int ntau = L; ntau >= 1; --ntau) {
    const int ntau1 = ntau - 1;
    if ((st = wrapgrdo_dev(c, ntau))) return st;
    if ((st = measure_dev(c, ntau1))) return st;          // main.F90:789-802, before stabilising
    if (nst >= 0 && s[nst] == ntau1 && ntau1 != 0) {
      if ((st = wrap_udv_dev(c, &c->var_l, ntau1, s[nst + 1], true))) return st;
      var_bind(c, &c->var_r, c->var_st[nst]);
      var_bind(c, &c->var_st[nst], c->var_l);
      const int nvar = (ntau1 > L / 2) ? 2 : 1;
      if ((st = cgr_dev(c, nvar, c->var_r, c->var_l, 1))) return st;
      --nst;
    }
  }
  if ((st = wrap_udv_dev(c, &c->var_l, s[0], s[1], true))) return st;
  if ((st = udv_reset_dev(c, &c->var_r, 'R'))) return st;
  if ((st = cgr_dev(c, 1, c->var_r, c->var_l, 1))) return st;
  if ((st = udv_reset_dev(c, &c->var_st[c->nstm], 'L'))) return st;
  if (c->tau_on && (st = tau_m_dev(c))) return st;        // Ltau = 1 (main.F90:874-876)
  return ST_NONE;
}

alfq_ctx::HostState host_state(const alfq_ctx* c) {
  alfq_ctx::HostState h;
  h.G = c->G; h.Gtmp = c->Gtmp; h.u = c->u;
  h.var_r = c->var_r; h.var_l = c->var_l; h.rounds = c->rounds; h.K = c->K; h.rng_mode = c->rng_mode;
  h.prec = c->prec;
  h.gemm_policy = c->gemm_policy;
  h.obs_on = c->obs_on;
  h.cb_on = c->cb_on;
  h.tau_on = c->tau_on;
  h.round_fusion = c->round_fusion;
  h.var_st = c->var_st;
  h.refs.reserve(c->pool.size());
  for (const auto& sl : c->pool) h.refs.push_back(sl.refs);
  return h;
}

void set_host_state(alfq_ctx* c, const alfq_ctx::HostState& h) {
  c->G = h.G; c->Gtmp = h.Gtmp;
  c->var_r = h.var_r; c->var_l = h.var_l;
  c->var_st = h.var_st;
  for (size_t i = 0; i < c->pool.size() && i < h.refs.size(); ++i) c->pool[i].refs = h.refs[i];
}

void drop_graphs(alfq_ctx* c) {
  for (auto& g : c->graphs) {
    hipGraphExecDestroy(g.exec);
    hipGraphDestroy(g.graph);
  }
  c->graphs.clear();
}

// One sweep.  After a first eager sweep (which sizes every buffer), a sweep is captured
// into a HIP graph the first time its starting host state (buffer parity, UDV slots,
// rounds, delay, random stream) is seen and replayed with one launch afterwards; the
// host bookkeeping then jumps to the state the captured sweep ended in.  Anything that
// fails during capture falls back to eager launches for the rest of the context's life.
int sweep_graph(alfq_ctx* c) {
  if (!c->graphs_enabled || c->chain_timer.on || c->eager_sweeps < 1) {
    const int st = sweep_dev(c);
    if (st == ST_NONE) c->eager_sweeps += 1;
    return st;
  }
  const alfq_ctx::HostState s0 = host_state(c);
  for (const auto& g : c->graphs) {
    if (g.start == s0) {
      HIP_OK(hipGraphLaunch(g.exec, c->s));
      set_host_state(c, g.end);
      c->rng_next += g.rng_advance;
      return ST_NONE;
    }
  }
  if (c->graphs.size() >= 8) drop_graphs(c);
  const unsigned long long rn0 = c->rng_next;
  std::unique_lock<std::recursive_mutex> capture_lock(null_stream_mutex());
  HIP_OK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  const int st = sweep_dev(c);
  c->capturing = false;
  hipGraph_t graph = nullptr;
  const hipError_t ee = hipStreamEndCapture(c->s, &graph);
  hipGraphExec_t exec = nullptr;
  bool ok = st == ST_NONE && ee == hipSuccess && graph != nullptr;
  if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
  capture_lock.unlock();
  if (!ok) {
    if (graph) hipGraphDestroy(graph);
    (void)hipGetLastError();
    set_host_state(c, s0);
    c->rng_next = rn0;
    c->graphs_enabled = false;
    return sweep_dev(c);
  }
  c->graphs.push_back({s0, host_state(c), c->rng_next - rn0, graph, exec});
  HIP_OK(hipGraphLaunch(exec, c->s));
  return ST_NONE;
}

// Ltrot of chain z (the batch's Ltrot unless the batch is uneven)
int chain_ltrot(const alfq_ctx* c, int z) { return c->h_lz.empty() ? c->L : c->h_lz[z]; }

// chain z's copy of a per-chain buffer
template <class T>
T* at_chain(const alfq_ctx* c, T* p, int z) {
  return (T*)((char*)p + (long long)z * c->cs);
}
template <class T>
T* at(const alfq_ctx* c, T* p) {
  return at_chain(c, p, c->cur);
}

int check_status(alfq_ctx* c) {
  int st;
  if ((st = gather_status(c, false))) return st;
  return status_error(c);
}

double phi_st(int s, int typ) {
  if (typ == 1) return (double)s;
  const double s6 = std::sqrt(6.0);
  const double a = std::sqrt(2.0 * (3.0 - s6)), b = std::sqrt(2.0 * (3.0 + s6));
  switch (s) {
    case -2: return -b;
    case -1: return -a;
    case 1: return a;
    default: return b;
  }
}
double gama_st(int s, int typ) {
  if (typ == 1) return 1.0;
  const double s6 = std::sqrt(6.0);
  return (s == 1 || s == -1) ? 1.0 + s6 / 3.0 : 1.0 - s6 / 3.0;
}

}  // namespace

// ============================================================================= C ABI

extern "C" {

const char* alfq_last_error(void) { return g_last_error.c_str(); }
const char* alfq_version(void) { return "alfq 0.1 gfx950 fp64"; }

}  // extern "C"

namespace {

// Checks one descriptor on its own and against the batch's first chain (shapes, operator
// sites and types must agree so that the chains advance in lockstep).
int check_desc(const alfq_model_desc* d, const alfq_model_desc* d0) {
  if (d->n_fl != 1) { alfq_set_error("N_FL must be 1"); return ST_HAMILTONIAN; }
  if (d->ndim < 1 || d->ltrot < 1 || d->nwrap < 1 || d->n_op < 0 || d->n_sun < 1) {
    alfq_set_error("invalid model sizes");
    return ST_HAMILTONIAN;
  }
  for (int i = 0; i < d->n_op; ++i) {
    if (d->op_P[i] < 1 || d->op_P[i] > d->ndim) { alfq_set_error("operator site out of range"); return ST_HAMILTONIAN; }
    if (d->op_type[i] != 1 && d->op_type[i] != 2) {
      alfq_set_error("only type-1 and type-2 (discrete) fields are supported");
      return ST_HAMILTONIAN;
    }
  }
  if ((d->expT_half == nullptr) != (d->expT_half_inv == nullptr)) {
    alfq_set_error("Symm: give both half-step exponentials or neither");
    return ST_HAMILTONIAN;
  }
  if (d0 && d != d0) {
    if ((d->expT_half == nullptr) != (d0->expT_half == nullptr)) {
      alfq_set_error("batch: the chains differ in Symm");
      return ST_HAMILTONIAN;
    }
    if (d->ndim != d0->ndim || d->nwrap != d0->nwrap || d->n_sun != d0->n_sun || d->n_op != d0->n_op ||
        d->t_is_real != d0->t_is_real) {
      alfq_set_error("batch: the chains differ in Ndim, Nwrap, N_SUN, the operator count or the type of T");
      return ST_HAMILTONIAN;
    }
    if (d->ltrot != d0->ltrot && d->expT_half != nullptr) {
      alfq_set_error("batch: chains of uneven Ltrot cannot use Symm");
      return ST_HAMILTONIAN;
    }
    for (int i = 0; i < d->n_op; ++i)
      if (d->op_P[i] != d0->op_P[i] || d->op_type[i] != d0->op_type[i]) {
        alfq_set_error("batch: the chains' operators act on different sites or have different field types");
        return ST_HAMILTONIAN;
      }
  }
  return ST_NONE;
}

// Per-operator tables of one chain's couplings: E_exp (5 per op), Z1-exponentials and
// Gama ratios (25 per op: old x new field), Op_phase factors (5 per op).
void coupling_tables(const alfq_model_desc* d, std::vector<cplx>& eexp, std::vector<cplx>& myexp,
                     std::vector<cplx>& expal, std::vector<cplx>& phf, std::vector<double>& gr) {
  const int no = d->n_op;
  eexp.assign(5 * no, cplx{}); myexp.assign(25 * no, cplx{}); expal.assign(25 * no, cplx{});
  phf.assign(5 * no, cplx{}); gr.assign(25 * no, 0.0);
  for (int i = 0; i < no; ++i) {
    const std::complex<double> g(d->op_g[2 * i], d->op_g[2 * i + 1]);
    const std::complex<double> al(d->op_alpha[2 * i], d->op_alpha[2 * i + 1]);
    const double E = d->op_E[i];
    const int T = d->op_type[i];
    for (int sp = -2; sp <= 2; ++sp) {
      std::complex<double> e(1.0, 0.0);
      if (sp != 0 && !(T == 1 && (sp == 2 || sp == -2))) {
        // E_exp(I) = exp(g E phi(I)), E_exp(-I) = 1/E_exp(I)   (Operator_mod.F90:436-466)
        const std::complex<double> ep = std::exp(g * E * phi_st(std::abs(sp), T));
        e = sp > 0 ? ep : 1.0 / ep;
      }
      eexp[5 * i + sp + 2] = make_double2(e.real(), e.imag());
      std::complex<double> pf(1.0, 0.0);
      if (sp != 0 && !(T == 1 && (sp == 2 || sp == -2))) {
        const double ang = std::imag(g * al * phi_st(sp, T));
        pf = std::complex<double>(std::cos(ang), std::sin(ang));
      }
      phf[5 * i + sp + 2] = make_double2(pf.real(), pf.imag());
    }
    for (int so = -2; so <= 2; ++so) {
      for (int sn = -2; sn <= 2; ++sn) {
        const int ti = 25 * i + (so + 2) * 5 + (sn + 2);
        std::complex<double> me(1.0, 0.0), ea(1.0, 0.0);
        double gra = 1.0;
        const bool valid = so != 0 && sn != 0 && !(T == 1 && (std::abs(so) == 2 || std::abs(sn) == 2));
        if (valid) {
          // Z1 = g (phi_new - phi_old); myexp = exp(Z1 E); exp(Z1 alpha)   (upgrade_mod.F90:160-180)
          const std::complex<double> z1 = g * (phi_st(sn, T) - phi_st(so, T));
          me = std::exp(z1 * E);
          ea = std::exp(z1 * al);
          gra = gama_st(sn, T) / gama_st(so, T);
        }
        myexp[ti] = make_double2(me.real(), me.imag());
        expal[ti] = make_double2(ea.real(), ea.imag());
        gr[ti] = gra;
      }
    }
  }
}

// Places every per-chain buffer of c in one chain's arena (sizes it when ar.base is null).
void place_chain(alfq_ctx* c, ArenaLayout& ar, int nslot) {
  const int n = c->n, no = std::max(c->n_op, 1);
  const size_t nn = (size_t)n * n;
  const size_t tsz = c->t_real ? sizeof(double) : sizeof(cplx);
  c->expT = ar.take<char>(nn * tsz);
  c->expTinv = ar.take<char>(nn * tsz);
  if (c->symm) {
    c->expTh = ar.take<char>(nn * tsz);
    c->expThinv = ar.take<char>(nn * tsz);
  }
  c->eexp = ar.take<cplx>(5 * (size_t)no);
  c->myexp = ar.take<cplx>(25 * (size_t)no);
  c->expalpha = ar.take<cplx>(25 * (size_t)no);
  c->phasefac = ar.take<cplx>(5 * (size_t)no);
  c->gama_ratio = ar.take<double>(25 * (size_t)no);
  c->f = ar.take<int8_t>((size_t)no * c->L);
  c->G = ar.take<cplx>(nn);
  c->Gtmp = ar.take<cplx>(nn);
  c->H = ar.take<cplx>(nn);
  c->X = ar.take<cplx>((size_t)n * CHAIN_KMAX);
  c->Y = ar.take<cplx>((size_t)n * CHAIN_KMAX);
  c->Wk = ar.take<cplx>((size_t)CHAIN_KMAX * CHAIN_KMAX);
  c->amapk = ar.take<int>(CHAIN_KMAX);
  c->st = ar.take<ChainState>(1);
  c->sc = ar.take<DevScalars>(1);
  c->log = ar.take<double>(3 * (size_t)LOG_CAP);
  c->dv_d = ar.take<cplx>((size_t)n * c->W);
  c->dv_dinv = ar.take<cplx>((size_t)n * c->W);
  c->dv_dconj = ar.take<cplx>((size_t)n * c->W);
  c->U0 = ar.take<cplx>(nn);
  c->U1 = ar.take<cplx>(nn);
  c->TP = ar.take<cplx>(nn);
  c->Qc = ar.take<cplx>(nn);
  c->DUP = ar.take<cplx>(n);
  c->DUPinv = ar.take<cplx>(n);
  c->Xc = ar.take<cplx>(nn);
  c->Yc = ar.take<cplx>(nn);
  c->cscal = ar.take<cplx>(4 * (size_t)n);
  c->u_arena = ar.take<double>((size_t)std::max<long long>(c->u_cap, 1));
  qr_work_place(c->qr, n, ar);
  c->pool.assign(nslot, UDVSlot{});
  for (auto& sl : c->pool) {
    sl.U = ar.take<cplx>(nn);
    sl.V = ar.take<cplx>(nn);
    sl.D = ar.take<cplx>(n);
    sl.sc = ar.take<cplx>(2);
  }
}

template <class T>
hipError_t put_chain(alfq_ctx* c, int z, T* dst, const T* src, size_t count) {
  return hipMemcpy((char*)dst + (long long)z * c->cs, src, count * sizeof(T), hipMemcpyHostToDevice);
}

// Bytes of one chain's arena for the model d (what a batch allocates per chain).
size_t chain_arena_bytes(const alfq_model_desc* d) {
  alfq_ctx t;
  t.n = d->ndim; t.L = d->ltrot; t.W = d->nwrap; t.n_op = d->n_op; t.n_sun = d->n_sun;
  t.t_real = d->t_is_real != 0;
  t.symm = d->expT_half != nullptr;
  t.nstm = (t.L % t.W == 0) ? t.L / t.W : t.L / t.W + 1;
  long long per = 0;
  for (int i = 0; i < d->n_op; ++i) per += (d->op_type[i] == 2) ? 3 : 2;
  t.u_cap = 2LL * t.L * per;
  ArenaLayout sizing;
  place_chain(&t, sizing, t.nstm + 6);
  return (sizing.off + 4095) & ~(size_t)4095;
}

int create_impl(const alfq_model_desc* descs, int nchain, alfq_ctx** out) {
  NULL_STREAM_LOCK;
  if (!descs || !out || nchain < 1) { alfq_set_error("null argument or empty batch"); return ST_GENERIC; }
  *out = nullptr;
  int st;
  for (int z = 0; z < nchain; ++z)
    if ((st = check_desc(&descs[z], &descs[0]))) return st;
  const alfq_model_desc* d = &descs[0];
  alfq_ctx* c = new alfq_ctx();
  c->B = nchain;
  c->n = d->ndim; c->L = d->ltrot; c->W = d->nwrap; c->n_op = d->n_op; c->n_sun = d->n_sun;
  for (int z = 1; z < nchain; ++z) c->L = std::max(c->L, (int)descs[z].ltrot);
  for (int z = 0; z < nchain; ++z)
    if (descs[z].ltrot != c->L) {
      c->h_lz.resize(nchain);
      for (int y = 0; y < nchain; ++y) c->h_lz[y] = descs[y].ltrot;
      break;
    }
  c->t_real = d->t_is_real != 0;
  c->symm = d->expT_half != nullptr;
  c->nstm = (c->L % c->W == 0) ? c->L / c->W : c->L / c->W + 1;
  c->stab.assign(c->nstm + 1, 0);
  for (int k = 1; k < c->nstm; ++k) c->stab[k] = c->W * k;
  c->stab[c->nstm] = c->L;
  c->K = std::max(1, std::min(chain_max_delay(d->ndim) > 0 ? chain_max_delay(d->ndim) : 32, std::max(1, d->n_op)));
  const int n = c->n, no = c->n_op;
  auto fail = [&](const char* what) {
    alfq_set_error(std::string("alfq_ctx_create: ") + what);
    alfq_ctx_destroy(c);
    return ST_GENERIC;
  };
  if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) return fail("stream");
  // shared operator structure: sites, types, per-site operator lists, uniform offsets
  std::vector<int> site(no), typ(no), sptr(n + 1, 0), sops(no);
  for (int i = 0; i < no; ++i) {
    site[i] = d->op_P[i] - 1;
    typ[i] = d->op_type[i];
    sptr[site[i] + 1] += 1;
  }
  for (int i = 0; i < n; ++i) sptr[i + 1] += sptr[i];
  {
    std::vector<int> fillp(sptr.begin(), sptr.end() - 1);
    for (int i = 0; i < no; ++i) sops[fillp[site[i]]++] = i;
  }
  c->h_op_type = typ;
  c->h_op_site = site;
  c->per_visit_sum = 0;
  for (int i = 0; i < no; ++i) c->per_visit_sum += (typ[i] == 2) ? 3 : 2;
  c->u_cap = c->u_arena_cap = 2LL * c->L * c->per_visit_sum;
  std::vector<int> up(no + 1, 0), dn(no + 1, 0);
  for (int v = 0; v < no; ++v) {
    up[v + 1] = up[v] + ((typ[v] == 2) ? 3 : 2);
    dn[v + 1] = dn[v] + ((typ[no - 1 - v] == 2) ? 3 : 2);
  }
  const int nop1 = std::max(no, 1);
  if (hipMalloc(&c->uoff_up, (no + 1) * sizeof(int)) || hipMalloc(&c->uoff_dn, (no + 1) * sizeof(int)) ||
      hipMalloc(&c->op_site, nop1 * sizeof(int)) || hipMalloc(&c->op_type, nop1 * sizeof(int)) ||
      hipMalloc(&c->site_ptr, (n + 1) * sizeof(int)) || hipMalloc(&c->site_ops, nop1 * sizeof(int)))
    return fail("hipMalloc tables");
  if (hipMemcpy(c->uoff_up, up.data(), (no + 1) * sizeof(int), hipMemcpyHostToDevice) ||
      hipMemcpy(c->uoff_dn, dn.data(), (no + 1) * sizeof(int), hipMemcpyHostToDevice) ||
      hipMemcpy(c->site_ptr, sptr.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice))
    return fail("copy tables");
  if (no > 0 && (hipMemcpy(c->op_site, site.data(), no * sizeof(int), hipMemcpyHostToDevice) ||
                 hipMemcpy(c->op_type, typ.data(), no * sizeof(int), hipMemcpyHostToDevice) ||
                 hipMemcpy(c->site_ops, sops.data(), no * sizeof(int), hipMemcpyHostToDevice)))
    return fail("copy tables");
  if (!c->h_lz.empty()) {
    const size_t esz = c->t_real ? sizeof(double) : sizeof(cplx);
    std::vector<char> id((size_t)n * n * esz, 0);
    for (int i = 0; i < n; ++i) *(double*)(id.data() + ((size_t)i * n + i) * esz) = 1.0;
    if (hipMalloc(&c->lz, nchain * sizeof(int)) || hipMalloc(&c->ident, id.size()) ||
        hipMemcpy(c->lz, c->h_lz.data(), nchain * sizeof(int), hipMemcpyHostToDevice) ||
        hipMemcpy(c->ident, id.data(), id.size(), hipMemcpyHostToDevice))
      return fail("uneven Ltrot tables");
  }
  // the chains' arenas: one allocation, chain z at z * cs bytes
  const int nslot = c->nstm + 6;
  ArenaLayout sizing;
  place_chain(c, sizing, nslot);
  c->cs = (long long)((sizing.off + 4095) & ~(size_t)4095);
  if (hipMalloc(&c->arena, (size_t)c->cs * nchain)) return fail("hipMalloc arena");
  if (hipMemset(c->arena, 0, (size_t)c->cs * nchain)) return fail("clear arena");
  ArenaLayout ar;
  ar.base = c->arena;
  place_chain(c, ar, nslot);
  c->qr.nchain = nchain;
  c->qr.cs = c->cs;
  c->u = c->u_arena;
  const size_t nn = (size_t)n * n;
  const size_t tsz = c->t_real ? sizeof(double) : sizeof(cplx);
  for (int z = 0; z < nchain; ++z) {
    const alfq_model_desc* dz = &descs[z];
    std::vector<cplx> eexp, myexp, expal, phf;
    std::vector<double> gr;
    coupling_tables(dz, eexp, myexp, expal, phf, gr);
    ChainState cs{};
    cs.phase = make_double2(1.0, 0.0);
    DevScalars ds{};
    ds.z_cgr = make_double2(1.0, 0.0);
    ds.opph = make_double2(1.0, 0.0);
    // fields +1 on the chain's own slices, none (0: identity exponentials) beyond them
    std::vector<int8_t> f1((size_t)std::max(no, 1) * c->L, 0);
    std::fill(f1.begin(), f1.begin() + (size_t)std::max(no, 1) * dz->ltrot, (int8_t)1);
    if (c->symm && (put_chain(c, z, (char*)c->expTh, (const char*)dz->expT_half, nn * tsz) ||
                    put_chain(c, z, (char*)c->expThinv, (const char*)dz->expT_half_inv, nn * tsz)))
      return fail("copy half-step exponentials");
    if (put_chain(c, z, (char*)c->expT, (const char*)dz->expT, nn * tsz) ||
        put_chain(c, z, (char*)c->expTinv, (const char*)dz->expT_inv, nn * tsz) ||
        put_chain(c, z, c->st, &cs, 1) || put_chain(c, z, c->sc, &ds, 1) ||
        put_chain(c, z, c->f, f1.data(), f1.size()))
      return fail("copy chain state");
    if (no > 0 && (put_chain(c, z, c->eexp, eexp.data(), eexp.size()) ||
                   put_chain(c, z, c->myexp, myexp.data(), myexp.size()) ||
                   put_chain(c, z, c->expalpha, expal.data(), expal.size()) ||
                   put_chain(c, z, c->phasefac, phf.data(), phf.size()) ||
                   put_chain(c, z, c->gama_ratio, gr.data(), gr.size())))
      return fail("copy coupling tables");
  }
  c->var_st.assign(c->nstm + 1, -1);
  if (udv_reset_dev(c, &c->var_r, 'R') || udv_reset_dev(c, &c->var_l, 'L')) return fail("udv");
  for (int k = 1; k <= c->nstm; ++k)
    if (udv_reset_dev(c, &c->var_st[k], 'L')) return fail("udv");
  if (hipEventCreate(&c->ev0) || hipEventCreate(&c->ev1)) return fail("events");
  if (hipMalloc(&c->stat_dev, 3 * sizeof(int) * nchain) ||
      hipHostMalloc((void**)&c->stat_host, 3 * sizeof(int) * nchain, hipHostMallocDefault))
    return fail("status buffers");
  if (hipStreamSynchronize(c->s)) return fail("sync");
  // ALFQ_SWEEP_GRAPHS=0 launches every kernel from the host (kernel tracers that do not
  // follow graph replays need this)
  if (const char* e = std::getenv("ALFQ_SWEEP_GRAPHS")) c->graphs_enabled = std::atoi(e) != 0;
  // ALFQ_ROUND_FUSION=0 keeps a round's X product and flush as two launches for few chains
  if (const char* e = std::getenv("ALFQ_ROUND_FUSION")) c->round_fusion = std::atoi(e) != 0;
  *out = c;
  return ST_NONE;
}

}  // namespace

extern "C" {

int alfq_ctx_create(const alfq_model_desc* d, alfq_ctx** out) { return create_impl(d, 1, out); }

int alfq_ctx_create_batch(const alfq_model_desc* descs, int32_t nchain, alfq_ctx** out) {
  return create_impl(descs, nchain, out);
}

int alfq_chain_bytes(const alfq_model_desc* d, int64_t* bytes) {
  if (!d || !bytes) return ST_GENERIC;
  int st;
  if ((st = check_desc(d, d))) return st;
  *bytes = (int64_t)chain_arena_bytes(d);
  return ST_NONE;
}

int alfq_chain_load(alfq_ctx* dst, int32_t dst_chain, alfq_ctx* src, int32_t src_chain) {
  NULL_STREAM_LOCK;
  if (!dst || !src || dst == src) { alfq_set_error("chain_load: two distinct contexts are needed"); return ST_GENERIC; }
  if (dst_chain < 0 || dst_chain >= dst->B || src_chain < 0 || src_chain >= src->B) {
    alfq_set_error("chain_load: chain index outside the batch");
    return ST_GENERIC;
  }
  if (dst->n != src->n || dst->L != src->L || dst->W != src->W || dst->n_op != src->n_op ||
      dst->n_sun != src->n_sun || dst->t_real != src->t_real || dst->symm != src->symm || dst->cs != src->cs ||
      dst->pool.size() != src->pool.size() || dst->h_op_type != src->h_op_type || dst->h_op_site != src->h_op_site) {
    alfq_set_error("chain_load: the contexts differ in shape or operator structure");
    return ST_HAMILTONIAN;
  }
  if (dst->lz || src->lz) {
    alfq_set_error("chain_load: batches of uneven Ltrot cannot exchange chains");
    return ST_GENERIC;
  }
  if (dst->obs.buf || src->obs.buf || dst->tau.mem || src->tau.mem || dst->cb_buf || src->cb_buf) {
    alfq_set_error("chain_load: contexts with measurements or a checkerboard Op_T cannot exchange chains");
    return ST_GENERIC;
  }
  if (dst->rng_mode != 2 || src->rng_mode != 2) {
    alfq_set_error("chain_load: both contexts need Philox streams");
    return ST_GENERIC;
  }
  HIP_OK(hipStreamSynchronize(src->s));
  // the chain's whole arena (model tables, fields, G, UDV slots, chain state, scalars, log)
  char* d0 = dst->arena + (long long)dst_chain * dst->cs;
  const char* s0 = src->arena + (long long)src_chain * src->cs;
  HIP_OK(hipMemcpyAsync(d0, s0, (size_t)dst->cs, hipMemcpyDeviceToDevice, dst->s));
  // the host's buffer and slot bindings may differ between the contexts: the current G and
  // every bound UDV variable go where the destination's bindings point
  auto rel = [](const alfq_ctx* c, const void* p) { return (long long)((const char*)p - c->arena); };
  const size_t gbytes = (size_t)dst->n * dst->n * sizeof(cplx);
  if (rel(dst, dst->G) != rel(src, src->G))
    HIP_OK(hipMemcpyAsync(d0 + rel(dst, dst->G), s0 + rel(src, src->G), gbytes, hipMemcpyDeviceToDevice, dst->s));
  auto move_var = [&](int vd, int vs) -> int {
    if (vd < 0 || vs < 0 || vd == vs) return ST_NONE;
    const UDVSlot &a = dst->pool[vd], &b = src->pool[vs];
    HIP_OK(hipMemcpyAsync(d0 + rel(dst, a.U), s0 + rel(src, b.U), gbytes, hipMemcpyDeviceToDevice, dst->s));
    HIP_OK(hipMemcpyAsync(d0 + rel(dst, a.V), s0 + rel(src, b.V), gbytes, hipMemcpyDeviceToDevice, dst->s));
    HIP_OK(hipMemcpyAsync(d0 + rel(dst, a.D), s0 + rel(src, b.D), dst->n * sizeof(cplx), hipMemcpyDeviceToDevice,
                          dst->s));
    HIP_OK(hipMemcpyAsync(d0 + rel(dst, a.sc), s0 + rel(src, b.sc), 2 * sizeof(cplx), hipMemcpyDeviceToDevice, dst->s));
    return ST_NONE;
  };
  int st;
  if ((st = move_var(dst->var_r, src->var_r)) || (st = move_var(dst->var_l, src->var_l))) return st;
  for (size_t k = 0; k < dst->var_st.size() && k < src->var_st.size(); ++k)
    if ((st = move_var(dst->var_st[k], src->var_st[k]))) return st;
  if (dst->oz_buf) {
    // the destination's digit planes of exp(-dtau T) and exp(dtau T), split again (shared by
    // the batch's chains only if the loaded chain has the same hopping)
    if ((st = oz_split_hopping(dst))) return st;
  }
  return ST_NONE;
}

int alfq_set_chain(alfq_ctx* c, int32_t chain) {
  if (!c) return ST_GENERIC;
  if (chain < 0 || chain >= c->B) { alfq_set_error("chain index outside the batch"); return ST_GENERIC; }
  c->cur = chain;
  return ST_NONE;
}

int alfq_batch_size(alfq_ctx* c, int32_t* nchain) {
  if (!c || !nchain) return ST_GENERIC;
  *nchain = c->B;
  return ST_NONE;
}

int alfq_ctx_destroy(alfq_ctx* c) {
  NULL_STREAM_LOCK;
  if (!c) return ST_NONE;
  if (c->s) hipStreamSynchronize(c->s);
  drop_graphs(c);
  if (c->obs.buf) hipFree(c->obs.buf);
  if (c->obs_tables) hipFree(c->obs_tables);
  if (c->cb_buf) hipFree(c->cb_buf);
  if (c->tau.mem) hipFree(c->tau.mem);
  if (c->tau.acc) hipFree(c->tau.acc);
  if (c->cb_sites) hipFree(c->cb_sites);
  if (c->oz_buf) hipFree(c->oz_buf);
  if (c->lz) hipFree(c->lz);
  if (c->ident) hipFree(c->ident);
  void* ptrs[] = {c->uoff_up, c->uoff_dn, c->op_site, c->op_type, c->site_ptr, c->site_ops, c->arena, c->u_tape,
                  c->stat_dev};
  for (void* p : ptrs)
    if (p) hipFree(p);
  for (auto& pr : c->chain_timer.ev) { hipEventDestroy(pr.first); hipEventDestroy(pr.second); }
  if (c->stat_host) hipHostFree(c->stat_host);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->s) hipStreamDestroy(c->s);
  delete c;
  return ST_NONE;
}

int alfq_set_delay(alfq_ctx* c, int32_t k) {
  if (!c) return ST_GENERIC;
  if (k < 1 || k > CHAIN_KMAX) { alfq_set_error("delay depth outside [1, 64]"); return ST_GENERIC; }
  c->K = k;
  return ST_NONE;
}

int alfq_set_gemm_policy(alfq_ctx* c, int32_t policy) {
  if (!c) return ST_GENERIC;
  if (policy < 0 || policy >= GEMM_POLICIES) { alfq_set_error("gemm policy outside 0..3"); return ST_GENERIC; }
  c->gemm_policy = policy;
  c->qr.policy = policy;
  c->tau.w2.policy = policy;
  return ST_NONE;
}

int alfq_check_wrap_args(alfq_ctx* c, int32_t propose_s0, int32_t nt_sequential_start, int32_t nt_sequential_end,
                         int32_t n_global_tau) {
  if (!c) return ST_GENERIC;
  const int no = (int)c->h_op_type.size();
  if (nt_sequential_start != 1 || nt_sequential_end != no) {
    alfq_set_error("wrap arguments: the engine updates every operator of a slice in sequence "
                   "(Nt_sequential_start = 1, Nt_sequential_end = size(Op_V, 1))");
    return ST_HAMILTONIAN;
  }
  if (n_global_tau != 0) {
    alfq_set_error("wrap arguments: global moves in the slice (N_Global_tau > 0) are not supported");
    return ST_HAMILTONIAN;
  }
  if (propose_s0)
    for (int t : c->h_op_type)
      if (t == 1) {
        alfq_set_error("wrap arguments: Propose_S0 with Ising (type 1) operators changes the proposal "
                       "probability; the engine proposes every visit");
        return ST_HAMILTONIAN;
      }
  return ST_NONE;
}

int alfq_set_round_fusion(alfq_ctx* c, int32_t on) {
  if (!c) return ST_GENERIC;
  c->round_fusion = on != 0;
  return ST_NONE;
}

int alfq_set_precision(alfq_ctx* c, int32_t mode) {
  if (!c) return ST_GENERIC;
  if (mode != ALFQ_PREC_FP64 && mode != ALFQ_PREC_FP32_WRAPS && mode != ALFQ_PREC_INT8_SPLIT) {
    alfq_set_error("precision: unknown mode");
    return ST_GENERIC;
  }
  if (mode == ALFQ_PREC_INT8_SPLIT && c->lz) {
    alfq_set_error("precision: the int8 split products need chains of one Ltrot");
    return ST_GENERIC;
  }
  if (mode == ALFQ_PREC_INT8_SPLIT && !c->oz_buf) {
    NULL_STREAM_LOCK;
    int st = oz_setup(c);
    if (st) return st;
  }
  c->prec = mode;
  return ST_NONE;
}

int alfq_obs_setup(alfq_ctx* c, const alfq_obs_desc* d) {
  NULL_STREAM_LOCK;
  if (!c || !d || !d->invlist || !d->imj || d->nnz < 0 || (d->nnz > 0 && (!d->o_row || !d->o_col || !d->o_val))) {
    alfq_set_error("obs_setup: null argument");
    return ST_GENERIC;
  }
  if (c->lz) { alfq_set_error("obs_setup: measurements need chains of one Ltrot"); return ST_GENERIC; }
  const int nc = d->ncell, no = d->norb, n = c->n, nnz = d->nnz;
  if (nc < 1 || no < 1 || (long long)nc * no != n) {
    alfq_set_error("obs_setup: ncell x norb must equal Ndim");
    return ST_HAMILTONIAN;
  }
  if (d->lobs_st < 1 || d->lobs_en > c->L || d->lobs_st > d->lobs_en) {
    alfq_set_error("obs_setup: measuring interval outside 1..Ltrot");
    return ST_GENERIC;
  }
  // shared tables: site of (cell, orbital), partner cell of (cell, distance class), O's pattern
  std::vector<int> key;
  key.reserve(3 + (size_t)nc * no + (size_t)nc * nc + 2 * (size_t)nnz);
  key.push_back(nc); key.push_back(no); key.push_back(nnz);
  std::vector<int> seen(n, 0);
  for (int I = 0; I < nc; ++I)
    for (int a = 0; a < no; ++a) {
      const int v = d->invlist[I + (long long)nc * a] - 1;
      if (v < 0 || v >= n || seen[v]++) { alfq_set_error("obs_setup: Invlist is not a permutation of the sites"); return ST_HAMILTONIAN; }
      key.push_back(v);
    }
  std::vector<int> jm((size_t)nc * nc, -1);
  for (int I = 0; I < nc; ++I)
    for (int J = 0; J < nc; ++J) {
      const int dc = d->imj[I + (long long)nc * J] - 1;
      if (dc < 0 || dc >= nc || jm[(size_t)I * nc + dc] >= 0) {
        alfq_set_error("obs_setup: imj(I, :) is not a permutation of the cells");
        return ST_HAMILTONIAN;
      }
      jm[(size_t)I * nc + dc] = J;
    }
  key.insert(key.end(), jm.begin(), jm.end());
  for (int e = 0; e < nnz; ++e) {
    const int i = d->o_row[e] - 1, j = d->o_col[e] - 1;
    if (i < 0 || i >= n || j < 0 || j >= n) { alfq_set_error("obs_setup: hopping entry outside the sites"); return ST_HAMILTONIAN; }
    key.push_back(i);
  }
  for (int e = 0; e < nnz; ++e) key.push_back(d->o_col[e] - 1);
  auto r256 = [](long long b) { return (b + 255) & ~255LL; };
  if (!c->obs.buf) {
    ObsDev o;
    o.n = n; o.ncell = nc; o.norb = no; o.nnz = nnz; o.n_sun = c->n_sun;
    o.off_u = 0;
    o.off_ov = 256;
    o.off_head = r256(o.off_ov + (long long)nnz * (long long)sizeof(cplx));
    o.off_den0 = o.off_head + r256(sizeof(ObsHead));
    o.off_latt = o.off_den0 + r256((long long)no * sizeof(cplx));
    const long long end = o.off_latt + 3LL * nc * no * no * (long long)sizeof(cplx);
    const long long ocs = (end + 4095) & ~4095LL;
    const size_t ntab = (size_t)nc * no + (size_t)nc * nc + 2 * (size_t)nnz;
    if (hipMalloc(&o.buf, (size_t)ocs * c->B) != hipSuccess) { alfq_set_error("obs_setup: hipMalloc"); return ST_GENERIC; }
    if (hipMalloc(&c->obs_tables, std::max<size_t>(ntab, 1) * sizeof(int)) != hipSuccess) {
      hipFree(o.buf);
      alfq_set_error("obs_setup: hipMalloc");
      return ST_GENERIC;
    }
    HIP_OK(hipMemset(o.buf, 0, (size_t)ocs * c->B));
    HIP_OK(hipMemcpy(c->obs_tables, key.data() + 3, ntab * sizeof(int), hipMemcpyHostToDevice));
    o.site_of = c->obs_tables;
    o.jm = o.site_of + (size_t)nc * no;
    o.oi = o.jm + (size_t)nc * nc;
    o.oj = o.oi + nnz;
    c->obs = o;
    c->obs_cs = ocs;
    c->obs_key = key;
  } else if (key != c->obs_key) {
    alfq_set_error("obs_setup: the lattice tables differ from the batch's");
    return ST_HAMILTONIAN;
  }
  char* blk = c->obs.buf + (long long)c->cur * c->obs_cs;
  const double u = d->ham_u;
  HIP_OK(hipMemcpy(blk + c->obs.off_u, &u, sizeof(double), hipMemcpyHostToDevice));
  if (nnz > 0) HIP_OK(hipMemcpy(blk + c->obs.off_ov, d->o_val, (size_t)nnz * sizeof(cplx), hipMemcpyHostToDevice));
  c->lobs_st = d->lobs_st;
  c->lobs_en = d->lobs_en;
  return ST_NONE;
}

int alfq_set_checkerboard(alfq_ctx* c, const alfq_checkerboard_desc* d) {
  NULL_STREAM_LOCK;
  if (!c || !d || !d->fam_len || !d->p || !d->expo || !d->expo_inv) {
    alfq_set_error("set_checkerboard: null argument");
    return ST_GENERIC;
  }
  if (c->lz) { alfq_set_error("set_checkerboard: a checkerboard Op_T needs chains of one Ltrot"); return ST_HAMILTONIAN; }
  const int n = c->n, nb = d->nbond, nf = d->nfam;
  if (nb < 1 || nf < 1 || nf > CB_MAXFAM) {
    alfq_set_error("set_checkerboard: need 1..16 families and at least one bond");
    return ST_HAMILTONIAN;
  }
  if (n > 1024) {
    alfq_set_error("set_checkerboard: Ndim above 1024");
    return ST_HAMILTONIAN;
  }
  // key: families, then the bond sites (0-based); a family's bonds must sit on disjoint sites
  std::vector<int> key;
  key.reserve(1 + nf + 2 * (size_t)nb);
  key.push_back(nf);
  CbFams fam{};
  fam.nfam = nf;
  int tot = 0;
  for (int f = 0; f < nf; ++f) {
    if (d->fam_len[f] < 0) { alfq_set_error("set_checkerboard: negative family length"); return ST_HAMILTONIAN; }
    fam.off[f] = tot;
    tot += d->fam_len[f];
    key.push_back(d->fam_len[f]);
  }
  fam.off[nf] = tot;
  if (tot != nb) { alfq_set_error("set_checkerboard: family lengths do not add up to nbond"); return ST_HAMILTONIAN; }
  std::vector<int> mark(n, -1);
  for (int f = 0; f < nf; ++f)
    for (int b = fam.off[f]; b < fam.off[f + 1]; ++b) {
      const int i = d->p[2 * (long long)b] - 1, j = d->p[2 * (long long)b + 1] - 1;
      if (i < 0 || i >= n || j < 0 || j >= n || i == j) {
        alfq_set_error("set_checkerboard: bond sites outside 1..Ndim or equal");
        return ST_HAMILTONIAN;
      }
      if (mark[i] == f || mark[j] == f) {
        alfq_set_error("set_checkerboard: non-commuting bonds in one family (shared site)");
        return ST_HAMILTONIAN;
      }
      mark[i] = mark[j] = f;
      key.push_back(i);
      key.push_back(j);
    }
  const bool half = d->expo_half != nullptr && d->expo_half_inv != nullptr;
  if ((d->expo_half == nullptr) != (d->expo_half_inv == nullptr)) {
    alfq_set_error("set_checkerboard: give both half-step bond matrices or neither");
    return ST_HAMILTONIAN;
  }
  key.push_back(half ? 1 : 0);
  const long long per = (half ? 4LL : 2LL) * 4 * nb * (long long)sizeof(cplx);
  const long long mstride = (per + 255) & ~255LL;
  if (!c->cb_buf) {
    char* buf = nullptr;
    int2* sites = nullptr;
    if (hipMalloc(&buf, (size_t)mstride * c->B) != hipSuccess) { alfq_set_error("set_checkerboard: hipMalloc"); return ST_GENERIC; }
    if (hipMalloc(&sites, (size_t)nb * sizeof(int2)) != hipSuccess) {
      hipFree(buf);
      alfq_set_error("set_checkerboard: hipMalloc");
      return ST_GENERIC;
    }
    c->cb_buf = buf;
    c->cb_sites = sites;
    HIP_OK(hipMemset(c->cb_buf, 0, (size_t)mstride * c->B));
    HIP_OK(hipMemcpy(c->cb_sites, key.data() + 1 + nf, (size_t)nb * sizeof(int2), hipMemcpyHostToDevice));
    c->cb_key = key;
    c->cb_set.assign(c->B, 0);
    CbOps o{};
    o.sites = c->cb_sites;
    o.mats = (const cplx*)c->cb_buf;
    o.mstride = mstride;
    o.fam = fam;
    o.nbond = nb;
    c->cb = o;
    o.mats = (const cplx*)(c->cb_buf + 4LL * nb * (long long)sizeof(cplx));
    c->cb_inv = o;
    if (half) {
      o.mats = (const cplx*)(c->cb_buf + 8LL * nb * (long long)sizeof(cplx));
      c->cb_half = o;
      o.mats = (const cplx*)(c->cb_buf + 12LL * nb * (long long)sizeof(cplx));
      c->cb_half_inv = o;
    }
    c->cb_symm = half;
  } else if (key != c->cb_key) {
    alfq_set_error("set_checkerboard: the bonds differ from the batch's");
    return ST_HAMILTONIAN;
  }
  char* blk = c->cb_buf + (long long)c->cur * mstride;
  const size_t mb = 4 * (size_t)nb * sizeof(cplx);
  HIP_OK(hipMemcpy(blk, d->expo, mb, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(blk + mb, d->expo_inv, mb, hipMemcpyHostToDevice));
  if (half) {
    HIP_OK(hipMemcpy(blk + 2 * mb, d->expo_half, mb, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(blk + 3 * mb, d->expo_half_inv, mb, hipMemcpyHostToDevice));
  }
  // real arithmetic (RealExpOpT) while every chain's matrices are real
  bool real = true;
  for (size_t i = 0; i < 4 * (size_t)nb && real; ++i) real = d->expo[2 * i + 1] == 0.0 && d->expo_inv[2 * i + 1] == 0.0;
  for (size_t i = 0; half && i < 4 * (size_t)nb && real; ++i)
    real = d->expo_half[2 * i + 1] == 0.0 && d->expo_half_inv[2 * i + 1] == 0.0;
  if (c->cb_real.empty()) c->cb_real.assign(c->B, 1);
  c->cb_real[c->cur] = real;
  const bool all_real = std::all_of(c->cb_real.begin(), c->cb_real.end(), [](char v) { return v != 0; });
  c->cb.is_real = c->cb_inv.is_real = c->cb_half.is_real = c->cb_half_inv.is_real = all_real;
  c->cb_set[c->cur] = 1;
  drop_graphs(c);                        // captured sweeps may hold the other arithmetic
  c->cb_on = std::all_of(c->cb_set.begin(), c->cb_set.end(), [](char v) { return v != 0; });
  return ST_NONE;
}

// Buffers of TAU_M: regions of cs bytes per chain (chain stride cs, like the arena) and the
// sums of the time-displaced measurements.
int tau_alloc(alfq_ctx* c) {
  auto& t = c->tau;
  if (t.mem) return ST_NONE;
  const int n = c->n, m = 2 * n;
  const size_t nn = (size_t)n * n, mm = (size_t)m * m;
  auto place = [&](ArenaLayout& ar) {
    for (auto& p : t.b) p = ar.take<cplx>(nn);
    t.Xv = ar.take<cplx>(nn);
    t.Qv = ar.take<cplx>(nn);
    t.F1 = ar.take<cplx>(nn);
    t.F2 = ar.take<cplx>(nn);
    t.H2 = ar.take<cplx>(mm);
    t.Q2 = ar.take<cplx>(mm);
    t.Y2 = ar.take<cplx>(mm);
    t.O2 = ar.take<cplx>(mm);
    t.DUPv = ar.take<cplx>(n);
    t.DUPvinv = ar.take<cplx>(n);
    t.DUP2 = ar.take<cplx>(m);
    t.DUP2inv = ar.take<cplx>(m);
    t.zs = ar.take<cplx>(1);
    t.flag = ar.take<int>(1);
    t.ctl = ar.take<TauCtl>(1);
    t.part = ar.take<double>(8 * (size_t)CGR22_PARTS);
    qr_work_place(t.w2, m, ar);
  };
  ArenaLayout sizing;
  sizing.cap = (size_t)c->cs;
  place(sizing);
  if (sizing.overflow) { alfq_set_error("tau: a 2n x 2n buffer exceeds the chain's arena stride"); return ST_GENERIC; }
  const int regions = sizing.region + 1;
  const size_t bytes = (size_t)regions * c->B * (size_t)c->cs;
  if (hipMalloc(&t.mem, bytes) != hipSuccess) { t.mem = nullptr; alfq_set_error("tau: hipMalloc"); return ST_GENERIC; }
  HIP_OK(hipMemset(t.mem, 0, bytes));
  ArenaLayout ar;
  ar.base = t.mem;
  ar.cap = (size_t)c->cs;
  ar.rstride = (long long)c->B * c->cs;
  place(ar);
  t.regions = regions;
  t.w2.nchain = c->B;
  t.w2.cs = c->cs;
  t.w2.policy = c->gemm_policy;
  // sums: head, Den0[norb], [3][Ltrot + 1][ncell][norb][norb]
  const int nc = c->obs.ncell, no = c->obs.norb;
  auto r256 = [](long long b) { return (b + 255) & ~255LL; };
  t.off_den0 = r256(sizeof(TauHead));
  t.off_latt = t.off_den0 + r256((long long)no * sizeof(cplx));
  const long long end = t.off_latt + 3LL * (c->L + 1) * nc * no * no * (long long)sizeof(cplx);
  t.acc_cs = (end + 4095) & ~4095LL;
  if (hipMalloc(&t.acc, (size_t)t.acc_cs * c->B) != hipSuccess) {
    t.acc = nullptr;
    alfq_set_error("tau: hipMalloc");
    return ST_GENERIC;
  }
  HIP_OK(hipMemset(t.acc, 0, (size_t)t.acc_cs * c->B));
  return ST_NONE;
}

int alfq_tau_enable(alfq_ctx* c, int32_t on) {
  NULL_STREAM_LOCK;
  if (!c) return ST_GENERIC;
  if (on) {
    if (!c->obs.buf) { alfq_set_error("tau_enable: call alfq_obs_setup first"); return ST_GENERIC; }
    if (c->symm) {
      alfq_set_error("tau_enable: time-displaced measurements with Symm = .T. are not built");
      return ST_HAMILTONIAN;
    }
    int st;
    if ((st = tau_alloc(c))) return st;
  }
  c->tau_on = on != 0;
  return ST_NONE;
}

int alfq_tau_m(alfq_ctx* c) {
  NULL_STREAM_LOCK;
  if (!c) return ST_GENERIC;
  int st;
  if ((st = tau_m_dev(c))) return st;
  return check_status(c);
}

int alfq_tau_get(alfq_ctx* c, alfq_tau_sums* sums, double* latt, double* den0) {
  NULL_STREAM_LOCK;
  if (!c || !sums) return ST_GENERIC;
  auto& t = c->tau;
  if (!t.acc) { alfq_set_error("tau_get: call alfq_tau_enable first"); return ST_GENERIC; }
  HIP_OK(hipStreamSynchronize(c->s));
  const char* blk = t.acc + (long long)c->cur * t.acc_cs;
  TauHead h;
  TauCtl k;
  HIP_OK(hipMemcpy(&h, blk, sizeof(h), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(&k, at(c, t.ctl), sizeof(k), hipMemcpyDeviceToHost));
  sums->sign = h.sign;
  sums->n = h.n;
  sums->ncg_tau = k.ncg;
  sums->xmax_tau = k.xmax;
  sums->xmean_tau_sum = k.xmean;
  const int nc = c->obs.ncell, no = c->obs.norb;
  if (latt)
    HIP_OK(hipMemcpy(latt, blk + t.off_latt, 3ULL * (c->L + 1) * nc * no * no * sizeof(cplx), hipMemcpyDeviceToHost));
  if (den0) HIP_OK(hipMemcpy(den0, blk + t.off_den0, (size_t)no * sizeof(cplx), hipMemcpyDeviceToHost));
  return ST_NONE;
}

int alfq_obs_enable(alfq_ctx* c, int32_t on) {
  if (!c) return ST_GENERIC;
  if (on && !c->obs.buf) { alfq_set_error("obs_enable: call alfq_obs_setup first"); return ST_GENERIC; }
  c->obs_on = on != 0;
  return ST_NONE;
}

int alfq_obs_reset(alfq_ctx* c) {
  NULL_STREAM_LOCK;
  if (!c) return ST_GENERIC;
  if (!c->obs.buf) return ST_NONE;
  HIP_OK(hipStreamSynchronize(c->s));
  for (int z = 0; z < c->B; ++z)
    HIP_OK(hipMemsetAsync(c->obs.buf + (long long)z * c->obs_cs + c->obs.off_head, 0,
                          (size_t)(c->obs_cs - c->obs.off_head), c->s));
  if (c->tau.acc) HIP_OK(hipMemsetAsync(c->tau.acc, 0, (size_t)c->tau.acc_cs * c->B, c->s));
  HIP_OK(hipStreamSynchronize(c->s));
  return ST_NONE;
}

int alfq_obs_get(alfq_ctx* c, alfq_obs_sums* sums, double* latt, double* den0) {
  NULL_STREAM_LOCK;
  if (!c || !sums) return ST_GENERIC;
  if (!c->obs.buf) { alfq_set_error("obs_get: call alfq_obs_setup first"); return ST_GENERIC; }
  HIP_OK(hipStreamSynchronize(c->s));
  const char* blk = c->obs.buf + (long long)c->cur * c->obs_cs;
  ObsHead h;
  HIP_OK(hipMemcpy(&h, blk + c->obs.off_head, sizeof(h), hipMemcpyDeviceToHost));
  for (int q = 0; q < 4; ++q) {
    sums->scal[2 * q] = h.scal[q].x;
    sums->scal[2 * q + 1] = h.scal[q].y;
  }
  sums->sign_scal = h.sign_scal;
  sums->n_scal = h.n_scal;
  sums->sign_eq = h.sign_eq;
  sums->n_eq = h.n_eq;
  const int nc = c->obs.ncell, no = c->obs.norb;
  if (latt)
    HIP_OK(hipMemcpy(latt, blk + c->obs.off_latt, 3ULL * nc * no * no * sizeof(cplx), hipMemcpyDeviceToHost));
  if (den0) HIP_OK(hipMemcpy(den0, blk + c->obs.off_den0, (size_t)no * sizeof(cplx), hipMemcpyDeviceToHost));
  return ST_NONE;
}

int alfq_fields_set(alfq_ctx* c, const int32_t* f, int32_t n_op, int32_t ltrot) {
  NULL_STREAM_LOCK;
  if (!c || !f) return ST_GENERIC;
  if (n_op != c->n_op || ltrot != chain_ltrot(c, c->cur)) { alfq_set_error("fields: shape mismatch"); return ST_FIELDS; }
  std::vector<int8_t> h((size_t)n_op * ltrot);
  for (size_t i = 0; i < h.size(); ++i) {
    const int v = f[i];
    const int t = c->h_op_type[i % n_op];
    if (v == 0 || v < -2 || v > 2 || (t == 1 && (v == 2 || v == -2))) {
      alfq_set_error("fields: value outside the field type's range");
      return ST_FIELDS;
    }
    h[i] = (int8_t)v;
  }
  HIP_OK(hipStreamSynchronize(c->s));
  HIP_OK(hipMemcpy(at(c, c->f), h.data(), h.size(), hipMemcpyHostToDevice));
  return ST_NONE;
}

int alfq_fields_get(alfq_ctx* c, int32_t* f, int32_t n_op, int32_t ltrot) {
  NULL_STREAM_LOCK;
  if (!c || !f) return ST_GENERIC;
  if (n_op != c->n_op || ltrot != chain_ltrot(c, c->cur)) { alfq_set_error("fields: shape mismatch"); return ST_FIELDS; }
  std::vector<int8_t> h((size_t)n_op * ltrot);
  HIP_OK(hipStreamSynchronize(c->s));
  HIP_OK(hipMemcpy(h.data(), at(c, c->f), h.size(), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < h.size(); ++i) f[i] = h[i];
  return ST_NONE;
}

int alfq_green_set(alfq_ctx* c, const double* G, int32_t ld) {
  NULL_STREAM_LOCK;
  if (!c || !G || ld < c->n) return ST_GENERIC;
  HIP_OK(hipStreamSynchronize(c->s));
  // a contiguous G (ld = n, the shims' case) moves as one copy: a 2-D copy from pageable memory
  // is staged row by row
  if (ld == c->n)
    HIP_OK(hipMemcpy(at(c, c->G), G, (size_t)c->n * c->n * sizeof(cplx), hipMemcpyHostToDevice));
  else
    HIP_OK(hipMemcpy2D(at(c, c->G), c->n * sizeof(cplx), G, (size_t)ld * sizeof(cplx), c->n * sizeof(cplx), c->n,
                       hipMemcpyHostToDevice));
  return ST_NONE;
}

int alfq_green_get(alfq_ctx* c, double* G, int32_t ld) {
  NULL_STREAM_LOCK;
  if (!c || !G || ld < c->n) return ST_GENERIC;
  HIP_OK(hipStreamSynchronize(c->s));
  if (ld == c->n)
    HIP_OK(hipMemcpy(G, at(c, c->G), (size_t)c->n * c->n * sizeof(cplx), hipMemcpyDeviceToHost));
  else
    HIP_OK(hipMemcpy2D(G, (size_t)ld * sizeof(cplx), at(c, c->G), c->n * sizeof(cplx), c->n * sizeof(cplx), c->n,
                       hipMemcpyDeviceToHost));
  return ST_NONE;
}

int alfq_phase_set(alfq_ctx* c, const double* phase) {
  NULL_STREAM_LOCK;
  if (!c || !phase) return ST_GENERIC;
  HIP_OK(hipStreamSynchronize(c->s));
  const cplx p = make_double2(phase[0], phase[1]);
  HIP_OK(hipMemcpy(&at(c, c->st)->phase, &p, sizeof(cplx), hipMemcpyHostToDevice));
  return ST_NONE;
}

int alfq_phase_get(alfq_ctx* c, double* phase) {
  NULL_STREAM_LOCK;
  if (!c || !phase) return ST_GENERIC;
  HIP_OK(hipStreamSynchronize(c->s));
  cplx p;
  HIP_OK(hipMemcpy(&p, &at(c, c->st)->phase, sizeof(cplx), hipMemcpyDeviceToHost));
  phase[0] = p.x;
  phase[1] = p.y;
  return ST_NONE;
}

int alfq_udv_set(alfq_ctx* c, int32_t var, const double* U, const double* D, const double* V, char side) {
  NULL_STREAM_LOCK;
  if (!c) return ST_GENERIC;
  if (c->B > 1) { alfq_set_error("udv_set: single chains only (a slot binding is shared by the batch)"); return ST_GENERIC; }
  int* v = var_ref(c, var);
  if (!v) { alfq_set_error("udv: no such variable"); return ST_GENERIC; }
  const int id = slot_alloc(c);
  if (id < 0) return ST_GENERIC;
  UDVSlot& sl = c->pool[id];
  sl.side = (side == 'L' || side == 'l') ? 'L' : 'R';
  const size_t nn = (size_t)c->n * c->n;
  HIP_OK(hipStreamSynchronize(c->s));
  HIP_OK(hipMemcpy(sl.U, U, nn * sizeof(cplx), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(sl.V, V, nn * sizeof(cplx), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(sl.D, D, c->n * sizeof(cplx), hipMemcpyHostToDevice));
  // det(U) of a caller-supplied U: LU on the host (the device keeps it updated afterwards)
  {
    std::vector<std::complex<double>> a(nn);
    std::memcpy(a.data(), U, nn * sizeof(cplx));
    const int n = c->n;
    std::complex<double> det(1.0, 0.0);
    for (int k = 0; k < n; ++k) {
      int p = k;
      for (int i = k + 1; i < n; ++i)
        if (std::abs(a[i + (size_t)k * n]) > std::abs(a[p + (size_t)k * n])) p = i;
      if (p != k) {
        for (int j = 0; j < n; ++j) std::swap(a[k + (size_t)j * n], a[p + (size_t)j * n]);
        det = -det;
      }
      const std::complex<double> piv = a[k + (size_t)k * n];
      det *= piv;
      if (std::abs(piv) == 0.0) break;
      for (int i = k + 1; i < n; ++i) {
        const std::complex<double> l = a[i + (size_t)k * n] / piv;
        for (int j = k + 1; j < n; ++j) a[i + (size_t)j * n] -= l * a[k + (size_t)j * n];
      }
    }
    const cplx sc[2] = {make_double2(1.0, 0.0), make_double2(det.real(), det.imag())};
    HIP_OK(hipMemcpy(sl.sc, sc, sizeof(sc), hipMemcpyHostToDevice));
  }
  var_bind(c, v, id);
  slot_release(c, id);
  return ST_NONE;
}

int alfq_udv_get(alfq_ctx* c, int32_t var, double* U, double* D, double* V) {
  NULL_STREAM_LOCK;
  if (!c) return ST_GENERIC;
  int* v = var_ref(c, var);
  if (!v) { alfq_set_error("udv: no such variable"); return ST_GENERIC; }
  const UDVSlot& sl = c->pool[*v];
  const size_t nn = (size_t)c->n * c->n;
  HIP_OK(hipStreamSynchronize(c->s));
  if (U) HIP_OK(hipMemcpy(U, at(c, sl.U), nn * sizeof(cplx), hipMemcpyDeviceToHost));
  if (V) HIP_OK(hipMemcpy(V, at(c, sl.V), nn * sizeof(cplx), hipMemcpyDeviceToHost));
  if (D) HIP_OK(hipMemcpy(D, at(c, sl.D), c->n * sizeof(cplx), hipMemcpyDeviceToHost));
  return ST_NONE;
}

int alfq_udv_reset(alfq_ctx* c, int32_t var, char side) {
  if (!c) return ST_GENERIC;
  int* v = var_ref(c, var);
  if (!v) { alfq_set_error("udv: no such variable"); return ST_GENERIC; }
  return udv_reset_dev(c, v, (side == 'L' || side == 'l') ? 'L' : 'R');
}

int alfq_udv_assign(alfq_ctx* c, int32_t dst, int32_t src) {
  if (!c) return ST_GENERIC;
  int* d = var_ref(c, dst);
  int* s = var_ref(c, src);
  if (!d || !s) { alfq_set_error("udv: no such variable"); return ST_GENERIC; }
  var_bind(c, d, *s);
  return ST_NONE;
}

int alfq_rng_set_tape(alfq_ctx* c, const double* u, int64_t n) {
  NULL_STREAM_LOCK;
  if (!c || (!u && n > 0) || n < 0) return ST_GENERIC;
  if (c->B > 1) { alfq_set_error("random tapes are for single chains; use alfq_rng_set_philox per chain"); return ST_GENERIC; }
  HIP_OK(hipStreamSynchronize(c->s));
  drop_graphs(c);
  if (c->u_tape) hipFree(c->u_tape);
  c->u_tape = nullptr;
  HIP_OK(hipMalloc(&c->u_tape, std::max<int64_t>(n, 1) * sizeof(double)));
  if (n > 0) HIP_OK(hipMemcpy(c->u_tape, u, n * sizeof(double), hipMemcpyHostToDevice));
  c->u = c->u_tape;
  c->u_cap = n;
  c->u_len = n;
  c->u_base = 0;
  c->rng_mode = 1;
  c->rng_next = 0;
  const unsigned long long zero = 0;
  const long long len = n;
  HIP_OK(hipMemcpy(&c->st->rng_pos, &zero, sizeof(zero), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(&c->st->u_base, &zero, sizeof(zero), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(&c->st->u_len, &len, sizeof(len), hipMemcpyHostToDevice));
  return ST_NONE;
}

int alfq_rng_set_philox(alfq_ctx* c, uint64_t seed, uint32_t stream) {
  NULL_STREAM_LOCK;
  if (!c) return ST_GENERIC;
  HIP_OK(hipStreamSynchronize(c->s));
  c->rng_mode = 2;
  c->seed = seed;
  c->stream_id = stream;
  c->rng_next = 0;
  c->u_len = 0;
  c->u_base = 0;
  c->u = c->u_arena;
  c->u_cap = c->u_arena_cap;
  // the chain's key and stream live in its device state
  ChainState* stc = at(c, c->st);
  const unsigned long long zero = 0;
  const long long len0 = 0;
  HIP_OK(hipMemcpy(&stc->rng_pos, &zero, sizeof(zero), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(&stc->u_base, &zero, sizeof(zero), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(&stc->u_len, &len0, sizeof(len0), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(&stc->seed, &seed, sizeof(seed), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(&stc->stream, &stream, sizeof(stream), hipMemcpyHostToDevice));
  return ST_NONE;
}

int alfq_rng_set_position(alfq_ctx* c, int64_t pos) {
  NULL_STREAM_LOCK;
  if (!c || pos < 0) return ST_GENERIC;
  if (c->rng_mode != 2) { alfq_set_error("rng_set_position: the chain has no Philox stream"); return ST_GENERIC; }
  HIP_OK(hipStreamSynchronize(c->s));
  ChainState* stc = at(c, c->st);
  const unsigned long long p = (unsigned long long)pos, zero = 0;
  const long long len0 = 0;
  HIP_OK(hipMemcpy(&stc->rng_pos, &p, sizeof(p), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(&stc->u_base, &zero, sizeof(zero), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(&stc->u_len, &len0, sizeof(len0), hipMemcpyHostToDevice));
  c->rng_next = p;
  return ST_NONE;
}

int alfq_rng_position(alfq_ctx* c, int64_t* pos) {
  NULL_STREAM_LOCK;
  if (!c || !pos) return ST_GENERIC;
  HIP_OK(hipStreamSynchronize(c->s));
  unsigned long long p;
  HIP_OK(hipMemcpy(&p, &at(c, c->st)->rng_pos, sizeof(p), hipMemcpyDeviceToHost));
  *pos = (int64_t)p;
  return ST_NONE;
}

static int phase_in(alfq_ctx* c, const double* phase) {
  if (!phase) return ST_NONE;
  return alfq_phase_set(c, phase);
}
static int phase_out(alfq_ctx* c, double* phase) {
  if (!phase) return ST_NONE;
  return alfq_phase_get(c, phase);
}

int alfq_wrapgrup(alfq_ctx* c, int32_t ntau, double* phase) {
  if (!c) return ST_GENERIC;
  if (ntau < 0 || ntau >= c->L) { alfq_set_error("WRAPGRUP: NTAU outside [0, Ltrot-1]"); return ST_GENERIC; }
  int st;
  if ((st = phase_in(c, phase))) return st;
  if ((st = wrapgrup_dev(c, ntau))) return st;
  if ((st = check_status(c))) return st;
  return phase_out(c, phase);
}

int alfq_wrapgrdo(alfq_ctx* c, int32_t ntau, double* phase) {
  if (!c) return ST_GENERIC;
  if (ntau < 1 || ntau > c->L) { alfq_set_error("WRAPGRDO: NTAU outside [1, Ltrot]"); return ST_GENERIC; }
  int st;
  if ((st = phase_in(c, phase))) return st;
  if ((st = wrapgrdo_dev(c, ntau))) return st;
  if ((st = check_status(c))) return st;
  return phase_out(c, phase);
}

int alfq_wrapur(alfq_ctx* c, int32_t ntau, int32_t ntau1) {
  if (!c) return ST_GENERIC;
  if (ntau < 0 || ntau1 > c->L) { alfq_set_error("WRAPUR: slice range"); return ST_GENERIC; }
  int st;
  if ((st = wrap_udv_dev(c, &c->var_r, ntau, ntau1, false))) return st;
  return check_status(c);
}

int alfq_wrapul(alfq_ctx* c, int32_t ntau1, int32_t ntau) {
  if (!c) return ST_GENERIC;
  if (ntau < 0 || ntau1 > c->L) { alfq_set_error("WRAPUL: slice range"); return ST_GENERIC; }
  int st;
  if ((st = wrap_udv_dev(c, &c->var_l, ntau, ntau1, true))) return st;
  return check_status(c);
}

int alfq_cgr(alfq_ctx* c, int32_t nvar, int32_t var_r, int32_t var_l, int32_t check, double* z) {
  NULL_STREAM_LOCK;
  if (!c) return ST_GENERIC;
  if (nvar != 1 && nvar != 2) { alfq_set_error("CGR: NVAR must be 1 or 2"); return ST_GENERIC; }
  int* r = var_ref(c, var_r);
  int* l = var_ref(c, var_l);
  if (!r || !l) { alfq_set_error("CGR: no such UDV variable"); return ST_GENERIC; }
  int st;
  if ((st = cgr_dev(c, nvar, *r, *l, check))) return st;
  if ((st = check_status(c))) return st;
  if (z) {
    cplx zz;
    HIP_OK(hipMemcpy(&zz, &at(c, c->sc)->z_cgr, sizeof(cplx), hipMemcpyDeviceToHost));
    z[0] = zz.x;
    z[1] = zz.y;
  }
  return ST_NONE;
}

int alfq_init(alfq_ctx* c) {
  if (!c) return ST_GENERIC;
  int st;
  if ((st = init_dev(c))) return st;
  return check_status(c);
}

int alfq_sweep(alfq_ctx* c, int32_t nsweep) {
  if (!c) return ST_GENERIC;
  int st;
  for (int i = 0; i < nsweep; ++i) {
    if ((st = sweep_graph(c))) return st;
    if ((st = end_of_sweep(c))) return st;
  }
  return ST_NONE;
}

int alfq_sweep_many(alfq_ctx* const* ctxs, int32_t count, int32_t nsweep, double* ms) {
  if (!ctxs || count < 1) { alfq_set_error("sweep_many: no contexts"); return ST_GENERIC; }
  for (int i = 0; i < count; ++i)
    if (!ctxs[i]) { alfq_set_error("sweep_many: null context"); return ST_GENERIC; }
  for (int i = 0; i < count; ++i) HIP_OK(hipStreamSynchronize(ctxs[i]->s));
  std::vector<int> status(count, ST_NONE);
  std::vector<std::string> errors(count);
  const auto t0 = std::chrono::steady_clock::now();
  auto work = [&](int i) {
    alfq_ctx* c = ctxs[i];
    int st = ST_NONE;
    for (int k = 0; k < nsweep && st == ST_NONE; ++k) {
      st = sweep_graph(c);
      if (st == ST_NONE) st = end_of_sweep(c);
    }
    status[i] = st;
    if (st) errors[i] = alfq_last_error();
  };
  std::vector<std::thread> th;
  for (int i = 1; i < count; ++i) th.emplace_back(work, i);
  work(0);
  for (auto& t : th) t.join();
  const auto t1 = std::chrono::steady_clock::now();
  if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  for (int i = 0; i < count; ++i)
    if (status[i]) { alfq_set_error("chain " + std::to_string(i) + ": " + errors[i]); return status[i]; }
  return ST_NONE;
}

int alfq_sweep_timed(alfq_ctx* c, int32_t nsweep, double* ms) {
  if (!c || !ms) return ST_GENERIC;
  int st;
  HIP_OK(hipStreamSynchronize(c->s));
  HIP_OK(hipEventRecord(c->ev0, c->s));
  for (int i = 0; i < nsweep; ++i)
    if ((st = sweep_graph(c))) return st;
  HIP_OK(hipEventRecord(c->ev1, c->s));
  // a timed run reads the status once, after its sweeps
  if ((st = end_of_sweep(c))) return st;
  HIP_OK(hipEventSynchronize(c->ev1));
  float t = 0.f;
  HIP_OK(hipEventElapsedTime(&t, c->ev0, c->ev1));
  *ms = t;
  return ST_NONE;
}

int alfq_sync(alfq_ctx* c) {
  if (!c) return ST_GENERIC;
  return check_status(c);
}

int alfq_control_get(alfq_ctx* c, alfq_control* out) {
  NULL_STREAM_LOCK;
  if (!c || !out) return ST_GENERIC;
  HIP_OK(hipStreamSynchronize(c->s));
  ChainState cs;
  DevScalars ds;
  HIP_OK(hipMemcpy(&cs, at(c, c->st), sizeof(cs), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(&ds, at(c, c->sc), sizeof(ds), hipMemcpyDeviceToHost));
  out->nc_up = cs.nc;
  out->acc_up = cs.acc;
  out->ncg = ds.ncg;
  out->xmean_g_sum = ds.xmeang;
  out->xmax_g = ds.xmaxg;
  out->xmax_p = ds.xmaxp;
  return ST_NONE;
}

int alfq_control_reset(alfq_ctx* c) {
  NULL_STREAM_LOCK;
  if (!c) return ST_GENERIC;
  HIP_OK(hipStreamSynchronize(c->s));
  ChainState cs;
  DevScalars ds;
  HIP_OK(hipMemcpy(&cs, at(c, c->st), sizeof(cs), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(&ds, at(c, c->sc), sizeof(ds), hipMemcpyDeviceToHost));
  cs.nc = cs.acc = 0;
  ds.ncg = 0;
  ds.xmeang = ds.xmaxg = ds.xmaxp = 0.0;
  ds.nlog = 0;
  ds.status = 0;
  HIP_OK(hipMemcpy(at(c, c->st), &cs, sizeof(cs), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(at(c, c->sc), &ds, sizeof(ds), hipMemcpyHostToDevice));
  if (c->tau.ctl) {
    const TauCtl k{};
    HIP_OK(hipMemcpy(at(c, c->tau.ctl), &k, sizeof(k), hipMemcpyHostToDevice));
  }
  return ST_NONE;
}

int alfq_precision_log(alfq_ctx* c, double* out3, int64_t cap, int64_t* count) {
  NULL_STREAM_LOCK;
  if (!c || !count) return ST_GENERIC;
  HIP_OK(hipStreamSynchronize(c->s));
  DevScalars ds;
  HIP_OK(hipMemcpy(&ds, at(c, c->sc), sizeof(ds), hipMemcpyDeviceToHost));
  const int64_t avail = std::min<int64_t>(ds.nlog, LOG_CAP);
  *count = ds.nlog;
  const int64_t k = std::min<int64_t>(avail, cap);
  if (out3 && k > 0) HIP_OK(hipMemcpy(out3, at(c, c->log), 3 * k * sizeof(double), hipMemcpyDeviceToHost));
  return ST_NONE;
}

int alfq_sweep_graphs(alfq_ctx* c, int32_t enable) {
  if (!c) return ST_GENERIC;
  HIP_OK(hipStreamSynchronize(c->s));
  if (!enable) drop_graphs(c);
  c->graphs_enabled = enable != 0;
  return ST_NONE;
}

int alfq_kernel_timing(alfq_ctx* c, int32_t enable) {
  if (!c) return ST_GENERIC;
  HIP_OK(hipStreamSynchronize(c->s));
  c->chain_timer.on = enable != 0;
  c->chain_timer.used = 0;
  return ST_NONE;
}

int alfq_kernel_timing_get(alfq_ctx* c, double* total_ms, int64_t* launches) {
  if (!c || !total_ms || !launches) return ST_GENERIC;
  HIP_OK(hipStreamSynchronize(c->s));
  double t = 0.0;
  for (size_t i = 0; i < c->chain_timer.used; ++i) {
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, c->chain_timer.ev[i].first, c->chain_timer.ev[i].second));
    t += ms;
  }
  *total_ms = t;
  *launches = (int64_t)c->chain_timer.used;
  return ST_NONE;
}

}  // extern "C"

// Staging memory of alfq_gemm_noalloc: one fixed array per device, three regions of
// ALFQ_NOALLOC_OPERAND_BYTES (A, B, C).  Host code reaches it only through the device
// address hipGetSymbolAddress returns, fetched once per device.
static __device__ __align__(256) unsigned char scratch[3 * (size_t)ALFQ_NOALLOC_OPERAND_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's staging array (looked up on its first use there).
int noalloc_scratch(char** out) {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) { alfq_set_error("gemm_noalloc: device index out of range"); return ST_GENERIC; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) { alfq_set_error("gemm_noalloc: no device address for the staging array"); return ST_GENERIC; }
    g_scratch_ptr[dev] = p;
  }
  *out = (char*)g_scratch_ptr[dev];
  return ST_NONE;
}

// C = alpha A B + beta C on the int8 split products, with digit planes allocated here
// (op(A) = A and op(B) = B; at least one operand complex).
int gemm_split_dev(int32_t M, int32_t N, int32_t K, int oa, int32_t a_is_real, const void* dA, int32_t lda, int ob,
                   int32_t b_is_real, const void* dB, int32_t ldb, cplx alpha, cplx beta, cplx* dC, int32_t ldc) {
  if (oa != OP_N || ob != OP_N || (a_is_real && b_is_real)) {
    alfq_set_error("gemm: the int8 split products take op = 'N' and at least one complex operand");
    return ST_GENERIC;
  }
  if (M <= 0 || N <= 0) return ST_NONE;
  OzOperand a, b;
  a.rows = M;
  a.planes = a_is_real ? 1 : (b_is_real ? 2 : 3);
  b.rows = N;
  b.planes = b_is_real ? 1 : 2;
  a.kp = b.kp = oz_pitch(K);
  const size_t ad = oz_bytes(M, K, a.planes), bd = oz_bytes(N, K, b.planes);
  char* buf = nullptr;
  const size_t aoff = 0, aeoff = (ad + 255) & ~(size_t)255, boff = (aeoff + M * sizeof(int) + 255) & ~(size_t)255,
               beoff = (boff + bd + 255) & ~(size_t)255;
  if (hipMalloc(&buf, beoff + N * sizeof(int)) != hipSuccess) {
    alfq_set_error("gemm: hipMalloc of the digit planes");
    return ST_GENERIC;
  }
  a.d = (int8_t*)(buf + aoff);
  a.e = (int*)(buf + aeoff);
  b.d = (int8_t*)(buf + boff);
  b.e = (int*)(buf + beoff);
  int st = 0;
  if (K > 0) {
    if (!(st = oz_split(dA, !a_is_real, lda, M, K, true, nullptr, 0, a, 1, 0)) &&
        !(st = oz_split(dB, !b_is_real, ldb, N, K, false, nullptr, 0, b, 1, 0)))
      st = oz_gemm(M, N, K, a, b, dC, ldc, alpha, beta, nullptr, nullptr, 0, 1, 0);
  } else {
    hipMemset(a.e, 0, M * sizeof(int));
    hipMemset(b.e, 0, N * sizeof(int));
    st = oz_gemm(M, N, 0, a, b, dC, ldc, alpha, beta, nullptr, nullptr, 0, 1, 0);
  }
  if (hipDeviceSynchronize() != hipSuccess && !st) {
    alfq_set_error("gemm: split product failed");
    st = ST_GENERIC;
  }
  hipFree(buf);
  return st;
}

// The copy-in, product and copy-out of alfq_gemm on device buffers dA, dB, dC.
int gemm_staged(int32_t M, int32_t N, int32_t K, int oa, int32_t a_is_real, const double* A, int32_t lda, size_t asz,
                int ob, int32_t b_is_real, const double* B, int32_t ldb, size_t bsz, const double* alpha,
                const double* beta, double* Cm, int32_t ldc, size_t csz, int32_t prec, void* dA, void* dB, void* dC) {
  HIP_OK(hipMemcpy(dA, A, asz, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dB, B, bsz, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dC, Cm, csz, hipMemcpyHostToDevice));
  int st;
  if (prec == ALFQ_PREC_INT8_SPLIT) {
    st = gemm_split_dev(M, N, K, oa, a_is_real, dA, lda, ob, b_is_real, dB, ldb, make_double2(alpha[0], alpha[1]),
                        make_double2(beta[0], beta[1]), (cplx*)dC, ldc);
  } else {
    GemmArgs g = gemm_args(M, N, K, dA, lda, oa, !a_is_real, dB, ldb, ob, !b_is_real, dC, ldc,
                           make_double2(alpha[0], alpha[1]), make_double2(beta[0], beta[1]));
    g.prec = prec == ALFQ_PREC_FP32_WRAPS ? 1 : 0;
    st = gemm_launch(g, 0);
  }
  if (!st) {
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(Cm, dC, csz, hipMemcpyDeviceToHost));
  }
  return st;
}
}  // namespace

extern "C" {

int alfq_gemm_noalloc(int32_t M, int32_t N, int32_t K, char opa, int32_t a_is_real, const double* A, int32_t lda,
                      char opb, int32_t b_is_real, const double* B, int32_t ldb, const double* alpha,
                      const double* beta, double* Cm, int32_t ldc, int32_t prec) {
  NULL_STREAM_LOCK;
  if (M < 0 || N < 0 || K < 0 || !A || !B || !Cm || !alpha || !beta) {
    alfq_set_error("gemm_noalloc: bad arguments");
    return ST_GENERIC;
  }
  auto op = [](char o) { return (o == 'T' || o == 't') ? OP_T : ((o == 'C' || o == 'c') ? OP_C : OP_N); };
  const int oa = op(opa), ob = op(opb);
  const long long acols = (oa == OP_N) ? K : M, bcols = (ob == OP_N) ? N : K;
  const size_t asz = (size_t)lda * acols * (a_is_real ? sizeof(double) : sizeof(cplx));
  const size_t bsz = (size_t)ldb * bcols * (b_is_real ? sizeof(double) : sizeof(cplx));
  const size_t csz = (size_t)ldc * N * sizeof(cplx);
  if (prec == ALFQ_PREC_INT8_SPLIT) {
    alfq_set_error("gemm_noalloc: the int8 split products need digit-plane workspace; use alfq_gemm");
    return ST_GENERIC;
  }
  const size_t cap = (size_t)ALFQ_NOALLOC_OPERAND_BYTES;
  if (asz > cap || bsz > cap || csz > cap) {
    alfq_set_error("gemm_noalloc: an operand exceeds the " + std::to_string(cap) +
                   "-byte staging limit (A " + std::to_string(asz) + ", B " + std::to_string(bsz) + ", C " +
                   std::to_string(csz) + " bytes)");
    return ST_GENERIC;
  }
  char* base = nullptr;
  int st = noalloc_scratch(&base);
  if (st) return st;
  return gemm_staged(M, N, K, oa, a_is_real, A, lda, asz, ob, b_is_real, B, ldb, bsz, alpha, beta, Cm, ldc, csz,
                     prec, base, base + cap, base + 2 * cap);
}

int alfq_gemm(int32_t M, int32_t N, int32_t K, char opa, int32_t a_is_real, const double* A, int32_t lda,
              char opb, int32_t b_is_real, const double* B, int32_t ldb, const double* alpha, const double* beta,
              double* Cm, int32_t ldc, int32_t prec) {
  NULL_STREAM_LOCK;
  if (M < 0 || N < 0 || K < 0 || !A || !B || !Cm || !alpha || !beta) { alfq_set_error("gemm: bad arguments"); return ST_GENERIC; }
  auto op = [](char o) { return (o == 'T' || o == 't') ? OP_T : ((o == 'C' || o == 'c') ? OP_C : OP_N); };
  const int oa = op(opa), ob = op(opb);
  const long long acols = (oa == OP_N) ? K : M, bcols = (ob == OP_N) ? N : K;
  const size_t asz = (size_t)lda * acols * (a_is_real ? sizeof(double) : sizeof(cplx));
  const size_t bsz = (size_t)ldb * bcols * (b_is_real ? sizeof(double) : sizeof(cplx));
  const size_t csz = (size_t)ldc * N * sizeof(cplx);
  void *dA = nullptr, *dB = nullptr, *dC = nullptr;
  if (hipMalloc(&dA, asz ? asz : 16) || hipMalloc(&dB, bsz ? bsz : 16) || hipMalloc(&dC, csz ? csz : 16)) {
    hipFree(dA);
    hipFree(dB);
    hipFree(dC);
    alfq_set_error("gemm: hipMalloc");
    return ST_GENERIC;
  }
  int st = gemm_staged(M, N, K, oa, a_is_real, A, lda, asz, ob, b_is_real, B, ldb, bsz, alpha, beta, Cm, ldc, csz,
                       prec, dA, dB, dC);
  hipFree(dA);
  hipFree(dB);
  hipFree(dC);
  return st;
}

}  // extern "C"
