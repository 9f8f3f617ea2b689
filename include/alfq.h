/*
 * alfq.h -- C ABI of the gfx950 finite-temperature DQMC sweep engine.
 *
 * Drop-in boundary for the sequential sweep of ALF (Prog/main.F90:714-887).
 * Every entry point returns an int32 status whose values are the reference's
 * error enum (Libraries/Modules/runtime_error_mod.F90:56-68): 0 = ERROR_NONE,
 * 1 = ERROR_GENERIC, 3 = ERROR_UNSTABLE_MATRIX, 5 = ERROR_FIELDS,
 * 6 = ERROR_HAMILTONIAN.  alfq_last_error() gives the message.  No C++
 * exception crosses this boundary.
 *
 * Conventions: matrices are column-major; complex numbers are interleaved
 * (re, im) doubles, i.e. the memory of Fortran COMPLEX(KIND=8); indices that
 * the reference passes as Fortran integers keep their meaning (NTAU, NVAR,
 * sites are 1-based where the reference's are).  One context = one Markov
 * chain with N_FL = 1 on one HIP stream; contexts are independent and a
 * context must be driven by one host thread at a time.
 */
#ifndef ALFQ_H
#define ALFQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct alfq_ctx alfq_ctx;

/* What the model plugin sets up (Hamiltonian_main_mod.F90:181-197):
 * the dense hopping exponential of Op_T (OpTTypes_mod.F90:92-133, 203-234:
 * exp(g T) with g = -dtau, Hermitian-symmetrised) and the 1x1 interaction
 * operators Op_V(1:n_op, 1) (Operator_mod.F90:56-90). */
typedef struct alfq_model_desc {
  int32_t ndim;            /* Ndim */
  int32_t ltrot;           /* Ltrot */
  int32_t nwrap;           /* Nwrap */
  int32_t n_sun;           /* N_SUN */
  int32_t n_fl;            /* N_FL; this engine supports 1 */
  int32_t n_op;            /* size(Op_V, 1) */
  int32_t t_is_real;       /* 1: RealExpOpT (real symmetric), 0: CmplxExpOpT */
  const double* expT;      /* exp(g T): ndim*ndim reals, or 2*ndim*ndim (complex) */
  const double* expT_inv;  /* exp(-g T), same layout */
  const int32_t* op_P;     /* Op_V(n)%P(1), 1-based site, n_op entries */
  const int32_t* op_type;  /* Op_V(n)%type: 1 (Ising) or 2 (discrete HS) */
  const double* op_E;      /* Op_V(n)%E(1) */
  const double* op_g;      /* Op_V(n)%g, complex, 2*n_op doubles */
  const double* op_alpha;  /* Op_V(n)%alpha, complex, 2*n_op doubles */
  /* Symm = .T. (main.F90:763-772, 794-802): the half steps exp(g T/2) and exp(-g T/2)
   * (mat_1D2, invmat_1D2 of RealExpOpT/CmplxExpOpT, OpTTypes_mod.F90:110-129, 213-232), same
   * layout as expT; the measurements then see Hop_mod_Symm(GR) = exp(g T/2) GR exp(-g T/2)
   * (Hop_mod.F90:279-299).  NULL (both): Symm = .F.  With a checkerboard Op_T the bond half
   * steps of alfq_checkerboard_desc are applied instead. */
  const double* expT_half;
  const double* expT_half_inv;
} alfq_model_desc;

/* Statistics of Control_upgrade / Control_PrecisionG / Control_PrecisionP
 * (Prog/control_mod.F90:164-176, 207-320). */
typedef struct alfq_control {
  int64_t nc_up, acc_up;
  int64_t ncg;
  double xmean_g_sum;      /* XMEANG: sum of per-call mean deviations */
  double xmax_g;           /* XMAXG */
  double xmax_p;           /* XMAXP */
} alfq_control;

/* UDV variables addressed by the fine-grained calls */
#define ALFQ_UDVR (-1)     /* udvr */
#define ALFQ_UDVL (-2)     /* udvl */
/* k in 1..NSTM addresses udvst(k) */

/* ---- lifecycle ----------------------------------------------------------- */
/* Replaces Ham_Set's output consumption + Hop_mod_init (Prog/Hop_mod.F90:118-139),
 * Op_set's E_exp tables (Prog/Operator_mod.F90:436-466) and the allocations of
 * main.F90:588-605. */
int alfq_ctx_create(const alfq_model_desc* desc, alfq_ctx** out);
int alfq_ctx_destroy(alfq_ctx* ctx);
/* Batches (no counterpart in the reference, which runs one chain per MPI rank): nchain
 * independent Markov chains that advance in lockstep, every kernel launch covering all
 * of them.  descs[z] describes chain z; Ndim, Nwrap, N_SUN, the operators' sites and field
 * types and the type of T must agree, the couplings (U, ...), exp(-dtau T) and Ltrot may
 * differ.  With uneven Ltrot the batch sweeps the largest one, Lmax, with the stabilisation
 * schedule of Lmax: a chain of Ltrot < Lmax treats the slices beyond its own as identities
 * (no fields, no visits, no random numbers, exp(-dtau T) replaced by 1), which leaves its
 * Green's functions and weights unchanged, so it follows its own trajectory up to rounding
 * (its stabilisation points and NVAR follow Lmax); fields are set and read per chain with
 * the chain's own Ltrot.  Uneven batches take no measurements, checkerboard Op_T, int8
 * split products or alfq_chain_load.  alfq_ctx_create(desc) is the batch of one.  Sweeps,
 * wraps and CGR act on every chain; the accessors (fields, G, PHASE, random stream, UDV reads, control,
 * precision log) act on the chain chosen with alfq_set_chain (chain 0 initially).
 * Random tapes and alfq_udv_set are single-chain only. */
int alfq_ctx_create_batch(const alfq_model_desc* descs, int32_t nchain, alfq_ctx** out);
/* Device bytes one chain of the model takes in a batch (its arena), for admission by
 * HBM footprint (no counterpart in the reference). */
int alfq_chain_bytes(const alfq_model_desc* desc, int64_t* bytes);
int alfq_set_chain(alfq_ctx* ctx, int32_t chain);
/* Continuous batching (no counterpart in the reference): chain src_chain of src replaces
 * chain dst_chain of dst.  Both contexts need the same shape (Ndim, Ltrot, Nwrap, N_SUN,
 * operator sites and types, type of T, Symm) and Philox streams; the chain's couplings and
 * exp(-dtau T), fields, G, PHASE, UDV stack, random stream and statistics move with it, so a
 * chain initialised in a small context (alfq_init) joins a running batch at a sweep
 * boundary and continues exactly as it would have alone.  Enqueued on dst's stream after
 * its earlier work (src is synchronised first).  Contexts with measurements or a
 * checkerboard Op_T are refused. */
int alfq_chain_load(alfq_ctx* dst, int32_t dst_chain, alfq_ctx* src, int32_t src_chain);
int alfq_batch_size(alfq_ctx* ctx, int32_t* nchain);
/* Delayed-update depth k (1..64) of the Sherman-Morrison chain. */
int alfq_set_delay(alfq_ctx* ctx, int32_t k);
/* GEMM tiling policy of the context (no counterpart in the reference; picked by the
 * autotuner of alf_amd/runtime.py): 0 the estimate-based default, 1 the 64 x 64 kernel from
 * one output tile per CU, 2 never the 64 x 64 kernel, 3 the direct-load kernel up to 512
 * tiles of 32 x 32.  Results agree to rounding across policies. */
int alfq_set_gemm_policy(alfq_ctx* ctx, int32_t policy);
/* Arithmetic of the hopping wraps of G in WRAPGRUP/WRAPGRDO (Hop_mod_mmthr/mmthl and
 * their inverses, Prog/Hop_mod.F90:143-250).  ALFQ_PREC_FP64 (the default) is the drop-in;
 * ALFQ_PREC_FP32_WRAPS rounds both operands of those products to FP32 and multiplies and
 * accumulates in FP32 on the matrix cores (no counterpart in the reference).  The UDV
 * propagation, QR, CGR, the Metropolis ratio and the delayed updates stay FP64 in both
 * modes, so every stabilisation point rebuilds G in FP64 and Precision Green measures the
 * drift the FP32 wraps accumulated since the last one. */
#define ALFQ_PREC_FP64 0
#define ALFQ_PREC_FP32_WRAPS 1
/* ALFQ_PREC_INT8_SPLIT (no counterpart in the reference): the same products as exact
 * integer sums on the int8 matrix cores.  Each operand row (left) or column (right) is
 * 2^e times 8 base-128 int8 digits rounded to nearest (56 bits of the row's range); the digit
 * pairs with s + t <= 9 are multiplied exactly (int32 sums) and combined in FP64.  The
 * product's error is about 2^-56 of |row| |column|, against 2^-24 for FP32 wraps.  The
 * first call allocates the digit planes and splits exp(-dtau T) and exp(dtau T) once. */
#define ALFQ_PREC_INT8_SPLIT 2
int alfq_set_precision(alfq_ctx* ctx, int32_t mode);

/* ---- state movement ------------------------------------------------------ */
/* nsigma%f (Fields_mod.F90:79-99): n_op x ltrot integers, column-major. */
int alfq_fields_set(alfq_ctx* ctx, const int32_t* f, int32_t n_op, int32_t ltrot);
int alfq_fields_get(alfq_ctx* ctx, int32_t* f, int32_t n_op, int32_t ltrot);
/* GR(:,:,1) */
int alfq_green_set(alfq_ctx* ctx, const double* G, int32_t ld);
int alfq_green_get(alfq_ctx* ctx, double* G, int32_t ld);
/* PHASE */
int alfq_phase_set(alfq_ctx* ctx, const double* phase);
int alfq_phase_get(alfq_ctx* ctx, double* phase);
/* UDV_State contents (udv_state_mod.F90:85-110); side is 'R' or 'L'. */
int alfq_udv_set(alfq_ctx* ctx, int32_t var, const double* U, const double* D, const double* V, char side);
int alfq_udv_get(alfq_ctx* ctx, int32_t var, double* U, double* D, double* V);
/* udv%reset(side) (udv_state_mod.F90:320-345) and dst = src (assign_UDV_state). */
int alfq_udv_reset(alfq_ctx* ctx, int32_t var, char side);
int alfq_udv_assign(alfq_ctx* ctx, int32_t dst, int32_t src);

/* ---- random numbers (replaces ranf_wrap/nranf, random_wrap_mod.F90:135-168) */
/* A tape of uniforms consumed in the reference's order: per type-2 field visit
 * nranf(3), the T0 proposal, the Metropolis test; type 1 skips the first. */
int alfq_rng_set_tape(alfq_ctx* ctx, const double* u, int64_t n);
/* Counter-based Philox4x32-10 stream (seed, stream id). */
int alfq_rng_set_philox(alfq_ctx* ctx, uint64_t seed, uint32_t stream);
int alfq_rng_position(alfq_ctx* ctx, int64_t* pos);
/* Continue the selected chain's Philox stream at position pos (a restart from a saved
 * configuration: the seed line of confout, Fields_mod.F90:739-760, holds seed, stream and
 * position for this engine). */
int alfq_rng_set_position(alfq_ctx* ctx, int64_t pos);

/* ---- the reference routines, same semantics ------------------------------ */
/* WRAPGRUP(GR, NTAU, PHASE, ...) (Prog/Wrapgr_mod.F90:81-156), NTAU in [0, Ltrot-1].
 * phase: in/out complex (2 doubles); NULL keeps the device-resident phase. */
int alfq_wrapgrup(alfq_ctx* ctx, int32_t ntau, double* phase);
/* WRAPGRDO(GR, NTAU, PHASE, ...) (Prog/Wrapgr_mod.F90:160-243), NTAU in [Ltrot, 1]. */
int alfq_wrapgrdo(alfq_ctx* ctx, int32_t ntau, double* phase);
/* The remaining arguments of WRAPGRUP/WRAPGRDO (Prog/Wrapgr_mod.F90:81, 160:
 * Propose_S0, Nt_sequential_start, Nt_sequential_end, N_Global_tau; main.F90:729, 788 passes
 * get_Propose_S0() etc.), checked against what the engine does: every operator of the slice
 * in sequence (start 1, end size(Op_V, 1), lines 115 and 198), no global moves in the slice
 * (N_Global_tau = 0, lines 149 and 190), and no S0-weighted proposals for Ising operators
 * (Propose_S0 with type-1 operators, lines 127-132 and 212; with type-2 operators it changes
 * nothing).  ERROR_HAMILTONIAN with a message when the caller asks for more.  The engine
 * takes S0 = 1 (the Hubbard plugins'). */
int alfq_check_wrap_args(alfq_ctx* ctx, int32_t propose_s0, int32_t nt_sequential_start,
                         int32_t nt_sequential_end, int32_t n_global_tau);
/* WRAPUR(NTAU, NTAU1, UDVR) (Prog/wrapur_mod.F90:37-128) on udvr. */
int alfq_wrapur(alfq_ctx* ctx, int32_t ntau, int32_t ntau1);
/* WRAPUL(NTAU1, NTAU, UDVL) (Prog/wrapul_mod.F90:36-133) on udvl. */
int alfq_wrapul(alfq_ctx* ctx, int32_t ntau1, int32_t ntau);
/* CGR(Z, NVAR, GR, udvr, udvl) (Prog/cgr1_mod.F90:36-449, STAB3): the fresh
 * Green's function replaces GR; the phase Z is returned in z (2 doubles).
 * With check != 0 the replaced GR is compared as in Control_PrecisionG and the
 * stabilisation step of main.F90:744-753 follows: Op_phase, Z**N_SUN,
 * Control_PrecisionP, PHASE = Z. */
int alfq_cgr(alfq_ctx* ctx, int32_t nvar, int32_t var_r, int32_t var_l, int32_t check, double* z);

/* ---- coarse entries ------------------------------------------------------- */
/* main.F90:589-631: fill udvst by WRAPUL, G and PHASE at tau = 0. */
int alfq_init(alfq_ctx* ctx);
/* nsweep iterations of main.F90:714-887 (up pass, down pass, slice-0 refresh). */
int alfq_sweep(alfq_ctx* ctx, int32_t nsweep);
/* Like alfq_sweep, timed with HIP events on the context's stream; *ms = elapsed. */
int alfq_sweep_timed(alfq_ctx* ctx, int32_t nsweep, double* ms);

/* nsweep sweeps of each of count independent chains (contexts), advanced
 * concurrently: one host thread per context enqueues on the context's own
 * stream, so the device overlaps the chains.  *ms (may be NULL) is the wall time.
 * Walker parallelism of the reference (one chain per MPI rank,
 * Documentation/performance.tex:29-40) inside one GPU. */
int alfq_sweep_many(alfq_ctx* const* ctxs, int32_t count, int32_t nsweep, double* ms);

/* ---- diagnostics ---------------------------------------------------------- */
int alfq_sync(alfq_ctx* ctx);
int alfq_control_get(alfq_ctx* ctx, alfq_control* out);
int alfq_control_reset(alfq_ctx* ctx);
/* Per-stabilisation record: (XMAX, XMEAN, |Z - PHASE|) of each checked CGR call. */
int alfq_precision_log(alfq_ctx* ctx, double* out3, int64_t cap, int64_t* count);
const char* alfq_last_error(void);
/* The engine's MFMA product on host arrays (copied in and out), for checking the
 * kernels in isolation: C = alpha op(A) op(B) + beta C with op in 'N', 'T', 'C'; A
 * and/or B may be real (a_is_real/b_is_real), C is complex; prec is an ALFQ_PREC_*
 * arithmetic (ALFQ_PREC_FP32_WRAPS: FP32 operands, products and sums; ALFQ_PREC_INT8_SPLIT:
 * op 'N' only, at least one complex operand). */
int alfq_gemm(int32_t M, int32_t N, int32_t K, char opa, int32_t a_is_real, const double* A, int32_t lda,
              char opb, int32_t b_is_real, const double* B, int32_t ldb, const double* alpha,
              const double* beta, double* C, int32_t ldc, int32_t prec);
/* alfq_gemm without any device allocation (no counterpart in the reference): for callers
 * that cannot allocate device memory (serving loops and harnesses that forbid hipMalloc).
 * Same arguments and result as alfq_gemm; A, B and C are staged through a fixed array of
 * 3 x ALFQ_NOALLOC_OPERAND_BYTES per device that the library holds until the process ends.
 * An operand larger than ALFQ_NOALLOC_OPERAND_BYTES (lda x columns x element size) is
 * rejected with ERROR_GENERIC.  Every call on one device shares that device's array: calls
 * that overlap on one device (from other processes or from code that bypasses this
 * library's lock) race. */
#define ALFQ_NOALLOC_OPERAND_BYTES (8 << 20)
int alfq_gemm_noalloc(int32_t M, int32_t N, int32_t K, char opa, int32_t a_is_real, const double* A,
                      int32_t lda, char opb, int32_t b_is_real, const double* B, int32_t ldb,
                      const double* alpha, const double* beta, double* C, int32_t ldc, int32_t prec);
/* Per-launch timing of the Metropolis chain kernel (start/stop events recorded
 * with each launch on the context's stream).  Enabling resets the record;
 * *total_ms is the summed kernel time of the recorded launches. */
int alfq_kernel_timing(alfq_ctx* ctx, int32_t enable);
int alfq_kernel_timing_get(alfq_ctx* ctx, double* total_ms, int64_t* launches);
/* Whole-sweep replay (no counterpart in the reference, which launches nothing): with
 * enable != 0 (the default) alfq_sweep, alfq_sweep_timed and alfq_sweep_many capture a
 * sweep into a HIP graph the first time its starting state is seen and replay it with
 * one launch afterwards; enable = 0 launches every kernel from the host. */
int alfq_sweep_graphs(alfq_ctx* ctx, int32_t enable);
/* One launch per delayed-update round (no counterpart in the reference): with on != 0 (the
 * default) a context of fewer than 32 chains and Ndim <= 256 forms a round's columns and
 * applies them to G in one kernel that reads G and writes the second G buffer; on = 0 uses
 * an X-product and a flush launch.  The serving loop's lanes use 0: a stream of 500-sweep
 * requests over nine concurrent lanes stalled twice with fused rounds, cause not found. */
int alfq_set_round_fusion(alfq_ctx* ctx, int32_t on);
/* Measured issue-rate ceiling (no counterpart in the reference; measurement only) of
 * kind 0: FP64 MFMA v_mfma_f64_16x16x4f64, 1: FP32 MFMA v_mfma_f32_16x16x4f32,
 * 2: vector FP64 FMA, in TFLOP/s, 3: int8 MFMA v_mfma_i32_32x32x32_i8, in TOP/s, over all
 * CUs of the current device. */
int alfq_mfma_rate(int32_t kind, double* tflops);
/* ---- equal-time measurements (Hubbard plugin, N_FL = 1) ------------------- */
/* Replaces ham%Obser (Prog/Hamiltonians/Hamiltonian_Hubbard_smod.F90:695-801) called by the
 * sweep on the measured slices LOBS_ST..LOBS_EN (main.F90:757-772, 789-802) with Symm = .F.:
 * scalars Kin, Pot, Part, Ener and the correlations Green, SpinZ, Den by distance class,
 * accumulated on the device per chain.  The geometry (Invlist, Latt%imj) is shared by the
 * chains of a batch; the hopping values and ham_U are set for the selected chain. */
typedef struct {
  int32_t ncell;           /* Latt%N */
  int32_t norb;            /* Latt_unit%Norb */
  int32_t nnz;             /* nonzero entries of the hopping matrix O of H_T = c^+ O c */
  int32_t lobs_st;         /* first and last measured slice NTAU1 (1-based) */
  int32_t lobs_en;
  const int32_t* invlist;  /* ncell x norb, column-major: Invlist(I, no), 1-based sites */
  const int32_t* imj;      /* ncell x ncell, column-major: Latt%imj(I, J), 1-based cells */
  const int32_t* o_row;    /* nnz 1-based rows and columns of the nonzero O entries */
  const int32_t* o_col;
  const double* o_val;     /* nnz complex values (re, im) */
  double ham_u;
} alfq_obs_desc;

typedef struct {
  double scal[8];          /* sums of Kin, Pot, Part, Ener times ZP ZS, (re, im) each */
  double sign_scal;        /* Obs_scal%Ave_sign sum, and Obs_scal%N */
  int64_t n_scal;
  double sign_eq;          /* Obs_eq%Ave_sign sum, and Obs_eq%N */
  int64_t n_eq;
} alfq_obs_sums;

int alfq_obs_setup(alfq_ctx* ctx, const alfq_obs_desc* desc);
/* on != 0: measure during alfq_sweep (off after alfq_obs_setup). */
int alfq_obs_enable(alfq_ctx* ctx, int32_t on);
/* Zero every chain's sums (the bin boundary of Print_bin_*). */
int alfq_obs_reset(alfq_ctx* ctx);
/* The selected chain's sums: latt[((q ncell + d) norb + a) norb + b] complex (re, im) with
 * q = 0 Green, 1 SpinZ, 2 Den and d the 0-based distance class; den0[norb] complex. */
int alfq_obs_get(alfq_ctx* ctx, alfq_obs_sums* sums, double* latt, double* den0);
/* ---- time-displaced measurements (Ltau = 1) ------------------------------- */
/* Replaces TAU_M (Prog/tau_m_mod.F90:56-195) with CGR2_2 (Prog/cgr2_2_mod.F90:196-428) and
 * ham%ObserT of the Hubbard plugin (Prog/Hamiltonians/Hamiltonian_Hubbard_smod.F90:826-857),
 * called by the sweep after the slice-0 refresh when Ltau = 1 (main.F90:874-876): G(tau, 0),
 * G(0, tau), G(tau, tau), G(0, 0) for tau = 0..Ltrot, and the correlations Green, SpinZ, Den
 * of tau by distance class, accumulated per chain with the geometry of alfq_obs_setup.
 * alfq_obs_reset clears these sums too. */
typedef struct {
  double sign;              /* Obs_tau%Ave_sign sum and Obs_tau%N (one per TAU_M call) */
  int64_t n;
  int64_t ncg_tau;          /* Control_Precision_tau: NCG_tau, XMAX_tau, XMEAN_tau (sum) */
  double xmax_tau;
  double xmean_tau_sum;
} alfq_tau_sums;

/* on != 0: run TAU_M at the end of every alfq_sweep (needs alfq_obs_setup). */
int alfq_tau_enable(alfq_ctx* ctx, int32_t on);
/* TAU_M now, on the current G, fields and stored left states (after a sweep). */
int alfq_tau_m(alfq_ctx* ctx);
/* The selected chain's sums: latt[(((q (Ltrot + 1) + tau) ncell + d) norb + a) norb + b]
 * complex (re, im), q = 0 Green, 1 SpinZ, 2 Den; den0[norb] complex. */
int alfq_tau_get(alfq_ctx* ctx, alfq_tau_sums* sums, double* latt, double* den0);

/* ---- checkerboard Op_T ---------------------------------------------------- */
/* Replaces the checkerboard branch of Predefined_Hoppings_set_OPT
 * (Prog/Predefined_Hop_mod.F90:1773-1834) as Hop_mod applies it (Prog/Hop_mod.F90:143-250):
 * Op_T(1:nbond) are 2 x 2 bond operators, grouped into nfam families of consecutive bonds
 * whose sites are disjoint (test_checkerboard_decomposition, Predefined_Hop_mod.F90:171-246).
 * After this call the wraps of G and of the UDV chains apply P = E_1 E_2 ... E_N in the
 * reference's bond order (mmthr: E_N first; mmthl: E_1 first; mmthlc: E_N ... E_1) instead
 * of the dense expT of the model description.  The sites and families are shared by the
 * chains of a batch; the matrices are set for the selected chain, and every chain of a
 * batch needs them before it is propagated. */
typedef struct {
  int32_t nbond;            /* size(Op_T, 1) */
  int32_t nfam;             /* families, 1..16 */
  const int32_t* fam_len;   /* bonds per family, in order; they add up to nbond */
  const int32_t* p;         /* 2 * nbond: Op_T(b)%P(1), Op_T(b)%P(2), 1-based sites */
  const double* expo;       /* nbond 2 x 2 complex, column-major: exp(g O_b) as applied
                               (the upper triangle mirrored, OpTTypes_mod.F90:92-133, 203-234) */
  const double* expo_inv;   /* exp(-g O_b), same layout */
  /* Symm = .T.: exp(g O_b / 2), exp(-g O_b / 2) of every bond (mat_1D2, invmat_1D2), same
   * layout, for Hop_mod_Symm: for bonds Ncheck..1, GR <- E_b^1/2 GR E_b^-1/2; the families
   * are then the symmetrised ones (Symmetrize_Families, Predefined_Hop_mod.F90:1422-1495).
   * NULL (both) when the model description has no half steps. */
  const double* expo_half;
  const double* expo_half_inv;
} alfq_checkerboard_desc;

int alfq_set_checkerboard(alfq_ctx* ctx, const alfq_checkerboard_desc* desc);

/* Library build identification. */
const char* alfq_version(void);

#ifdef __cplusplus
}
#endif
#endif /* ALFQ_H */
