"""ctypes binding of the C ABI declared in include/alfq.h.

This is the Python side of the drop-in boundary: the same entry points a
Fortran host binds with ISO_C_BINDING (alf_amd/fortran/alfq_mod.F90).
The library is the in-tree ``alf_amd/libalfq.so`` built for gfx950; there is
no fallback -- loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libalfq.so")

ALFQ_UDVR = -1
ALFQ_UDVL = -2

PREC_FP64, PREC_FP32_WRAPS, PREC_INT8_SPLIT = 0, 1, 2     # alfq_set_precision modes

ERROR_NAMES = {0: "ERROR_NONE", 1: "ERROR_GENERIC", 2: "ERROR_FILE_NOT_FOUND", 3: "ERROR_UNSTABLE_MATRIX",
               4: "ERROR_MISSING_OBS", 5: "ERROR_FIELDS", 6: "ERROR_HAMILTONIAN"}

# every symbol include/alfq.h declares
EXPORTS = [
    "alfq_ctx_create", "alfq_ctx_destroy", "alfq_set_delay",
    "alfq_fields_set", "alfq_fields_get", "alfq_green_set", "alfq_green_get",
    "alfq_phase_set", "alfq_phase_get", "alfq_udv_set", "alfq_udv_get", "alfq_udv_reset", "alfq_udv_assign",
    "alfq_rng_set_tape", "alfq_rng_set_philox", "alfq_rng_position",
    "alfq_wrapgrup", "alfq_wrapgrdo", "alfq_wrapur", "alfq_wrapul", "alfq_cgr",
    "alfq_init", "alfq_sweep", "alfq_sweep_timed", "alfq_sweep_many",
    "alfq_sync", "alfq_control_get", "alfq_control_reset", "alfq_precision_log",
    "alfq_last_error", "alfq_version", "alfq_gemm", "alfq_kernel_timing", "alfq_kernel_timing_get",
    "alfq_sweep_graphs", "alfq_set_round_fusion", "alfq_ctx_create_batch", "alfq_set_chain", "alfq_batch_size",
    "alfq_set_precision", "alfq_mfma_rate", "alfq_chain_bytes",
    "alfq_obs_setup", "alfq_obs_enable", "alfq_obs_reset", "alfq_obs_get", "alfq_rng_set_position",
    "alfq_set_checkerboard", "alfq_tau_enable", "alfq_tau_m", "alfq_tau_get", "alfq_gemm_noalloc", "alfq_chain_load",
    "alfq_set_gemm_policy", "alfq_check_wrap_args",
]


class ModelDesc(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32), ("ltrot", C.c_int32), ("nwrap", C.c_int32), ("n_sun", C.c_int32),
        ("n_fl", C.c_int32), ("n_op", C.c_int32), ("t_is_real", C.c_int32),
        ("expT", C.POINTER(C.c_double)), ("expT_inv", C.POINTER(C.c_double)),
        ("op_P", C.POINTER(C.c_int32)), ("op_type", C.POINTER(C.c_int32)),
        ("op_E", C.POINTER(C.c_double)), ("op_g", C.POINTER(C.c_double)), ("op_alpha", C.POINTER(C.c_double)),
        ("expT_half", C.POINTER(C.c_double)), ("expT_half_inv", C.POINTER(C.c_double)),
    ]


class ObsDesc(C.Structure):
    _fields_ = [("ncell", C.c_int32), ("norb", C.c_int32), ("nnz", C.c_int32), ("lobs_st", C.c_int32),
                ("lobs_en", C.c_int32), ("invlist", C.POINTER(C.c_int32)), ("imj", C.POINTER(C.c_int32)),
                ("o_row", C.POINTER(C.c_int32)), ("o_col", C.POINTER(C.c_int32)),
                ("o_val", C.POINTER(C.c_double)), ("ham_u", C.c_double)]


class CheckerboardDesc(C.Structure):
    _fields_ = [("nbond", C.c_int32), ("nfam", C.c_int32), ("fam_len", C.POINTER(C.c_int32)),
                ("p", C.POINTER(C.c_int32)), ("expo", C.POINTER(C.c_double)), ("expo_inv", C.POINTER(C.c_double)),
                ("expo_half", C.POINTER(C.c_double)), ("expo_half_inv", C.POINTER(C.c_double))]


class TauSums(C.Structure):
    _fields_ = [("sign", C.c_double), ("n", C.c_int64), ("ncg_tau", C.c_int64), ("xmax_tau", C.c_double),
                ("xmean_tau_sum", C.c_double)]


class ObsSums(C.Structure):
    _fields_ = [("scal", C.c_double * 8), ("sign_scal", C.c_double), ("n_scal", C.c_int64),
                ("sign_eq", C.c_double), ("n_eq", C.c_int64)]


class Control(C.Structure):
    _fields_ = [("nc_up", C.c_int64), ("acc_up", C.c_int64), ("ncg", C.c_int64),
                ("xmean_g_sum", C.c_double), ("xmax_g", C.c_double), ("xmax_p", C.c_double)]


class AlfqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libalfq.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() (hipcc, gfx950)")
    lib = C.CDLL(path)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    sig = {
        "alfq_ctx_create": [C.POINTER(ModelDesc), C.POINTER(P)],
        "alfq_ctx_destroy": [P],
        "alfq_set_delay": [P, C.c_int32],
        "alfq_set_precision": [P, C.c_int32],
        "alfq_mfma_rate": [C.c_int32, dp],
        "alfq_chain_bytes": [C.POINTER(ModelDesc), C.POINTER(C.c_int64)],
        "alfq_obs_setup": [P, C.POINTER(ObsDesc)],
        "alfq_rng_set_position": [P, C.c_int64],
        "alfq_set_checkerboard": [P, C.POINTER(CheckerboardDesc)],
        "alfq_tau_enable": [P, C.c_int32],
        "alfq_tau_m": [P],
        "alfq_tau_get": [P, C.POINTER(TauSums), dp, dp],
        "alfq_obs_enable": [P, C.c_int32],
        "alfq_obs_reset": [P],
        "alfq_obs_get": [P, C.POINTER(ObsSums), dp, dp],
        "alfq_fields_set": [P, ip, C.c_int32, C.c_int32],
        "alfq_fields_get": [P, ip, C.c_int32, C.c_int32],
        "alfq_green_set": [P, dp, C.c_int32],
        "alfq_green_get": [P, dp, C.c_int32],
        "alfq_phase_set": [P, dp],
        "alfq_phase_get": [P, dp],
        "alfq_udv_set": [P, C.c_int32, dp, dp, dp, C.c_char],
        "alfq_udv_get": [P, C.c_int32, dp, dp, dp],
        "alfq_udv_reset": [P, C.c_int32, C.c_char],
        "alfq_udv_assign": [P, C.c_int32, C.c_int32],
        "alfq_rng_set_tape": [P, dp, C.c_int64],
        "alfq_rng_set_philox": [P, C.c_uint64, C.c_uint32],
        "alfq_rng_position": [P, C.POINTER(C.c_int64)],
        "alfq_wrapgrup": [P, C.c_int32, dp],
        "alfq_wrapgrdo": [P, C.c_int32, dp],
        "alfq_wrapur": [P, C.c_int32, C.c_int32],
        "alfq_wrapul": [P, C.c_int32, C.c_int32],
        "alfq_cgr": [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, dp],
        "alfq_init": [P],
        "alfq_sweep": [P, C.c_int32],
        "alfq_sweep_timed": [P, C.c_int32, dp],
        "alfq_sweep_many": [C.POINTER(P), C.c_int32, C.c_int32, dp],
        "alfq_sync": [P],
        "alfq_control_get": [P, C.POINTER(Control)],
        "alfq_control_reset": [P],
        "alfq_precision_log": [P, dp, C.c_int64, C.POINTER(C.c_int64)],
        "alfq_kernel_timing": [P, C.c_int32],
        "alfq_sweep_graphs": [P, C.c_int32],
        "alfq_set_round_fusion": [P, C.c_int32],
        "alfq_check_wrap_args": [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32],
        "alfq_ctx_create_batch": [C.POINTER(ModelDesc), C.c_int32, C.POINTER(P)],
        "alfq_set_chain": [P, C.c_int32],
        "alfq_chain_load": [P, C.c_int32, P, C.c_int32],
        "alfq_set_gemm_policy": [P, C.c_int32],
        "alfq_batch_size": [P, C.POINTER(C.c_int32)],
        "alfq_kernel_timing_get": [P, dp, C.POINTER(C.c_int64)],
        "alfq_gemm": [C.c_int32, C.c_int32, C.c_int32, C.c_char, C.c_int32, dp, C.c_int32,
                      C.c_char, C.c_int32, dp, C.c_int32, dp, dp, dp, C.c_int32, C.c_int32],
        "alfq_gemm_noalloc": [C.c_int32, C.c_int32, C.c_int32, C.c_char, C.c_int32, dp, C.c_int32,
                              C.c_char, C.c_int32, dp, C.c_int32, dp, dp, dp, C.c_int32, C.c_int32],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.alfq_last_error.restype = C.c_char_p
    lib.alfq_last_error.argtypes = []
    lib.alfq_version.restype = C.c_char_p
    lib.alfq_version.argtypes = []
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def check(lib, code):
    if code != 0:
        raise AlfqError(code, lib.alfq_last_error().decode())


def model_desc(model):
    """alfq_model_desc of a host model, with the arrays it points to (keep them alive)."""
    ops = model.Op_V
    t_real = bool(model.Op_T.is_real)
    if t_real:
        expT = np.asfortranarray(np.real(model.Op_T.mat), dtype=np.float64)
        expTi = np.asfortranarray(np.real(model.Op_T.invmat), dtype=np.float64)
    else:
        expT = np.asfortranarray(model.Op_T.mat, dtype=np.complex128)
        expTi = np.asfortranarray(model.Op_T.invmat, dtype=np.complex128)
    if any(o.N != 1 for o in ops):
        raise NotImplementedError("the device sweep handles 1x1 interaction operators")
    keep = dict(expT=expT, expTi=expTi,
                P=np.array([int(o.P[0]) + 1 for o in ops], dtype=np.int32),
                type=np.array([o.type for o in ops], dtype=np.int32),
                E=np.array([float(o.E[0]) for o in ops], dtype=np.float64),
                g=np.array([o.g for o in ops], dtype=np.complex128),
                alpha=np.array([o.alpha for o in ops], dtype=np.complex128))
    half = (None, None)
    if getattr(model, "extra", None) and model.extra.get("symm"):
        # Symm = .T.: the half steps mat_1D2 / invmat_1D2 for Hop_mod_Symm before the measurements
        conv = (lambda m: np.asfortranarray(np.real(m), dtype=np.float64)) if t_real else \
            (lambda m: np.asfortranarray(m, dtype=np.complex128))
        keep["expTh"] = conv(model.Op_T.mat_1D2)
        keep["expThi"] = conv(model.Op_T.invmat_1D2)
        half = (_dptr(keep["expTh"]), _dptr(keep["expThi"]))
    d = ModelDesc(model.Ndim, model.Ltrot, model.Nwrap, model.N_SUN, model.N_FL, len(ops), int(t_real),
                  _dptr(keep["expT"]), _dptr(keep["expTi"]), _iptr(keep["P"]), _iptr(keep["type"]),
                  _dptr(keep["E"]), _dptr(keep["g"]), _dptr(keep["alpha"]), *half)
    return d, keep


def checkerboard_desc(model):
    """alfq_checkerboard_desc of a host model built with Checkerboard = True (None otherwise),
    with the arrays it points to."""
    cb = model.extra.get("checkerboard") if getattr(model, "extra", None) else None
    if cb is None:
        return None, None
    keep = dict(fam=np.asarray(cb.fam_len, dtype=np.int32),
                p=np.ascontiguousarray(np.stack([cb.p1, cb.p2], axis=1) + 1, dtype=np.int32),
                E=np.ascontiguousarray(np.transpose(np.asarray(cb.E, dtype=np.complex128), (0, 2, 1))),
                Ei=np.ascontiguousarray(np.transpose(np.asarray(cb.Einv, dtype=np.complex128), (0, 2, 1))))
    half = (None, None)
    if cb.symm:
        keep["Eh"] = np.ascontiguousarray(np.transpose(np.asarray(cb.Eh, dtype=np.complex128), (0, 2, 1)))
        keep["Eih"] = np.ascontiguousarray(np.transpose(np.asarray(cb.Eih, dtype=np.complex128), (0, 2, 1)))
        half = (_dptr(keep["Eh"]), _dptr(keep["Eih"]))
    d = CheckerboardDesc(len(cb.p1), len(cb.fam_len), _iptr(keep["fam"]), _iptr(keep["p"]), _dptr(keep["E"]),
                         _dptr(keep["Ei"]), *half)
    return d, keep


def obs_desc(model, lobs=None):
    """alfq_obs_desc of a host Hubbard model (hopping matrix, lattice tables, U) with the
    arrays it points to; lobs = (LOBS_ST, LOBS_EN), default every slice 1..Ltrot."""
    from alf_amd.host import lattices
    latt, unit = model.latt, model.unit
    O = np.asarray(model.hop_matrix, dtype=np.complex128)
    rows, cols = np.nonzero(O)
    imj = lattices.distance_classes(latt) + 1
    keep = dict(invlist=np.asfortranarray(np.asarray(model.Invlist) + 1, dtype=np.int32),
                imj=np.asfortranarray(imj, dtype=np.int32),
                row=(rows + 1).astype(np.int32), col=(cols + 1).astype(np.int32),
                val=np.ascontiguousarray(O[rows, cols], dtype=np.complex128))
    st, en = lobs if lobs is not None else (1, model.Ltrot)
    d = ObsDesc(latt.N, unit.Norb, len(rows), int(st), int(en), _iptr(keep["invlist"]), _iptr(keep["imj"]),
                _iptr(keep["row"]), _iptr(keep["col"]), _dptr(keep["val"]), float(model.params.Ham_U))
    return d, keep


_TUNED = {}     # autotuner results per (Ndim, Ltrot, chains, real T) for this process


class Engine:
    """One Markov chain on the GPU behind the C ABI.

    ``model`` is the host plugin object (alf_amd.host.hamiltonian.HamBase).
    """

    def __init__(self, model, delay: int | None = None):
        self.lib = load_library()
        self.model = model
        d, self._keep = model_desc(model)
        self._desc = d
        h = C.c_void_p()
        check(self.lib, self.lib.alfq_ctx_create(C.byref(d), C.byref(h)))
        self.ctx = h
        self.n = model.Ndim
        self.nchain = 1
        self.set_checkerboard(model)
        if delay is not None:
            self.set_delay(delay)

    def set_checkerboard(self, model=None):
        """Hand the bond operators of a checkerboard model to the selected chain (no-op for
        a model with the dense hopping exponential)."""
        d, keep = checkerboard_desc(model if model is not None else self.model)
        if d is not None:
            self._c(self.lib.alfq_set_checkerboard(self.ctx, C.byref(d)))
        del keep

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.alfq_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, code):
        check(self.lib, code)

    # ---- state ----
    def set_delay(self, k):
        self._c(self.lib.alfq_set_delay(self.ctx, int(k)))
        self.delay = int(k)

    def set_gemm_policy(self, policy):
        self._c(self.lib.alfq_set_gemm_policy(self.ctx, int(policy)))
        self.gemm_policy = int(policy)

    def autotune(self, delays=(8, 16, 32), policies=(0, 1, 2, 3), timed=2):
        """Measured choice of the delayed-update depth k and the GEMM tiling policy for this
        context's shape (Ndim, Ltrot, chains), cached per shape for the process.

        Every trial runs real sweeps of the chains (the trajectory does not depend on k, and
        the tiling changes only the rounding of the products), so tuning costs no wasted
        work: per candidate two sweeps capture the sweep graphs of both buffer parities, then
        `timed` replayed sweeps are timed with HIP events.  k is tuned first at policy 0, then
        the policy at the best k.  Needs a context after init with a random stream."""
        key = (self.model.Ndim, self.model.Ltrot, getattr(self, "nchain", 1), bool(self.model.Op_T.is_real))
        hit = _TUNED.get(key)
        if hit is not None:
            self.set_delay(hit["delay"])
            self.set_gemm_policy(hit["policy"])
            return hit

        def trial(k, pol):
            self.set_delay(k)
            self.set_gemm_policy(pol)
            self.sweep(2)
            return self.sweep_timed(timed) / timed

        table = {}
        best_k, best_t = None, float("inf")
        for k in delays:
            t = trial(k, 0)
            table[f"k={k},policy=0"] = t
            if t < best_t:
                best_k, best_t = k, t
        best_p = 0
        for pol in policies:
            if pol == 0:
                continue
            t = trial(best_k, pol)
            table[f"k={best_k},policy={pol}"] = t
            if t < best_t:
                best_p, best_t = pol, t
        self.set_delay(best_k)
        self.set_gemm_policy(best_p)
        hit = {"delay": best_k, "policy": best_p, "ms_per_sweep": best_t, "trials_ms_per_sweep": table}
        _TUNED[key] = hit
        return hit

    # ---- equal-time measurements ----
    def obs_setup(self, model=None, lobs=None):
        """Measurement tables of `model` (default: this engine's) for the selected chain."""
        d, keep = obs_desc(model if model is not None else self.model, lobs)
        self._c(self.lib.alfq_obs_setup(self.ctx, C.byref(d)))
        self._obs_shape = (int(d.ncell), int(d.norb))
        del keep

    def obs_enable(self, on=True):
        self._c(self.lib.alfq_obs_enable(self.ctx, int(bool(on))))

    # ---- time-displaced measurements (Ltau = 1) ----
    def tau_enable(self, on=True):
        """TAU_M after every sweep (needs obs_setup)."""
        self._c(self.lib.alfq_tau_enable(self.ctx, int(bool(on))))

    def tau_m(self):
        """TAU_M now, after a sweep."""
        self._c(self.lib.alfq_tau_m(self.ctx))

    def tau_get(self):
        """The selected chain's time-displaced sums: latt[3 (Green, SpinZ, Den), Ltrot + 1,
        ncell, norb, norb], den0[norb], the count and sign sum, and Precision tau."""
        nc, no = self._obs_shape
        nt = self.model.Ltrot + 1
        s = TauSums()
        latt = np.zeros(3 * nt * nc * no * no, dtype=np.complex128)
        den0 = np.zeros(no, dtype=np.complex128)
        self._c(self.lib.alfq_tau_get(self.ctx, C.byref(s), _dptr(latt), _dptr(den0)))
        return {"latt": latt.reshape(3, nt, nc, no, no), "den0": den0, "n": int(s.n), "sign": float(s.sign),
                "ncg_tau": int(s.ncg_tau), "xmax_tau": float(s.xmax_tau), "xmean_tau_sum": float(s.xmean_tau_sum)}

    def obs_reset(self):
        self._c(self.lib.alfq_obs_reset(self.ctx))

    def obs_get(self):
        """The selected chain's measurement sums (Obs_scal / Obs_eq of the reference)."""
        nc, no = self._obs_shape
        s = ObsSums()
        latt = np.zeros(3 * nc * no * no, dtype=np.complex128)
        den0 = np.zeros(no, dtype=np.complex128)
        self._c(self.lib.alfq_obs_get(self.ctx, C.byref(s), _dptr(latt), _dptr(den0)))
        sc = np.array(s.scal[:]).view(np.complex128)
        return {"scal": sc, "sign_scal": s.sign_scal, "n_scal": s.n_scal, "latt": latt.reshape(3, nc, no, no),
                "den0": den0, "sign_eq": s.sign_eq, "n_eq": s.n_eq}

    def set_precision(self, mode):
        """PREC_FP64 (drop-in), PREC_FP32_WRAPS (hopping wraps of G on FP32 MFMA) or
        PREC_INT8_SPLIT (the same wraps as exact int8 digit products summed in FP64)."""
        self._c(self.lib.alfq_set_precision(self.ctx, int(mode)))

    def set_fields(self, f):
        a = np.asfortranarray(f, dtype=np.int32)
        self._c(self.lib.alfq_fields_set(self.ctx, _iptr(a), a.shape[0], a.shape[1]))

    def get_fields(self):
        m = self.model
        a = np.zeros((len(m.Op_V), m.Ltrot), dtype=np.int32, order="F")
        self._c(self.lib.alfq_fields_get(self.ctx, _iptr(a), a.shape[0], a.shape[1]))
        return a

    def set_green(self, G):
        a = np.asfortranarray(G, dtype=np.complex128)
        self._c(self.lib.alfq_green_set(self.ctx, _dptr(a), self.n))

    def get_green(self):
        a = np.zeros((self.n, self.n), dtype=np.complex128, order="F")
        self._c(self.lib.alfq_green_get(self.ctx, _dptr(a), self.n))
        return a

    def set_phase(self, z):
        a = np.array([complex(z).real, complex(z).imag])
        self._c(self.lib.alfq_phase_set(self.ctx, _dptr(a)))

    def get_phase(self):
        a = np.zeros(2)
        self._c(self.lib.alfq_phase_get(self.ctx, _dptr(a)))
        return complex(a[0], a[1])

    def set_udv(self, var, U, D, V, side):
        Ua = np.asfortranarray(U, dtype=np.complex128)
        Va = np.asfortranarray(V, dtype=np.complex128)
        Da = np.ascontiguousarray(D, dtype=np.complex128)
        self._c(self.lib.alfq_udv_set(self.ctx, var, _dptr(Ua), _dptr(Da),
                                      _dptr(Va), side.encode()))

    def get_udv(self, var):
        n = self.n
        U = np.zeros((n, n), dtype=np.complex128, order="F")
        V = np.zeros((n, n), dtype=np.complex128, order="F")
        D = np.zeros(n, dtype=np.complex128)
        self._c(self.lib.alfq_udv_get(self.ctx, var, _dptr(U), _dptr(D),
                                      _dptr(V)))
        return U, D, V

    def udv_reset(self, var, side):
        self._c(self.lib.alfq_udv_reset(self.ctx, var, side.encode()))

    def udv_assign(self, dst, src):
        self._c(self.lib.alfq_udv_assign(self.ctx, dst, src))

    def set_tape(self, u):
        a = np.ascontiguousarray(u, dtype=np.float64)
        self._c(self.lib.alfq_rng_set_tape(self.ctx, _dptr(a), a.size))

    def set_philox(self, seed, stream=0):
        self._c(self.lib.alfq_rng_set_philox(self.ctx, C.c_uint64(seed), C.c_uint32(stream)))

    def rng_position(self):
        p = C.c_int64()
        self._c(self.lib.alfq_rng_position(self.ctx, C.byref(p)))
        return p.value

    def set_rng_position(self, pos):
        """Continue the Philox stream at `pos` (restart from a saved configuration)."""
        self._c(self.lib.alfq_rng_set_position(self.ctx, int(pos)))

    # ---- reference routines ----
    def wrapgrup(self, ntau, phase=None):
        a = None if phase is None else np.array([complex(phase).real, complex(phase).imag])
        self._c(self.lib.alfq_wrapgrup(self.ctx, ntau, _dptr(a) if a is not None else None))
        return None if a is None else complex(a[0], a[1])

    def wrapgrdo(self, ntau, phase=None):
        a = None if phase is None else np.array([complex(phase).real, complex(phase).imag])
        self._c(self.lib.alfq_wrapgrdo(self.ctx, ntau, _dptr(a) if a is not None else None))
        return None if a is None else complex(a[0], a[1])

    def wrapur(self, ntau, ntau1):
        self._c(self.lib.alfq_wrapur(self.ctx, ntau, ntau1))

    def wrapul(self, ntau1, ntau):
        self._c(self.lib.alfq_wrapul(self.ctx, ntau1, ntau))

    def cgr(self, nvar, var_r=ALFQ_UDVR, var_l=ALFQ_UDVL, check_precision=False):
        z = np.zeros(2)
        self._c(self.lib.alfq_cgr(self.ctx, nvar, var_r, var_l, int(check_precision), _dptr(z)))
        return complex(z[0], z[1])

    # ---- coarse ----
    def init(self):
        self._c(self.lib.alfq_init(self.ctx))

    def sweep(self, nsweep=1):
        self._c(self.lib.alfq_sweep(self.ctx, nsweep))

    def sweep_timed(self, nsweep=1):
        ms = C.c_double()
        self._c(self.lib.alfq_sweep_timed(self.ctx, nsweep, C.byref(ms)))
        return ms.value

    def sync(self):
        self._c(self.lib.alfq_sync(self.ctx))

    def control(self):
        c = Control()
        self._c(self.lib.alfq_control_get(self.ctx, C.byref(c)))
        return {"NC_up": c.nc_up, "ACC_up": c.acc_up, "NCG": c.ncg, "XMEANG": c.xmean_g_sum,
                "XMAXG": c.xmax_g, "XMAXP": c.xmax_p,
                "precision_green_mean": (c.xmean_g_sum / c.ncg) if c.ncg else 0.0}

    def sweep_graphs(self, enable=True):
        """Replay sweeps as captured HIP graphs (default) or launch every kernel from the host."""
        self._c(self.lib.alfq_sweep_graphs(self.ctx, int(enable)))

    def check_wrap_args(self, propose_s0=False, nt_sequential_start=1, nt_sequential_end=None, n_global_tau=0):
        """The trailing WRAPGRUP/WRAPGRDO arguments of the reference (Propose_S0,
        Nt_sequential_start/end, N_Global_tau) checked against what the engine does
        (alfq_check_wrap_args): raises AlfqError (ERROR_HAMILTONIAN) for anything else."""
        end = len(self.model.Op_V) if nt_sequential_end is None else nt_sequential_end
        self._c(self.lib.alfq_check_wrap_args(self.ctx, int(bool(propose_s0)), int(nt_sequential_start), int(end),
                                              int(n_global_tau)))

    def round_fusion(self, on=True):
        """One launch per delayed-update round for contexts of fewer than 32 chains (default)
        or an X-product and a flush launch (alfq_set_round_fusion)."""
        self._c(self.lib.alfq_set_round_fusion(self.ctx, int(on)))

    def kernel_timing(self, enable=True):
        self._c(self.lib.alfq_kernel_timing(self.ctx, int(enable)))

    def kernel_timing_get(self):
        ms = C.c_double()
        n = C.c_int64()
        self._c(self.lib.alfq_kernel_timing_get(self.ctx, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def control_reset(self):
        self._c(self.lib.alfq_control_reset(self.ctx))

    def precision_log(self, cap=1 << 16):
        out = np.zeros(3 * cap)
        cnt = C.c_int64()
        self._c(self.lib.alfq_precision_log(self.ctx, _dptr(out), cap, C.byref(cnt)))
        k = min(cnt.value, cap)
        return out[:3 * k].reshape(k, 3)


def gemm(A, B, opa="N", opb="N", alpha=1.0, beta=0.0, Cin=None, prec=PREC_FP64):
    """C = alpha op(A) op(B) + beta C through the engine's MFMA kernels (alfq_gemm)."""
    lib = load_library()
    a_real = not np.iscomplexobj(A)
    b_real = not np.iscomplexobj(B)
    Aa = np.asfortranarray(A, dtype=np.float64 if a_real else np.complex128)
    Ba = np.asfortranarray(B, dtype=np.float64 if b_real else np.complex128)
    M = Aa.shape[0] if opa == "N" else Aa.shape[1]
    K = Aa.shape[1] if opa == "N" else Aa.shape[0]
    N = Ba.shape[1] if opb == "N" else Ba.shape[0]
    Cc = np.zeros((M, N), dtype=np.complex128, order="F") if Cin is None else np.array(Cin, dtype=np.complex128, order="F")
    al = np.array([complex(alpha).real, complex(alpha).imag])
    be = np.array([complex(beta).real, complex(beta).imag])
    check(lib, lib.alfq_gemm(M, N, K, opa.encode(), int(a_real), _dptr(Aa), max(Aa.shape[0], 1), opb.encode(),
                             int(b_real), _dptr(Ba), max(Ba.shape[0], 1), _dptr(al), _dptr(be), _dptr(Cc), max(M, 1),
                             int(prec)))
    return Cc


class BatchEngine(Engine):
    """Several independent Markov chains advancing in lockstep in one context: every
    kernel launch covers all of them.  The per-chain accessors inherited from Engine
    (fields, G, phase, random stream, control, ...) act on the chain selected with
    ``chain(z)``; sweeps act on every chain."""

    def __init__(self, models, delay: int | None = None):
        self.lib = load_library()
        self.models = list(models)
        self.model = self.models[0]
        descs, self._keep = [], []
        for m in self.models:
            d, k = model_desc(m)
            descs.append(d)
            self._keep.append(k)
        arr = (ModelDesc * len(descs))(*descs)
        self._descs = arr
        h = C.c_void_p()
        check(self.lib, self.lib.alfq_ctx_create_batch(arr, len(descs), C.byref(h)))
        self.ctx = h
        self.n = self.model.Ndim
        self.nchain = len(descs)
        for z, m in enumerate(self.models):
            self.chain(z)
            self.set_checkerboard(m)
        self.chain(0)
        if delay is not None:
            self.set_delay(delay)

    def obs_setup(self, model=None, lobs=None):
        """Measurement tables for every chain of the batch, each with its own model."""
        for z, m in enumerate(self.models):
            self.chain(z)
            Engine.obs_setup(self, m, lobs)
        self.chain(0)

    def chain(self, z):
        """Select chain z for the per-chain accessors; returns self for chaining calls."""
        self._c(self.lib.alfq_set_chain(self.ctx, int(z)))
        self._cur = int(z)
        return self

    def get_fields(self):
        """The selected chain's fields, n_op x its own Ltrot (chains of a batch may differ in Ltrot)."""
        m = self.models[getattr(self, "_cur", 0)]
        a = np.zeros((len(m.Op_V), m.Ltrot), dtype=np.int32, order="F")
        self._c(self.lib.alfq_fields_get(self.ctx, _iptr(a), a.shape[0], a.shape[1]))
        return a

    def load_chain(self, z, src, src_chain=0):
        """Chain src_chain of the context `src` (same shape, initialised) replaces chain z
        (alfq_chain_load): a new chain joins the running batch at a sweep boundary."""
        self._c(self.lib.alfq_chain_load(self.ctx, int(z), src.ctx, int(src_chain)))
        self.models[z] = src.models[src_chain] if hasattr(src, "models") else src.model


def sweep_many(engines, nsweep=1):
    """Advance several independent chains concurrently (alfq_sweep_many); returns wall ms."""
    lib = load_library()
    arr = (C.c_void_p * len(engines))(*[e.ctx for e in engines])
    ms = C.c_double()
    check(lib, lib.alfq_sweep_many(arr, len(engines), nsweep, C.byref(ms)))
    return ms.value


def mfma_rate(kind):
    """Measured ceiling: 0 FP64 MFMA, 1 FP32 MFMA, 2 vector FP64 FMA (TFLOP/s), 3 int8 MFMA
    (TOP/s) (alfq_mfma_rate)."""
    lib = load_library()
    v = C.c_double()
    check(lib, lib.alfq_mfma_rate(int(kind), C.byref(v)))
    return v.value


def chain_bytes(model):
    """Device bytes of one chain of `model` in a batch (alfq_chain_bytes)."""
    lib = load_library()
    d, keep = model_desc(model)
    v = C.c_int64()
    check(lib, lib.alfq_chain_bytes(C.byref(d), C.byref(v)))
    del keep
    return int(v.value)
