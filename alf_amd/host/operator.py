"""The Operator plugin type and the hopping exponentials.

Host-side mirror of ALF's ``Type(Operator)``: the model plugin fills
``N, P, O, g, alpha, type`` and ``op_set`` derives ``E, U, diag, N_non_zero``
and the exponential tables, exactly what the device sweep consumes.

Follows
  * ``Prog/Operator_mod.F90:56-90``   (Type Operator)
  * ``Prog/Operator_mod.F90:193-212`` (Op_make)
  * ``Prog/Operator_mod.F90:259-472`` (Op_set: Hermiticity check, diagonalisation,
    non-zero eigenvalues first, det(U)=1, E_exp tables for types 1 and 2)
  * ``Prog/Operator_mod.F90:491-529`` (Op_exp with Kahan summation)
  * ``Prog/Operator_mod.F90:1103-1120`` (Op_is_real)
  * ``Prog/OpTTypes_mod.F90:92-133, 203-234`` (exp(+-g T), exp(+-g T/2), upper-triangle
    Hermitian symmetrisation)
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .fields import phi_st

EPS_MACHINE = 1.0e-12    # natural_constants_mod.F90:8
EPS_SMALL = 1.0e-10      # natural_constants_mod.F90:9


@dataclass
class Operator:
    N: int
    P: np.ndarray                      # (N,) 0-based sites
    O: np.ndarray                      # (N, N) complex
    g: complex = 0j
    alpha: complex = 0j
    type: int = 0
    flip_protocol: int = 1
    diag: bool = False
    N_non_zero: int = 0
    E: np.ndarray = field(default_factory=lambda: np.zeros(0))
    U: np.ndarray = field(default_factory=lambda: np.zeros((0, 0), complex))
    E_exp: dict = field(default_factory=dict)    # sp -> (N,) complex, sp in -type..type


def op_make(N: int) -> Operator:
    """Op_make (Operator_mod.F90:193-212)."""
    return Operator(N=N, P=np.zeros(N, dtype=np.int64), O=np.zeros((N, N), dtype=complex), N_non_zero=N)


def op_set(op: Operator) -> Operator:
    """Op_set (Operator_mod.F90:259-472)."""
    if op.type < 0 or op.type > 4:
        raise ValueError(f"Op_set: invalid operator type {op.type}")
    if np.any(op.P < 0):
        raise ValueError("Op_set: projector index out of bounds")
    N = op.N
    O = op.O
    op.diag = True
    if N > 1:
        for i in range(N):
            for j in range(i + 1, N):
                dev = abs(O[i, j] - np.conj(O[j, i]))
                if dev > EPS_MACHINE * max(abs(O[i, j]), abs(O[j, i]), 1e-30):
                    raise ValueError("Op_set: operator matrix is not Hermitian")
                if op.diag and O[i, j] != 0:
                    op.diag = False
        for i in range(N):
            if abs(O[i, i].imag) > EPS_MACHINE * max(abs(O[i, i]), 1e-30):
                raise ValueError("Op_set: complex diagonal element")
    elif abs(O[0, 0].imag) > EPS_MACHINE * max(abs(O[0, 0]), 1e-30):
        raise ValueError("Op_set: 1x1 operator with non-zero imaginary part")

    op.E = np.zeros(N)
    op.U = np.zeros((N, N), dtype=complex)
    if N > 1:
        if op.diag:
            op.E = np.real(np.diag(O)).copy()
            op.U = np.eye(N, dtype=complex)
            op.N_non_zero = N
        else:
            E, U = np.linalg.eigh(O)
            np_, nz = 0, 0
            for i in range(N):
                if abs(E[i]) > EPS_SMALL:
                    op.U[:, np_] = U[:, i]
                    op.E[np_] = E[i]
                    np_ += 1
                else:
                    op.U[:, N - 1 - nz] = U[:, i]
                    op.E[N - 1 - nz] = E[i]
                    nz += 1
            op.N_non_zero = np_
            z = np.linalg.det(op.U)
            if abs(z) < 1e-30:
                raise ValueError("Op_set: eigenvector matrix has near-zero determinant")
            op.U[:, 0] = op.U[:, 0] / z
    else:
        op.E[0] = O[0, 0].real
        op.U[0, 0] = 1.0
        op.N_non_zero = 1
        op.diag = True

    op.E_exp = {}
    if op.type in (1, 2):
        for I in range(1, op.type + 1):
            pos = np.ones(N, dtype=complex)
            neg = np.ones(N, dtype=complex)
            for n in range(op.N_non_zero):
                pos[n] = np.exp(op.g * op.E[n] * phi_st(I, op.type))
                neg[n] = 1.0 / pos[n]
            op.E_exp[I] = pos
            op.E_exp[-I] = neg
    return op


def op_exp(g: complex, op: Operator) -> np.ndarray:
    """Op_exp (Operator_mod.F90:491-529): exp(g O) = U diag(exp(g E)) U^dagger with a
    Kahan-compensated accumulation over eigen-components."""
    N = op.N
    if op.diag:
        return np.diag(np.exp(g * op.E).astype(complex))
    mat = np.zeros((N, N), dtype=complex)
    c = np.zeros((N, N), dtype=complex)
    for n in range(N):
        z = np.exp(g * op.E[n])
        y = np.outer(op.U[:, n], z * np.conj(op.U[:, n])) - c
        t = mat + y
        c = (t - mat) - y
        mat = t
    return mat


def op_is_real(op: Operator) -> bool:
    """Op_is_real (Operator_mod.F90:1103-1120)."""
    ret = abs(np.imag(op.g)) < abs(op.g) * np.finfo(float).eps
    myzero = np.max(np.abs(op.E)) * np.finfo(float).eps if op.E.size else 0.0
    return bool(ret and np.all(np.abs(np.imag(op.O)) < myzero))


@dataclass
class ExpOpT:
    """Stored hopping exponentials of one Op_T (OpTTypes_mod.F90:50-89).

    ``mat = exp(g O)``, ``invmat = exp(-g O)`` and the half-step versions, each
    Hermitian-symmetrised from its upper triangle, as the reference's
    ZDSLSYMM/ZSLHEMM(..., 'U', ...) calls only read that triangle.  Stored here as
    full dense n x n arrays in the embedding given by ``P``.
    """
    is_real: bool
    P: np.ndarray
    mat: np.ndarray
    invmat: np.ndarray
    mat_1D2: np.ndarray
    invmat_1D2: np.ndarray
    g: complex


def _symmetrise_upper(m: np.ndarray, real: bool) -> np.ndarray:
    """Upper triangle (i<=j) := (m_ij + conj(m_ji))/2, lower := conj(upper)^T."""
    n = m.shape[0]
    out = m.copy()
    iu = np.triu_indices(n)
    if real:
        out[iu] = (m[iu] + m.T[iu]) / 2.0
    else:
        out[iu] = (m[iu] + np.conj(m.T[iu])) / 2.0
    upper = np.triu(out)
    full = upper + np.conj(np.triu(out, 1)).T
    if not real:
        # the HEMM convention reads only the real part of the diagonal
        full[np.diag_indices(n)] = np.real(np.diag(full))
    return full


def exp_opt_init(op_t: Operator) -> ExpOpT:
    """RealExpOpT_init / CmplxExpOpT_init (OpTTypes_mod.F90:92-133, 203-234)."""
    real = op_is_real(op_t)
    g = op_t.g
    invmat = op_exp(-g, op_t)
    mat = op_exp(g, op_t)
    invmat2 = op_exp(-g / 2.0, op_t)
    mat2 = op_exp(g / 2.0, op_t)
    if real:
        mats = [np.real(m).astype(complex) for m in (mat, invmat, mat2, invmat2)]
    else:
        mats = [mat, invmat, mat2, invmat2]
    mats = [_symmetrise_upper(m, real) for m in mats]
    return ExpOpT(real, op_t.P.copy(), mats[0], mats[1], mats[2], mats[3], g)
