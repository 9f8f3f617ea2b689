"""Hopping matrices (dense, non-checkerboard) for the Hubbard plugin.

Follows
  * ``Prog/Predefined_Hop_mod.F90:257-362``   square defaults (two bonds, -t)
  * ``Prog/Predefined_Hop_mod.F90:827-906``   honeycomb defaults (three bonds, -t)
  * ``Prog/Predefined_Hop_mod.F90:1719-1772`` dense Op_T assembly (Checkerboard=.F.)
  * ``Prog/Predefined_Hop_mod.F90:2061-2151`` Generic_hopping: twist and Landau-gauge flux
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .lattices import EPS_SMALL, TWOPI, npbc_b


@dataclass
class HoppingMatrix:
    N_bonds: int
    List: np.ndarray        # (N_bonds, 4): no_1, no_2, n_1, n_2 (orbitals 0-based)
    T: np.ndarray           # (N_bonds,) complex
    T_Loc: np.ndarray       # (Norb,) complex
    N_Phi: int
    Phi_X: float
    Phi_Y: float
    Bulk: bool


def default_hopping_square(ham_t, ham_chem, phi_x, phi_y, bulk, n_phi, latt, unit):
    """Set_Default_hopping_parameters_square (Predefined_Hop_mod.F90:257-362), 2D branch."""
    if abs(latt.L2_p[1] - latt.a2_p[1]) < EPS_SMALL and abs(latt.L2_p[0] - latt.a2_p[0]) < EPS_SMALL:
        raise NotImplementedError("one-leg ladders are outside this build")
    bonds, T = [], []
    if abs(ham_t) > EPS_SMALL:
        bonds = [(0, 0, 0, 1), (0, 0, 1, 0)]
        T = [-ham_t + 0j, -ham_t + 0j]
    return HoppingMatrix(len(bonds), np.array(bonds, dtype=np.int64).reshape(-1, 4),
                         np.array(T, dtype=complex), np.full(unit.Norb, -ham_chem + 0j),
                         n_phi, phi_x, phi_y, bulk)


def default_hopping_honeycomb(ham_t, ham_chem, phi_x, phi_y, bulk, n_phi, latt, unit):
    """Set_Default_hopping_parameters_honeycomb (Predefined_Hop_mod.F90:827-906)."""
    bonds = [(0, 1, 0, 0), (1, 0, 0, 1), (0, 1, 1, -1)]
    T = [-ham_t + 0j] * 3
    return HoppingMatrix(3, np.array(bonds, dtype=np.int64), np.array(T, dtype=complex),
                         np.full(unit.Norb, -ham_chem + 0j), n_phi, phi_x, phi_y, bulk)


def _chi(L_p, X_p, B):
    return -TWOPI * B * L_p[1] * X_p[0]


def generic_hopping(i, no_i, del_1, del_2, no_j, n_phi, flux_1, flux_2, bulk, latt, unit):
    """Peierls phase of the bond (i,no_i) -> (i + del, no_j) (Predefined_Hop_mod.F90:2061-2134)."""
    z = 1.0 + 0.0j
    xj_p = (latt.list[i, 0] + del_1) * latt.a1_p + (latt.list[i, 1] + del_2) * latt.a2_p
    xjp_p, N1, N2 = npbc_b(xj_p, latt.L1_p, latt.L2_p)
    XB_p = N1 * latt.L1_p + N2 * latt.L2_p
    xj_p = xj_p + unit.Orb_pos_p[no_j]
    xjp_p = xjp_p + unit.Orb_pos_p[no_j]
    xi_p = latt.list[i, 0] * latt.a1_p + latt.list[i, 1] * latt.a2_p + unit.Orb_pos_p[no_i]
    del_p = xj_p - xi_p
    A_p = (flux_1 * np.linalg.norm(latt.a1_p) * latt.BZ1_p / np.linalg.norm(latt.L1_p)
           + flux_2 * np.linalg.norm(latt.a2_p) * latt.BZ2_p / np.linalg.norm(latt.L2_p))
    if bulk:
        z *= np.exp(1j * float(A_p @ del_p))
    else:
        z *= np.exp(1j * float(A_p @ XB_p))
    V = abs(latt.L1_p[0] * latt.L2_p[1] - latt.L1_p[1] * latt.L2_p[0])
    if V > EPS_SMALL:
        B = n_phi / V
        z *= np.exp(-1j * TWOPI * B * del_p[0] * (xj_p[1] + xi_p[1]) / 2.0)
        x_p = N2 * latt.L2_p
        x1_p = xjp_p + N1 * latt.L1_p
        z *= np.exp(-1j * _chi(x_p, x1_p, B))
        x_p = N1 * latt.L1_p
        x1_p = xjp_p
        z *= np.exp(-1j * _chi(x_p, x1_p, B))
    return z


def dense_hopping_matrix(hop: HoppingMatrix, Invlist, latt, unit) -> np.ndarray:
    """The n x n single-particle matrix O of the one dense Op_T
    (Predefined_Hop_mod.F90:1719-1772, Checkerboard=.F.)."""
    ndim = latt.N * unit.Norb
    O = np.zeros((ndim, ndim), dtype=complex)
    for I in range(latt.N):
        for nb in range(hop.N_bonds):
            no_i, no_j, n_1, n_2 = (int(v) for v in hop.List[nb])
            J = int(latt.nnlist[I, n_1 + 1, n_2 + 1])
            Z = generic_hopping(I, no_i, n_1, n_2, no_j, hop.N_Phi, hop.Phi_X, hop.Phi_Y, hop.Bulk, latt, unit)
            I1 = int(Invlist[I, no_i])
            J1 = int(Invlist[J, no_j])
            O[I1, J1] = hop.T[nb] * Z
            O[J1, I1] = np.conj(hop.T[nb] * Z)
        for no_i in range(unit.Norb):
            I1 = int(Invlist[I, no_i])
            O[I1, I1] = hop.T_Loc[no_i]
    return O
