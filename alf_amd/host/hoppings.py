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


@dataclass
class CheckerboardOpT:
    """Checkerboard Op_T (Predefined_Hop_mod.F90:1773-1834, Symm = .F.): 2 x 2 bond
    operators in the reference's order, family after family.  ``E[b]`` = exp(g O_b) and
    ``Einv[b]`` = exp(-g O_b) with g = -dtau, symmetrised as RealExpOpT/CmplxExpOpT do;
    on sites ``p1[b], p2[b]``.  The propagation exp(-dtau T) is P = E_1 E_2 ... E_N
    (Hop_mod_mmthr applies E_N first, Prog/Hop_mod.F90:143-160)."""
    p1: np.ndarray
    p2: np.ndarray
    O: np.ndarray             # (N, 2, 2) bond matrices
    E: np.ndarray             # (N, 2, 2)
    Einv: np.ndarray          # (N, 2, 2)
    fam_len: list
    is_real: bool
    # half steps exp(g O_b / 2), exp(-g O_b / 2) (mat_1D2, invmat_1D2) of every bond, used by
    # Hop_mod_Symm before the measurements with Symm = .T.
    Eh: np.ndarray = None
    Eih: np.ndarray = None
    symm: bool = False

    def dense(self, n: int, which: str = "P") -> np.ndarray:
        """P = E_1 ... E_N, Pinv = E_N^-1 ... E_1^-1 or Pc = P^+ = E_N ... E_1 as n x n; with
        Symm = .T. also Q = E_1^1/2 ... E_N^1/2 ("Q") and Q^-1 ("Qinv"), so that Hop_mod_Symm
        (Hop_mod.F90:279-299) is G -> Q G Q^-1."""
        M = np.eye(n, dtype=complex)
        N = len(self.p1)
        if which == "P":
            order, mats = range(N - 1, -1, -1), self.E
        elif which == "Pinv":
            order, mats = range(N), self.Einv
        elif which == "Q":
            order, mats = range(N - 1, -1, -1), self.Eh
        elif which == "Qinv":
            order, mats = range(N), self.Eih
        else:
            order, mats = range(N), self.E
        for b in order:
            r = [int(self.p1[b]), int(self.p2[b])]
            M[r, :] = mats[b] @ M[r, :]
        return M


def checkerboard_families(hop: HoppingMatrix, latt, lattice_type: str):
    """List_Fam of the square (4 families: even/odd cells x the two bonds; commuting for even L1, L2)
    and honeycomb (3 families: bond nf of every cell) defaults
    (Predefined_Hop_mod.F90:326-355, 893-903): lists of (cell, bond)."""
    t = lattice_type.upper()
    if t == "SQUARE":
        fams = [[], [], [], []]
        for I in range(latt.N):
            even = (int(latt.list[I, 0]) + int(latt.list[I, 1])) % 2 == 0
            base = 0 if even else 2
            fams[base].append((I, 0))
            fams[base + 1].append((I, 1))
        return fams
    if t == "HONEYCOMB":
        return [[(I, nf) for I in range(latt.N)] for nf in range(3)]
    raise NotImplementedError(lattice_type)


def symmetrize_families(fams):
    """Symmetrize_Families (Predefined_Hop_mod.F90:1422-1495): families 1..N-1, N, N-1..1 with
    time step 1/2 each and 1 for the middle one, the longest family placed in the middle.
    Returns (families, Prop_Fam)."""
    nc = len(fams)
    order = list(range(nc)) + list(range(nc - 2, -1, -1))
    lmax, fmax = 0, -1
    for f, fam in enumerate(fams):
        if len(fam) > lmax:
            lmax, fmax = len(fam), f
    prop = [0.5] * len(order)
    prop[nc - 1] = 1.0
    if fmax != nc - 1:
        order[nc - 1] = fmax
        order[0] = nc - 1
        order[-1] = nc - 1
    return [fams[f] for f in order], prop


def _bond_sites(hop: HoppingMatrix, Invlist, latt, I: int, nb: int):
    no_i, no_j, n_1, n_2 = (int(v) for v in hop.List[nb])
    return int(Invlist[I, no_i]), int(Invlist[int(latt.nnlist[I, n_1 + 1, n_2 + 1]), no_j])


def test_checkerboard_decomposition(hop: HoppingMatrix, fams, Invlist, latt) -> list:
    """test_checkerboard_decomposition (Predefined_Hop_mod.F90:171-246): every bond in
    exactly one family and the bonds of a family on disjoint sites.  Returns the errors."""
    errs = []
    seen = np.zeros((latt.N, hop.N_bonds), dtype=bool)
    for f, fam in enumerate(fams):
        for I, nb in fam:
            if seen[I, nb]:
                errs.append(f"bond ({I}, {nb}) is present twice")
            seen[I, nb] = True
    for I, nb in zip(*np.nonzero(~seen)):
        errs.append(f"bond ({I}, {nb}) is missing")
    for f, fam in enumerate(fams):
        used = {}
        for I, nb in fam:
            for s in _bond_sites(hop, Invlist, latt, I, nb):
                if s in used:
                    errs.append(f"non-commuting hoppings in family {f}: ({I}, {nb}) and {used[s]} share site {s}")
                used[s] = (I, nb)
    return errs


def checkerboard_opt(hop: HoppingMatrix, Invlist, latt, unit, lattice_type: str, dtau: float,
                     symm: bool = False) -> CheckerboardOpT:
    """The checkerboard branch of Predefined_Hoppings_set_OPT (Predefined_Hop_mod.F90:1773-1834),
    with Symmetrize_Families first when Symm = .T."""
    from .operator import exp_opt_init, op_make, op_set
    fams = checkerboard_families(hop, latt, lattice_type)
    errs = test_checkerboard_decomposition(hop, fams, Invlist, latt)
    if errs:
        raise ValueError("inconsistent checkerboard decomposition: " + "; ".join(errs[:4]))
    prop = [1.0] * len(fams)
    if symm:
        fams, prop = symmetrize_families(fams)
    mult = np.zeros(unit.Norb)
    for nb in range(hop.N_bonds):
        mult[int(hop.List[nb, 0])] += 1
        mult[int(hop.List[nb, 1])] += 1
    p1, p2, Os, Es, Eis, Ehs, Eihs, fam_len = [], [], [], [], [], [], [], []
    real = True
    for fam, pf in zip(fams, prop):
        fam_len.append(len(fam))
        for I, nb in fam:
            no_i, no_j, n_1, n_2 = (int(v) for v in hop.List[nb])
            J = int(latt.nnlist[I, n_1 + 1, n_2 + 1])
            I1, J1 = int(Invlist[I, no_i]), int(Invlist[J, no_j])
            Z = generic_hopping(I, no_i, n_1, n_2, no_j, hop.N_Phi, hop.Phi_X, hop.Phi_Y, hop.Bulk, latt, unit)
            op = op_make(2)
            op.P[:] = [I1, J1]
            op.O[0, 1] = hop.T[nb] * Z
            op.O[1, 0] = np.conj(hop.T[nb] * Z)
            op.O[0, 0] = hop.T_Loc[no_i] / mult[no_i]
            op.O[1, 1] = hop.T_Loc[no_j] / mult[no_j]
            op.g = -dtau * pf + 0j
            op.alpha = 0j
            Os.append(op.O.copy())
            op_set(op)
            ex = exp_opt_init(op)
            real = real and ex.is_real
            p1.append(I1)
            p2.append(J1)
            Es.append(np.array(ex.mat))
            Eis.append(np.array(ex.invmat))
            Ehs.append(np.array(ex.mat_1D2))
            Eihs.append(np.array(ex.invmat_1D2))
    return CheckerboardOpT(np.array(p1), np.array(p2), np.array(Os), np.array(Es), np.array(Eis), fam_len, real,
                           np.array(Ehs), np.array(Eihs), symm)
