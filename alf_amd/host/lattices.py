"""Bravais lattices and unit cells for the Hubbard plugin.

Host-side set-up only: it fixes the site order, which in turn fixes the
layout of every n x n matrix the sweep engine holds on the device.

Follows
  * ``Libraries/Modules/lattices_v3_mod.F90:88-336``  (Make_lattice)
  * ``Libraries/Modules/lattices_v3_mod.F90:384-460`` (npbc_R, npbc_R_B)
  * ``Prog/Predefined_Latt_mod.F90:100-262``           (Predefined_Latt)

Sites are enumerated over the Wigner-Seitz cell spanned by L1_p, L2_p with
i1 outer and i2 inner; orbitals are the inner index of the flat site list.
All indices here are 0-based (the reference is 1-based).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

EPS_SMALL = 1.0e-10      # natural_constants_mod.F90:9
TWOPI = 2.0 * np.pi


@dataclass
class Lattice:
    L1_p: np.ndarray
    L2_p: np.ndarray
    a1_p: np.ndarray
    a2_p: np.ndarray
    BZ1_p: np.ndarray
    BZ2_p: np.ndarray
    b1_p: np.ndarray
    b2_p: np.ndarray
    N: int
    L: int
    list: np.ndarray          # (N, 2) integer coordinates (i1, i2) of each cell
    invlist: dict             # (i1, i2) -> 0-based cell index
    nnlist: np.ndarray        # (N, 3, 3) neighbour cell index for (d1+1, d2+1)


@dataclass
class UnitCell:
    Norb: int
    N_coord: int
    Orb_pos_p: np.ndarray     # (Norb, 2)


def _in_ws_cell(x_p, L1_p, L2_p):
    for a_p in (L2_p, L1_p, L2_p - L1_p, L2_p + L1_p):
        s = float(x_p @ a_p)
        h = float(a_p @ a_p) / 2.0
        if not (s <= h + EPS_SMALL and s >= -h + EPS_SMALL):
            return False
    return True


def npbc(n_p, L1_p, L2_p):
    """Fold a position into the Wigner-Seitz cell (lattices_v3_mod.F90:384-411)."""
    nr = np.array(n_p, dtype=float)
    for x_p in (L2_p, L1_p, L2_p - L1_p, L2_p + L1_p):
        X = float(nr @ x_p) / float(x_p @ x_p)
        if X >= 0.5 + EPS_SMALL:
            nr = nr - x_p
        if X <= -0.5 + EPS_SMALL:
            nr = nr + x_p
    return nr


def npbc_b(n_p, L1_p, L2_p):
    """npbc that also counts the boundary crossings N1, N2 with
    n_p = nr_p + N1*L1_p + N2*L2_p (lattices_v3_mod.F90:414-460)."""
    nr = np.array(n_p, dtype=float)
    N1 = N2 = 0
    for d1, d2 in ((0, 1), (1, 0), (-1, 1), (1, 1)):
        x_p = d2 * L2_p + d1 * L1_p
        X = float(nr @ x_p) / float(x_p @ x_p)
        if X >= 0.5 + EPS_SMALL:
            nr = nr - x_p
            N2 += d2
            N1 += d1
        if X <= -0.5 + EPS_SMALL:
            nr = nr + x_p
            N2 -= d2
            N1 -= d1
    return nr, N1, N2


def make_lattice(L1_p, L2_p, a1_p, a2_p) -> Lattice:
    """Make_lattice (lattices_v3_mod.F90:88-336), real-space part."""
    L1_p = np.asarray(L1_p, float)
    L2_p = np.asarray(L2_p, float)
    a1_p = np.asarray(a1_p, float)
    a2_p = np.asarray(a2_p, float)
    M = np.array([[a1_p[0], a1_p[1]], [a2_p[0], a2_p[1]]])
    X = M[0, 0] * M[1, 1] - M[1, 0] * M[0, 1]
    Minv = np.array([[M[1, 1] / X, -M[0, 1] / X], [-M[1, 0] / X, M[0, 0] / X]])
    BZ1_p = TWOPI * np.array([Minv[0, 0], Minv[1, 0]])
    BZ2_p = TWOPI * np.array([Minv[0, 1], Minv[1, 1]])
    Xk = TWOPI / (float(BZ1_p @ L1_p) * float(BZ2_p @ L2_p) - float(BZ2_p @ L1_p) * float(BZ1_p @ L2_p))
    Xk = abs(Xk)
    b1_p = Xk * (float(BZ2_p @ L2_p) * BZ1_p - float(BZ1_p @ L2_p) * BZ2_p)
    b2_p = Xk * (float(BZ1_p @ L1_p) * BZ2_p - float(BZ2_p @ L1_p) * BZ1_p)

    L = 0
    for bz, lv in ((BZ1_p, L1_p), (BZ2_p, L1_p), (BZ1_p, L2_p), (BZ2_p, L2_p)):
        L = max(L, abs(int(np.rint(float(bz @ lv) / TWOPI))))

    cells = []
    invlist = {}
    for i1 in range(-L, L + 1):
        for i2 in range(-L, L + 1):
            x_p = i1 * a1_p + i2 * a2_p
            if _in_ws_cell(x_p, L1_p, L2_p):
                invlist[(i1, i2)] = len(cells)
                cells.append((i1, i2))
    lst = np.array(cells, dtype=np.int64)
    N = len(cells)

    nnlist = np.zeros((N, 3, 3), dtype=np.int64)
    for nr in range(N):
        for nd1 in (-1, 0, 1):
            for nd2 in (-1, 0, 1):
                d_p = nd1 * a1_p + nd2 * a2_p
                x_p = lst[nr, 0] * a1_p + lst[nr, 1] * a2_p + d_p
                x1 = npbc(x_p, L1_p, L2_p)
                x_p = npbc(x1, L1_p, L2_p)
                x1 = npbc(x_p, L1_p, L2_p)
                x_p = npbc(x1, L1_p, L2_p)
                n1 = int(np.rint(float(BZ1_p @ x_p) / TWOPI))
                n2 = int(np.rint(float(BZ2_p @ x_p) / TWOPI))
                if (n1, n2) not in invlist:
                    raise ValueError("make_lattice: neighbour outside the cell list")
                nnlist[nr, nd1 + 1, nd2 + 1] = invlist[(n1, n2)]
    return Lattice(L1_p, L2_p, a1_p, a2_p, BZ1_p, BZ2_p, b1_p, b2_p, N, L, lst, invlist, nnlist)


def predefined_latt(lattice_type: str, L1: int, L2: int):
    """Predefined_Latt (Predefined_Latt_mod.F90:100-262) for the lattices the
    sweep engine is exercised on: SQUARE and HONEYCOMB.

    Returns (Ndim, List, Invlist, Latt, Latt_unit) with List[(site)] = (cell, orb)
    and Invlist[cell, orb] = site, all 0-based.
    """
    t = lattice_type.upper()
    if t == "SQUARE":
        if L2 == 1 and L1 > 1:
            n_coord = 1
        elif L2 > 1 and L1 > 1:
            n_coord = 2
        else:
            raise ValueError("For one-dimensional lattices set L2=1.")
        unit = UnitCell(1, n_coord, np.zeros((1, 2)))
        a1_p = np.array([1.0, 0.0])
        a2_p = np.array([0.0, 1.0])
    elif t == "HONEYCOMB":
        if L1 == 1 or L2 == 1:
            raise ValueError("The Honeycomb lattice cannot be one-dimensional.")
        a1_p = np.array([1.0, 0.0])
        a2_p = np.array([0.5, np.sqrt(3.0) / 2.0])
        pos = np.zeros((2, 2))
        pos[1, :] = (a2_p - 0.5 * a1_p) * 2.0 / 3.0
        unit = UnitCell(2, 3, pos)
    else:
        raise NotImplementedError(f"lattice {lattice_type!r} is not supported by this build")
    latt = make_lattice(L1 * a1_p, L2 * a2_p, a1_p, a2_p)
    ndim = latt.N * unit.Norb
    List = np.zeros((ndim, 2), dtype=np.int64)
    Invlist = np.zeros((latt.N, unit.Norb), dtype=np.int64)
    nc = 0
    for I in range(latt.N):
        for no in range(unit.Norb):
            List[nc] = (I, no)
            Invlist[I, no] = nc
            nc += 1
    return ndim, List, Invlist, latt, unit


def momentum_list(latt: Lattice) -> np.ndarray:
    """Listk (lattices_v3_mod.F90:243-275): integer coordinates (m, n) of the N momenta
    m b1 + n b2 in the Wigner-Seitz cell of the reciprocal lattice spanned by BZ1, BZ2
    (m outer, n inner)."""
    L = latt.L
    out = []
    for m in range(-L, L + 1):
        for n in range(-L, L + 1):
            xk = m * latt.b1_p + n * latt.b2_p
            ok = True
            for b_p in (latt.BZ2_p, latt.BZ1_p, latt.BZ2_p - latt.BZ1_p, latt.BZ2_p + latt.BZ1_p):
                s = float(xk @ b_p)
                h = float(b_p @ b_p) / 2.0
                if not (s <= h + EPS_SMALL and s >= -h + EPS_SMALL):
                    ok = False
                    break
            if ok:
                out.append((m, n))
    if len(out) != latt.N:
        raise ValueError(f"momentum_list: {len(out)} momenta for {latt.N} cells")
    return np.array(out, dtype=np.int64)


def distance_classes(latt: Lattice) -> np.ndarray:
    """imj (lattices_v3_mod.F90:316-330): imj[I, J] is the cell whose position equals
    R_I - R_J folded back into the cluster (0-based)."""
    N = latt.N
    imj = np.zeros((N, N), dtype=np.int64)
    for nr in range(N):
        x_p = latt.list[nr, 0] * latt.a1_p + latt.list[nr, 1] * latt.a2_p
        for nr1 in range(N):
            x1_p = latt.list[nr1, 0] * latt.a1_p + latt.list[nr1, 1] * latt.a2_p
            d_p = x_p - x1_p
            x1 = npbc(d_p, latt.L1_p, latt.L2_p)
            d_p = npbc(x1, latt.L1_p, latt.L2_p)
            i1 = int(np.rint(float(latt.BZ1_p @ d_p) / TWOPI))
            i2 = int(np.rint(float(latt.BZ2_p @ d_p) / TWOPI))
            imj[nr, nr1] = latt.invlist[(i1, i2)]
    return imj


def correlation_partner(imj: np.ndarray) -> np.ndarray:
    """jm[I, d] = the cell J with imj[I, J] = d (one per (I, d) on a periodic cluster)."""
    N = imj.shape[0]
    jm = np.full((N, N), -1, dtype=np.int64)
    for I in range(N):
        jm[I, imj[I]] = np.arange(N)
    if (jm < 0).any():
        raise ValueError("distance classes are not a bijection per cell")
    return jm
