"""The oracle's own model set-up: site order, hopping matrix and exp(-dtau T).

TEST INFRASTRUCTURE ONLY (the oracle): nothing in the product path imports this
file.  It builds the Hubbard plugin's one-body matrices from geometry alone,
independently of the product's host code (alf_amd/host), so that the parity
tests of the sweep (which feed the product's model to both sides) rest on a
model that is itself checked:

* ``ws_cells``: the cells of the Wigner-Seitz supercell in the reference's order
  (Libraries/Modules/lattices_v3_mod.F90:184-237: i1 outer, i2 inner over
  [-L, L], a cell is kept when |x.a| <= |a|^2/2 + Eps_small for a in L2, L1,
  L2 - L1, L2 + L1, with L from the reciprocal vectors, 173-181); orbitals are
  the inner index of the site list (Prog/Predefined_Latt_mod.F90:251-260).
* ``hopping_matrix``: -t on every nearest-neighbour pair, found from the site
  positions and all periodic images of the supercell, -mu on the diagonal.  A pair
  joined through two images (a length-2 direction) carries -t once, not -2t: the
  reference assigns each bond's entry rather than adding it
  (Prog/Predefined_Hop_mod.F90:1743-1744).  No Peierls phases:
  the flux case is checked through gauge-invariant face fluxes instead
  (``face_fluxes``), since its entries depend on the reference's gauge.
* ``exp_hopping``: scipy's expm of -dtau T and dtau T (a Pade method, not the
  eigen-decomposition with Kahan sums the product uses, Operator_mod.F90:510-527).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

EPS_SMALL = 1.0e-10


def lattice_vectors(lattice_type: str, L1: int, L2: int):
    """(a1, a2, L1_p, L2_p, orbital offsets) of Predefined_Latt (Predefined_Latt_mod.F90:100-206)."""
    t = lattice_type.upper()
    if t == "SQUARE":
        a1, a2 = np.array([1.0, 0.0]), np.array([0.0, 1.0])
        orbs = np.zeros((1, 2))
    elif t == "HONEYCOMB":
        a1, a2 = np.array([1.0, 0.0]), np.array([0.5, np.sqrt(3.0) / 2.0])
        orbs = np.array([[0.0, 0.0], (a2 - 0.5 * a1) * 2.0 / 3.0])
    else:
        raise NotImplementedError(lattice_type)
    return a1, a2, L1 * a1, L2 * a2, orbs


def ws_cells(a1, a2, L1_p, L2_p):
    """Integer coordinates (i1, i2) of the supercell's cells in the reference's order."""
    A = np.array([a1, a2])
    BZ = 2.0 * np.pi * np.linalg.inv(A)          # columns: BZ1_p, BZ2_p (a_i . BZ_j = 2 pi delta_ij)
    L = max(abs(int(np.rint(v))) for v in (BZ[:, 0] @ L1_p / (2 * np.pi), BZ[:, 1] @ L1_p / (2 * np.pi),
                                            BZ[:, 0] @ L2_p / (2 * np.pi), BZ[:, 1] @ L2_p / (2 * np.pi)))
    i1, i2 = np.meshgrid(np.arange(-L, L + 1), np.arange(-L, L + 1), indexing="ij")
    i1, i2 = i1.ravel(), i2.ravel()                # i1 outer, i2 inner
    x = np.outer(i1, a1) + np.outer(i2, a2)
    keep = np.ones(len(i1), dtype=bool)
    for a in (L2_p, L1_p, L2_p - L1_p, L2_p + L1_p):
        s = x @ a
        h = a @ a / 2.0
        keep &= (s <= h + EPS_SMALL) & (s >= -h + EPS_SMALL)
    return np.stack([i1[keep], i2[keep]], axis=1)


def site_positions(lattice_type: str, L1: int, L2: int) -> np.ndarray:
    """(Ndim, 2) positions of the sites in the reference's order (cell outer, orbital inner)."""
    a1, a2, L1_p, L2_p, orbs = lattice_vectors(lattice_type, L1, L2)
    cells = ws_cells(a1, a2, L1_p, L2_p)
    base = np.outer(cells[:, 0], a1) + np.outer(cells[:, 1], a2)
    return (base[:, None, :] + orbs[None, :, :]).reshape(-1, 2)


def _images(L1_p, L2_p, reach=2):
    return [m1 * L1_p + m2 * L2_p for m1 in range(-reach, reach + 1) for m2 in range(-reach, reach + 1)]


def nn_distance(lattice_type: str) -> float:
    return 1.0 if lattice_type.upper() == "SQUARE" else 1.0 / np.sqrt(3.0)


def hopping_matrix(lattice_type: str, L1: int, L2: int, t: float = 1.0, mu: float = 0.0) -> np.ndarray:
    """T with H_T = c^dagger T c: -t on each nearest-neighbour pair (through any periodic image),
    -mu on the diagonal."""
    _, _, L1_p, L2_p, _ = lattice_vectors(lattice_type, L1, L2)
    x = site_positions(lattice_type, L1, L2)
    n = len(x)
    d0 = nn_distance(lattice_type)
    bond = np.zeros((n, n), dtype=bool)
    for R in _images(L1_p, L2_p):
        d = np.linalg.norm(x[None, :, :] + R[None, None, :] - x[:, None, :], axis=2)
        bond |= np.abs(d - d0) < 1e-8
    return -t * bond.astype(float) - mu * np.eye(n)


def exp_hopping(T: np.ndarray, dtau: float):
    """(exp(-dtau T), exp(dtau T)) by scipy's expm."""
    return sla.expm(-dtau * T), sla.expm(dtau * T)


def hexagon_sites(L1: int, L2: int) -> list:
    """For every cell of the honeycomb supercell, the six sites of the hexagon that has the
    cell's A site as its lowest vertex, counter-clockwise: the edges (length 1/sqrt 3) point
    at -30, 30, 90, 150, 210 and 270 degrees."""
    _, _, L1_p, L2_p, _ = lattice_vectors("Honeycomb", L1, L2)
    x = site_positions("Honeycomb", L1, L2)
    images = _images(L1_p, L2_p)

    def find(p):
        for R in images:
            d = np.linalg.norm(x - (p + R)[None, :], axis=1)
            k = int(np.argmin(d))
            if d[k] < 1e-8:
                return k
        raise ValueError("no site at this position")

    r = nn_distance("Honeycomb")
    edges = [r * np.array([np.cos(np.radians(a)), np.sin(np.radians(a))]) for a in (-30, 30, 90, 150, 210, 270)]
    hexes = []
    for k in range(0, len(x), 2):
        p = x[k].copy()
        ring = []
        for e in edges:
            ring.append(find(p))
            p = p + e
        assert np.linalg.norm(p - x[k]) < 1e-9
        hexes.append(ring)
    return hexes


def face_fluxes(T: np.ndarray, hexes: list) -> np.ndarray:
    """Phase of the product of hoppings -T around each hexagon (counter-clockwise)."""
    out = []
    for ring in hexes:
        z = 1.0 + 0j
        for i, j in zip(ring, ring[1:] + ring[:1]):
            z *= -T[j, i]
        out.append(np.angle(z))
    return np.array(out)
