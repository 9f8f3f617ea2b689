"""The product's model set-up (alf_amd/host: site order, hopping matrix, exp(-dtau T),
checkerboard bonds) against the oracle's own construction from geometry
(oracle/model_setup.py), so that a mistake in the host code that both sides of the sweep
parity tests share cannot pass unseen."""
import numpy as np
import pytest

from alf_amd.host.hamiltonian import HubbardParams, hubbard_ham_set
from oracle import model_setup as M


CASES = [("Square", 4, 4, 0.0), ("Square", 6, 6, -0.2), ("Square", 2, 4, 0.3), ("Square", 3, 5, 0.0),
         ("Square", 16, 16, 0.0), ("Honeycomb", 2, 2, 0.0), ("Honeycomb", 3, 3, 0.1), ("Honeycomb", 4, 6, 0.0),
         ("Honeycomb", 12, 12, 0.0)]


@pytest.mark.parametrize("lattice,L1,L2,mu", CASES)
def test_site_order_and_hopping_exponential(lattice, L1, L2, mu):
    ham = hubbard_ham_set(HubbardParams(Lattice_type=lattice, L1=L1, L2=L2, Ham_chem=mu, Beta=1.0, Dtau=0.1))
    # site order: the product's cell list and orbital order give the oracle's positions
    a1, a2, _, _, orbs = M.lattice_vectors(lattice, L1, L2)
    cells = np.asarray(ham.latt.list)
    pos = (np.outer(cells[:, 0], a1) + np.outer(cells[:, 1], a2))[:, None, :] + orbs[None]
    assert np.abs(pos.reshape(-1, 2) - M.site_positions(lattice, L1, L2)).max() < 1e-12
    # hopping matrix and its exponentials entry by entry
    T = M.hopping_matrix(lattice, L1, L2, 1.0, mu)
    assert np.abs(ham.hop_matrix - T).max() < 1e-14
    e, ei = M.exp_hopping(T, 0.1)
    assert np.abs(ham.Op_T.mat - e).max() < 1e-13
    assert np.abs(ham.Op_T.invmat - ei).max() < 1e-13
    assert ham.Op_T.is_real


@pytest.mark.parametrize("L,nphi", [(12, 18), (6, 6), (4, 2), (6, 12)])
def test_flux_through_every_hexagon(L, nphi):
    """Honeycomb with N_Phi flux quanta: the bonds are the oracle's nearest-neighbour pairs with
    |T_ij| = t, and every hexagon of the torus, those across the boundary included, encloses
    the same flux 2 pi N_Phi / (L1 L2); exp(-dtau T) is scipy's expm of the product's T.
    Even N_Phi only (the 12 x 12, N_Phi = 18 workload among them): with odd N_Phi the Landau
    gauge's boundary terms of Generic_hopping leave a flux offset of pi on the hexagons along
    the supercell boundary in this restatement, and the reference module that would settle it
    (Predefined_Hop_mod, which needs the generated Hamiltonian_main) does not build here, so
    odd N_Phi is parity unpinned."""
    ham = hubbard_ham_set(HubbardParams(Lattice_type="Honeycomb", L1=L, L2=L, N_Phi=nphi, Beta=1.0, Dtau=0.1))
    T = ham.hop_matrix
    T0 = M.hopping_matrix("Honeycomb", L, L)
    assert np.abs(np.abs(T) - np.abs(T0)).max() < 1e-14
    hexes = M.hexagon_sites(L, L)
    assert len(hexes) == L * L
    fl = M.face_fluxes(T, hexes)
    want = 2 * np.pi * nphi / (L * L)
    wrapped = np.angle(np.exp(1j * (fl - fl[0])))
    assert np.abs(wrapped).max() < 1e-10                      # a uniform field
    assert min(abs(np.angle(np.exp(1j * (fl[0] - s * want)))) for s in (1, -1)) < 1e-10
    e, ei = M.exp_hopping(T, 0.1)
    assert np.abs(ham.Op_T.mat - e).max() < 1e-13
    assert np.abs(ham.Op_T.invmat - ei).max() < 1e-13


@pytest.mark.parametrize("lattice,L,nphi", [("Square", 4, 0), ("Square", 6, 0), ("Honeycomb", 3, 0), ("Honeycomb", 4, 3)])
def test_checkerboard_bonds_cover_the_hopping(lattice, L, nphi):
    """The checkerboard bond operators sit on the oracle's nearest-neighbour pairs, each pair
    once, and their product is exp(-dtau T) up to the splitting error."""
    ham = hubbard_ham_set(HubbardParams(Lattice_type=lattice, L1=L, L2=L, N_Phi=nphi, Beta=1.0, Dtau=0.1,
                                        Checkerboard=True))
    cb = ham.extra["checkerboard"]
    bond = M.hopping_matrix(lattice, L, L) != 0
    seen = np.zeros_like(bond)
    for p1, p2 in zip(cb.p1, cb.p2):
        assert bond[p1, p2]
        assert not seen[p1, p2]
        seen[p1, p2] = seen[p2, p1] = True
    assert np.array_equal(seen, bond)
    e, _ = M.exp_hopping(ham.hop_matrix, 0.1)
    assert np.abs(ham.Op_T.mat - e).max() < 0.02               # first-order splitting at dtau = 0.1
