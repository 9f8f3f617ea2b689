"""Checkerboard Op_T on the host and in the oracle (Predefined_Hop_mod.F90:1773-1834,
Hop_mod.F90:143-250), pinned by what the mathematics fixes: the bond operators add up to
the hopping matrix, each bond exponential is the matrix exponential of its 2 x 2 block,
the product over bonds is exp(-dtau T) up to O(dtau^2), the bonds of a family commute,
and a sweep with the decomposed propagator reproduces (1 + B...B)^-1 computed by brute
force from the same propagator.  No device needed."""
import numpy as np
import pytest
from scipy.linalg import expm

import oracle.dqmc_oracle as O
from alf_amd.host import hoppings
from alf_amd.host.hamiltonian import HubbardParams, hubbard_ham_set
from helpers import fields_for, hubbard, oracle, rel, tape

CASES = [("Square", 4, 0, 0.0), ("Square", 6, 0, 0.3), ("Honeycomb", 3, 1, 0.0), ("Honeycomb", 4, 2, 0.2)]


def embed(n, cb, b, M):
    out = np.zeros((n, n), dtype=complex)
    r = [int(cb.p1[b]), int(cb.p2[b])]
    out[np.ix_(r, r)] = M
    return out


@pytest.mark.parametrize("lattice,L,nphi,chem", CASES)
def test_bonds_partition_the_hopping_and_exponentiate(lattice, L, nphi, chem):
    ham = hubbard(L1=L, L2=L, lattice=lattice, n_phi=nphi, chem=chem, checkerboard=True)
    cb = ham.extra["checkerboard"]
    n = ham.Ndim
    nb = len(cb.p1)
    assert nb == ham.latt.N * (2 if lattice == "Square" else 3)
    assert sum(cb.fam_len) == nb and len(cb.fam_len) == (4 if lattice == "Square" else 3)
    # sum over bonds of the embedded 2 x 2 operators is the dense hopping matrix
    S = sum(embed(n, cb, b, cb.O[b]) for b in range(nb))
    assert np.abs(S - ham.hop_matrix).max() < 1e-14
    for b in range(nb):
        assert np.abs(cb.E[b] - expm(-ham.params.Dtau * cb.O[b])).max() < 1e-14
        assert np.abs(cb.E[b] @ cb.Einv[b] - np.eye(2)).max() < 1e-14
        assert np.abs(cb.E[b] - cb.E[b].conj().T).max() < 1e-15          # symmetrised
    # the bonds of one family sit on disjoint sites, so their product is order independent
    off = np.cumsum([0] + cb.fam_len)
    rng = np.random.default_rng(1)
    for f in range(len(cb.fam_len)):
        sites = np.concatenate([cb.p1[off[f]:off[f + 1]], cb.p2[off[f]:off[f + 1]]])
        assert len(np.unique(sites)) == len(sites)
        M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        a, b = M.copy(), M.copy()
        for q in range(off[f], off[f + 1]):
            a = embed(n, cb, q, cb.E[q]) @ a + (a - embed(n, cb, q, np.eye(2)) @ a)
        for q in rng.permutation(np.arange(off[f], off[f + 1])):
            b = embed(n, cb, q, cb.E[q]) @ b + (b - embed(n, cb, q, np.eye(2)) @ b)
        assert rel(a, b) < 1e-14


@pytest.mark.parametrize("lattice,L,nphi", [("Square", 4, 0), ("Honeycomb", 3, 1)])
def test_product_is_exp_minus_dtau_t_to_second_order(lattice, L, nphi):
    errs = []
    for dtau in (0.1, 0.05, 0.025):
        ham = hubbard(L1=L, L2=L, lattice=lattice, n_phi=nphi, dtau=dtau, checkerboard=True)
        errs.append(np.abs(ham.Op_T.mat - expm(-dtau * ham.hop_matrix)).max())
        assert np.abs(ham.Op_T.mat @ ham.Op_T.invmat - np.eye(ham.Ndim)).max() < 1e-13
    for a, b in zip(errs, errs[1:]):
        assert 3.6 < a / b < 4.4                                          # O(dtau^2) splitting error


def test_odd_square_lattice_is_refused():
    """With odd L the even/odd families share sites; the reference terminates
    (test_checkerboard_decomposition, Predefined_Hop_mod.F90:1687)."""
    with pytest.raises(ValueError, match="non-commuting"):
        hubbard(L1=5, L2=4, checkerboard=True)


@pytest.mark.parametrize("lattice,L,nphi", [("Square", 4, 0), ("Honeycomb", 3, 1)])
def test_oracle_hop_products_against_dense_products(lattice, L, nphi):
    """mmthr = P M, mmthr_m1 = P^-1 M, mmthl = M P, mmthl_m1 = M P^-1, mmthlc = P^+ M with
    P = E_1 E_2 ... E_N built here from scipy's expm of every bond."""
    ham = hubbard(L1=L, L2=L, lattice=lattice, n_phi=nphi, checkerboard=True)
    cb = ham.extra["checkerboard"]
    n, dt = ham.Ndim, ham.params.Dtau
    P = np.eye(n, dtype=complex)
    for b in range(len(cb.p1)):
        P = P @ (np.eye(n) + embed(n, cb, b, expm(-dt * cb.O[b]) - np.eye(2)))
    Pi = np.linalg.inv(P)
    m = O.model_from_ham(ham)
    M = np.random.default_rng(2).standard_normal((n, n)) + 1j * np.random.default_rng(3).standard_normal((n, n))
    assert rel(m.mmthr(M), P @ M) < 1e-13
    assert rel(m.mmthr_m1(M), Pi @ M) < 1e-13
    assert rel(m.mmthl(M), M @ P) < 1e-13
    assert rel(m.mmthl_m1(M), M @ Pi) < 1e-13
    assert rel(m.mmthlc(M), P.conj().T @ M) < 1e-13
    if lattice == "Honeycomb":
        assert rel(P.conj().T, P) > 1e-6                                 # P is not Hermitian here


@pytest.mark.parametrize("lattice,L,nphi", [("Square", 4, 0), ("Honeycomb", 3, 1)])
def test_checkerboard_chains_and_cgr_against_brute_force(lattice, L, nphi):
    """Right chain (mmthr) and left chain (mmthlc) of a checkerboard model meet in CGR at the
    brute-force (1 + B(tau, 0) B(beta, tau))^-1 of B = D P; the left chain with P in place
    of P^+ would not."""
    ham = hubbard(L1=L, L2=L, lattice=lattice, n_phi=nphi, beta=0.8, nwrap=4, checkerboard=True)
    f = fields_for(ham, 6)
    s = oracle(ham, f, tape(10))
    m = s.m
    ur = s.wrapur(0, 4, O.UDV.identity(m.ndim, "R"))
    ul = s.wrapul(8, 4, O.UDV.identity(m.ndim, "L"))
    z1, G1 = O.cgr(1, ur, ul)
    Gb, ph = O.brute_force_green(m, f, 4, dps=30)
    assert rel(G1, Gb) < 1e-11 and abs(z1 - ph) < 1e-11


def test_checkerboard_sweep_wraps_track_the_fresh_green():
    """A full sweep: WRAPGRUP/WRAPGRDO with the bond products agree with the Green's
    functions recomputed from the UDV chains at every stabilisation, and the final G
    matches the brute force."""
    ham = hubbard(L1=4, L2=4, beta=1.0, nwrap=5, U=4.0, checkerboard=True)
    f = fields_for(ham, 8)
    s = oracle(ham, f, tape(20000))
    s.init()
    s.sweep()
    assert s.ctrl.XMAXG < 1e-9 and s.ctrl.XMAXP < 1e-9
    Gb, _ = O.brute_force_green(s.m, s.f, 0, dps=30)
    assert rel(s.G, Gb) < 1e-10


def test_symmetrised_families_follow_the_reference_rule():
    """Symmetrize_Families (Predefined_Hop_mod.F90:1422-1495): families 1..N-1, N, N-1..1 with
    time steps 1/2 and 1 in the middle; the longest family goes to the middle."""
    from alf_amd.host.hoppings import symmetrize_families
    # equal lengths (the predefined square and honeycomb families): the first family is the
    # longest, so it takes the middle and the last family both ends
    out, prop = symmetrize_families([["a"], ["b"], ["c"], ["d"]])
    assert out == [["d"], ["b"], ["c"], ["a"], ["c"], ["b"], ["d"]]
    assert prop == [0.5, 0.5, 0.5, 1.0, 0.5, 0.5, 0.5]
    out, prop = symmetrize_families([["a"], ["b"], ["c"]])
    assert out == [["c"], ["b"], ["a"], ["b"], ["c"]]
    # the longest family last: the plain palindrome
    out, prop = symmetrize_families([["a"], ["b"], ["c", "c"]])
    assert out == [["a"], ["b"], ["c", "c"], ["b"], ["a"]] and prop == [0.5, 0.5, 1.0, 0.5, 0.5]
    # unequal lengths with the longest elsewhere: the reference's placement overwrites the first
    # and the last entries, so the longest family appears three times and the first one not at all
    out, _ = symmetrize_families([["a"], ["b", "b"], ["c"]])
    assert out == [["c"], ["b", "b"], ["b", "b"], ["b", "b"], ["c"]]


@pytest.mark.parametrize("lattice,L,nphi", [("Square", 6, 0), ("Honeycomb", 3, 0), ("Honeycomb", 4, 2)])
def test_symmetric_splitting_is_second_order(lattice, L, nphi):
    """With Symm the checkerboard product P = E_1 ... E_N of the symmetrised families and the
    half-step product Q of Hop_mod_Symm approach exp(-dtau T) and exp(-dtau T/2) with an error
    of order dtau^3 per step (falls about 8x per halving of dtau), against first order without
    Symm; Q Q^-1 = 1 and Q^dagger = Q for a palindromic product of Hermitian bond steps."""
    import scipy.linalg as sla
    errs, errq, err1 = [], [], []
    for dtau in (0.2, 0.1, 0.05):
        p = HubbardParams(Lattice_type=lattice, L1=L, L2=L, N_Phi=nphi, Beta=1.0, Dtau=dtau, Checkerboard=True,
                          Symm=True)
        ham = hubbard_ham_set(p)
        T = ham.hop_matrix
        errs.append(np.abs(ham.Op_T.mat - sla.expm(-dtau * T)).max())
        errq.append(np.abs(ham.Op_T.mat_1D2 - sla.expm(-dtau / 2 * T)).max())
        Q, Qi = ham.Op_T.mat_1D2, ham.Op_T.invmat_1D2
        assert np.abs(Q @ Qi - np.eye(ham.Ndim)).max() < 1e-13
        assert np.abs(Q - Q.conj().T).max() < 1e-13
        p1 = HubbardParams(Lattice_type=lattice, L1=L, L2=L, N_Phi=nphi, Beta=1.0, Dtau=dtau, Checkerboard=True)
        h1 = hubbard_ham_set(p1)
        err1.append(np.abs(h1.Op_T.mat - sla.expm(-dtau * T)).max())
    for e in (errs, errq):
        assert 6.0 < e[0] / e[1] < 10.0 and 6.0 < e[1] / e[2] < 10.0, e
    assert 3.0 < err1[0] / err1[1] < 5.0, err1


def test_dense_symm_half_steps():
    """Dense Op_T with Symm: mat_1D2 = exp(-dtau T/2), invmat_1D2 its inverse (the oracle's own
    expm of the geometric hopping matrix), and the oracle's Hop_mod_Symm is their similarity."""
    from oracle import model_setup as M
    from oracle.dqmc_oracle import model_from_ham
    ham = hubbard_ham_set(HubbardParams(L1=4, L2=4, Beta=1.0, Dtau=0.1, Symm=True))
    T = M.hopping_matrix("Square", 4, 4)
    h, hi = M.exp_hopping(T, 0.05)
    assert np.abs(ham.Op_T.mat_1D2 - h).max() < 1e-13 and np.abs(ham.Op_T.invmat_1D2 - hi).max() < 1e-13
    m = model_from_ham(ham)
    G = np.random.default_rng(0).standard_normal((16, 16)) + 0j
    assert np.abs(m.hop_symm(G) - h @ G @ hi).max() < 1e-13
