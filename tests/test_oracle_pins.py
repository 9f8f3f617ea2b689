"""Pins of the CPU oracle against what the mathematics fixes.

Each check is chosen so a plausible mistake in the oracle (a dropped term, a
wrong sign or index, a transposed operand) fails it: brute-force Green's
functions in mpmath, Sherman-Morrison against recomputation with one field
changed, determinant ratios, QR invariants, closed-form spectra and the
moments of the four-state HS transformation.
"""
import math

import numpy as np
import pytest

from helpers import fields_for, hubbard, oracle, rel, tape
from oracle import dqmc_oracle as O


# ---------------------------------------------------------------- field tables
def test_hs_moments_match_gaussian():
    """sum_s gama(s) phi(s)^(2k) / 4 = E[eta^(2k)] for eta ~ N(0, 2), k = 0..3
    (the defining property of the four-state discrete HS transformation), and
    odd moments vanish."""
    for k, want in zip(range(4), (1.0, 2.0, 12.0, 120.0)):
        got = sum(O.gama(s, 2) * O.phi(s, 2) ** (2 * k) for s in (-2, -1, 1, 2)) / 4.0
        assert got == pytest.approx(want, rel=1e-13)
    for k in range(4):
        assert abs(sum(O.gama(s, 2) * O.phi(s, 2) ** (2 * k + 1) for s in (-2, -1, 1, 2))) < 1e-12


def test_flip_table_and_nranf():
    for s, alts in O.FLIP2.items():
        assert sorted(alts + (s,)) == [-2, -1, 1, 2]
    # nranf(3) partitions [0,1) into thirds (Fortran nint rounds halves away from zero)
    assert [O.nranf(u, 3) for u in (0.0, 0.3333, 1 / 3, 0.5, 0.6666, 2 / 3, 0.99999)] == [1, 1, 2, 2, 2, 3, 3]


# ---------------------------------------------------------------- hopping exponentials
def test_square_hopping_spectrum():
    """exp(-dtau T) on the periodic L x L square lattice has eigenvalues
    exp(2 dtau t (cos kx + cos ky))."""
    ham = hubbard(L1=6, L2=6, beta=1.0)
    ev = np.sort(np.linalg.eigvalsh(np.real(ham.Op_T.mat)))
    ks = 2 * np.pi * np.arange(6) / 6
    want = np.sort([math.exp(2 * 0.1 * (math.cos(a) + math.cos(b))) for a in ks for b in ks])
    assert np.abs(ev - want).max() < 1e-13
    assert ham.Op_T.is_real
    assert np.abs(ham.Op_T.mat @ ham.Op_T.invmat - np.eye(36)).max() < 1e-13


def test_honeycomb_flux_hopping():
    """Honeycomb with N_Phi flux quanta: T is Hermitian, each hexagon encloses the
    flux 2 pi N_Phi / (L1 L2) (gauge-invariant product of hoppings), and the
    spectrum of exp(-dtau T) is exp(-dtau eps) of T's."""
    L, nphi = 6, 9
    ham = hubbard(L1=L, L2=L, beta=1.0, lattice="Honeycomb", n_phi=nphi)
    T = ham.hop_matrix
    assert np.abs(T - T.conj().T).max() < 1e-14
    assert not ham.Op_T.is_real
    want = 2 * np.pi * nphi / (L * L)
    # find all 6-cycles through site 0 using the bond graph and check the enclosed phases
    n = T.shape[0]
    adj = {i: [j for j in range(n) if j != i and abs(T[i, j]) > 1e-12] for i in range(n)}
    phases = set()
    start = 0
    def walk(path):
        if len(path) == 6:
            if start in adj[path[-1]]:
                z = 1.0 + 0j
                cyc = path + [start]
                for i, j in zip(cyc[:-1], cyc[1:]):
                    z *= -T[i, j]
                phases.add(float(np.angle(z)))
            return
        for j in adj[path[-1]]:
            if j not in path:
                walk(path + [j])
    walk([start])
    # both orientations of every hexagon through site 0 carry +-want (mod 2 pi)
    assert phases
    for p in phases:
        d = min(abs(((p - want + np.pi) % (2 * np.pi)) - np.pi), abs(((p + want + np.pi) % (2 * np.pi)) - np.pi))
        assert d < 1e-10
    ev = np.sort(np.linalg.eigvalsh(T))
    evx = np.sort(np.linalg.eigvalsh(ham.Op_T.mat))
    assert np.abs(np.sort(np.exp(-0.1 * ev)) - evx).max() < 1e-12


def test_op_set_hubbard_coupling():
    ham = hubbard(L1=4, L2=4, beta=1.0, U=4.0)
    op = ham.Op_V[3]
    assert op.type == 2 and op.N == 1 and op.diag
    assert op.g == pytest.approx(1j * math.sqrt(0.1 * 4.0 / 2.0))
    assert op.alpha == -0.5
    # E_exp(-s) = 1 / E_exp(s) and |E_exp| = 1 for imaginary g
    for s in (1, 2):
        assert abs(op.E_exp[s][0] * op.E_exp[-s][0] - 1) < 1e-15
        assert abs(abs(op.E_exp[s][0]) - 1) < 1e-15


# ---------------------------------------------------------------- Green's functions
@pytest.mark.parametrize("presort", [False, True])
def test_init_green_matches_brute_force(presort):
    ham = hubbard(L1=2, L2=4, beta=1.0, nwrap=3)
    f = fields_for(ham, 3)
    s = oracle(ham, f, tape(10), presort=presort)
    s.init()
    Gb, ph = O.brute_force_green(s.m, s.f, 0, dps=40)
    assert rel(s.G, Gb) < 1e-11
    z = ph
    for k in range(s.m.n_op):
        for nt in range(s.m.ltrot):
            z *= np.exp(1j * (s.m.op_g[k] * s.m.op_alpha[k] * O.phi(s.f[k, nt], 2)).imag)
    assert abs(z ** 2 - s.phase) < 1e-11


def test_cgr_nvar1_equals_nvar2_and_brute_force():
    ham = hubbard(L1=2, L2=4, beta=2.0, nwrap=5)
    f = fields_for(ham, 5)
    s = oracle(ham, f, tape(10), presort=False)
    m = s.m
    # UDV states for tau = 10: right chain over slices 1..10, left over 20..11
    ur = s.wrapur(0, 5, O.UDV.identity(m.ndim, "R"))
    ur = s.wrapur(5, 10, ur)
    ul = s.wrapul(20, 15, O.UDV.identity(m.ndim, "L"))
    ul = s.wrapul(15, 10, ul)
    z1, G1 = O.cgr(1, ur, ul)
    z2, G2 = O.cgr(2, ur, ul)
    Gb, ph = O.brute_force_green(m, f, 10)
    assert rel(G1, Gb) < 1e-11 and rel(G2, Gb) < 1e-11
    assert abs(z1 - ph) < 1e-11 and abs(z2 - ph) < 1e-11


def test_decompose_invariants():
    rng = np.random.default_rng(7)
    n = 12
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    scales = np.exp(np.linspace(-20, 20, n))
    A = A * scales[None, :]
    for presort in (False, True):
        for side in ("R", "L"):
            u = O.UDV(A.copy(), np.ones(n, complex), np.eye(n, dtype=complex), side)
            out = O.decompose(u, presort)
            assert np.abs(out.U.conj().T @ out.U - np.eye(n)).max() < 1e-13
            if side == "R":
                recon = (out.U * out.D) @ out.V
            else:
                recon = (out.U * out.D) @ out.V.conj().T
            assert rel(recon, A) < 1e-13
            assert abs(np.linalg.det(out.V) - 1) < 1e-8           # det V = 1 normalisation
            assert abs(out.detU - np.linalg.det(out.U)) < 1e-12   # reflector-product determinant


# ---------------------------------------------------------------- Metropolis / Sherman-Morrison
def test_upgrade2_rank1_and_ratio():
    """An accepted Upgrade2 leaves G equal to the Green's function recomputed from
    scratch with the new field, and the Metropolis ratio equals the ratio of
    fermion determinants (times the gamma and alpha factors)."""
    ham = hubbard(L1=2, L2=2, beta=0.5, nwrap=5)
    f = fields_for(ham, 11)
    m = O.model_from_ham(ham)
    G0, _ = O.brute_force_green(m, f, 1, dps=40)
    s = O.OracleSweep(m, f, O.Tape([0.0]))   # u = 0: always accept
    n, nt = 2, 0
    new = O.FLIP2[int(f[n, nt])][1]
    P = int(m.op_P[n])
    # ratio from the determinants of M = 1 + B...B  (det M = 1/det G)

    def detM(ff):
        Gm, _ = O.brute_force_green(m, ff, 1, dps=40)
        return 1.0 / np.linalg.det(Gm)
    f2 = f.copy()
    f2[n, nt] = new
    det_ratio = detM(f2) / detM(f)
    G = G0.copy()
    old = int(f[n, nt])
    Z1 = m.op_g[n] * (O.phi(new, 2) - O.phi(old, 2))
    ratio = (1 + (np.exp(Z1) - 1) * (1 - G[P, P]))
    assert abs(ratio - det_ratio) < 1e-10 * abs(det_ratio)
    s.phase = 1.0 + 0j
    assert s.upgrade2(G, n, nt, new)
    G1, _ = O.brute_force_green(m, f2, 1, dps=40)
    assert rel(G, G1) < 1e-11


def test_wrapgrup_matches_brute_force_slice_by_slice():
    ham = hubbard(L1=2, L2=4, beta=1.0, nwrap=5)
    f = fields_for(ham, 2)
    s = oracle(ham, f, tape(5000), presort=False)
    s.init()
    for ntau in range(3):
        s.G = s.wrapgrup(s.G, ntau)
        Gb, _ = O.brute_force_green(s.m, s.f, ntau + 1, dps=40)
        assert rel(s.G, Gb) < 1e-10


@pytest.mark.slow
def test_full_sweep_against_brute_force_and_variants():
    """A complete sweep: the wrapped and the fresh G agree at every stabilisation
    (Precision Green), the final G equals the brute force of the final fields,
    and the default (ZGEQP3) and presorted pivoting give the same trajectory."""
    ham = hubbard(L1=4, L2=4, beta=2.0, nwrap=10)
    f = fields_for(ham, 4)
    u = tape(20000)
    a = oracle(ham, f, u, presort=False)
    b = oracle(ham, f, u, presort=True)
    for s in (a, b):
        s.init()
        s.sweep()
        assert s.ctrl.XMAXG < 1e-9
        assert s.ctrl.XMAXP < 1e-9
    assert np.array_equal(a.f, b.f)
    assert a.ctrl.ACC_up == b.ctrl.ACC_up
    assert rel(a.G, b.G) < 1e-11
    Gb, _ = O.brute_force_green(a.m, a.f, 0, dps=30)
    assert rel(a.G, Gb) < 1e-10


# ---------------------------------------------------------------- equal-time observables
def _obs_oracle(ham):
    from alf_amd.host import lattices as LT
    from oracle.observables import HubbardObs
    imj = LT.distance_classes(ham.latt)
    return HubbardObs(ham.hop_matrix, ham.params.Ham_U, ham.N_SUN, imj, ham.List[:, 0], ham.List[:, 1],
                      ham.unit.Norb), imj


def _positions(latt, cells):
    return np.array([latt.list[c, 0] * latt.a1_p + latt.list[c, 1] * latt.a2_p for c in cells])


@pytest.mark.parametrize("lattice,L", [("Square", 4), ("Square", 6), ("Honeycomb", 3)])
def test_distance_classes_and_momenta(lattice, L):
    """imj[I, J] is R_I - R_J modulo the cluster periods; the momenta are commensurate with
    the cluster, distinct modulo the reciprocal lattice, and orthogonal on the cells."""
    from alf_amd.host import lattices as LT
    ham = hubbard(L1=L, L2=L, lattice=lattice)
    latt = ham.latt
    imj = LT.distance_classes(latt)
    R = _positions(latt, range(latt.N))
    P = np.array([latt.L1_p, latt.L2_p]).T
    for I in range(latt.N):
        for J in range(latt.N):
            c = np.linalg.solve(P, R[I] - R[J] - R[imj[I, J]])
            assert np.allclose(c, np.rint(c), atol=1e-9)
    k = LT.momentum_list(latt)
    K = np.array([m * latt.b1_p + n * latt.b2_p for m, n in k])
    assert np.allclose(np.exp(1j * K @ latt.L1_p), 1) and np.allclose(np.exp(1j * K @ latt.L2_p), 1)
    A = np.exp(1j * K @ R.T)
    S = A @ A.conj().T / latt.N                                           # sum_r e^{i (k - k') r} / N
    assert np.allclose(S, np.eye(latt.N), atol=1e-9)


def _free_green(ham, beta):
    e, V = np.linalg.eigh(ham.hop_matrix)
    f = 1.0 / (1.0 + np.exp(beta * e))
    G = V @ np.diag(1.0 - f) @ V.conj().T           # <c_i c^+_j> of H = c^+ O c
    return G, e, f


@pytest.mark.parametrize("lattice,L,chem", [("Square", 4, 0.3), ("Square", 6, -0.7), ("Honeycomb", 3, 0.4)])
def test_observables_of_free_fermions(lattice, L, chem):
    """U -> 0: Kin, Part, the Green correlation at every k and the k = 0 spin and density
    correlations have closed forms in the eigenvalues of the hopping matrix."""
    from alf_amd.host import lattices as LT
    from oracle.observables import fourier_r_to_k
    ham = hubbard(L1=L, L2=L, lattice=lattice, chem=chem, beta=2.0)
    obs, imj = _obs_oracle(ham)
    G, e, f = _free_green(ham, 2.0)
    obs.measure(G, 1.0 + 0j)
    obs.measure(G, 1.0 + 0j)
    b = obs.bins()
    ns = ham.N_SUN
    part = ns * f.sum()
    assert abs(b["scal"][0] - ns * np.sum(e * f)) < 1e-10
    assert abs(b["scal"][2] - part) < 1e-10
    assert abs(b["scal"][3] - b["scal"][0] - b["scal"][1]) < 1e-12 and b["sign_scal"] == 1.0
    latt = ham.latt
    R = _positions(latt, range(latt.N))
    k = LT.momentum_list(latt)
    K = np.array([m * latt.b1_p + n * latt.b2_p for m, n in k])
    X = fourier_r_to_k(b["latt"].reshape(3, latt.N, -1).transpose(1, 0, 2), R, K)   # (Nk, 3, norb^2)
    k0 = int(np.argmin(np.linalg.norm(K, axis=1)))
    ff = np.sum(f * (1 - f))
    # sum rules at k = 0 (all orbital pairs): spin N_SUN sum f(1-f) / N, density (Part^2 + that) / N
    assert abs(X[k0, 1].sum() - ns * ff / latt.N) < 1e-10
    assert abs(X[k0, 2].sum() - (part ** 2 + ns * ff) / latt.N) < 1e-9
    assert np.allclose(b["den0"].sum() * latt.N, part)
    if lattice == "Square":
        # one orbital: Green(k) = N_SUN f(eps(k)), eps(k) = sum_j O_0j exp(i k (R_j - R_0))
        O = ham.hop_matrix
        site0 = ham.Invlist[0, 0]
        for ik in range(latt.N):
            eps = sum(O[site0, ham.Invlist[j, 0]] * np.exp(1j * K[ik] @ (R[j] - R[0])) for j in range(latt.N))
            assert abs(X[ik, 0, 0] - ns / (1 + np.exp(2.0 * eps.real))) < 1e-10


def test_potential_energy_at_half_filling_and_phase_factors():
    """Pot = U n / 4 for free fermions at half filling on a bipartite lattice; a negative
    real phase flips every accumulated value and the sign; a complex phase multiplies them
    by PHASE / Re(PHASE)."""
    ham = hubbard(L1=4, L2=4, U=3.0, beta=3.0)
    obs, _ = _obs_oracle(ham)
    G, e, f = _free_green(ham, 3.0)
    obs.measure(G, 1.0 + 0j)
    ref = obs.bins()
    assert abs(ref["scal"][1] - 3.0 * ham.Ndim / 4) < 1e-10
    obs.reset()
    obs.measure(G, -2.0 + 0j)
    neg = obs.bins()
    assert neg["sign_scal"] == -1.0 and neg["sign_eq"] == -1.0
    assert np.allclose(neg["scal"], -ref["scal"]) and np.allclose(neg["latt"], -ref["latt"])
    obs.reset()
    obs.measure(G, 2.0 + 1.0j)
    cpx = obs.bins()
    assert np.allclose(cpx["scal"], ref["scal"] * (1 + 0.5j)) and cpx["sign_scal"] == 1.0


# ---------------------------------------------------------------- exact diagonalisation
def test_ed_free_fermions_atomic_limit_and_trotter_limit():
    from oracle.ed import HubbardED
    ham = hubbard(L1=2, L2=2, U=4.0, beta=2.0, chem=0.2)
    O = ham.hop_matrix
    # U = 0: Kin = 2 sum_e e f(e), green[i, j] = 2 f(O)_ji
    ed0 = HubbardED(O, 0.0)
    r = ed0.observables(ed0.thermal_rho(2.0))
    e, V = np.linalg.eigh(O)
    f = 1 / (1 + np.exp(2.0 * e))
    assert abs(r["Kin"] - 2 * np.sum(e * f)) < 1e-10
    assert np.allclose(r["Green"], 2 * (V @ np.diag(f) @ V.conj().T).T, atol=1e-10)
    # t = 0: each site alone, Z = 2 + 2 exp(-beta U / 2), <n_up n_dn> = exp(-beta U / 2) / Z
    edA = HubbardED(np.zeros((4, 4)), 4.0)
    a = edA.observables(edA.thermal_rho(2.0))
    x = np.exp(-2.0 * 2.0)
    assert abs(a["Pot"] - 4 * 4.0 * x / (2 + 2 * x)) < 1e-12 and abs(a["Part"] - 4) < 1e-12
    assert np.allclose(np.diag(a["SpinZ"]), 2 / (2 + 2 * x))
    # Trotter product -> exp(-beta H) as dtau -> 0, with the error shrinking with dtau
    ed = HubbardED(O, 4.0)
    ex = ed.observables(ed.thermal_rho(2.0))["Ener"]
    d1 = abs(ed.observables(ed.trotter_rho(0.1, 20))["Ener"] - ex)
    d2 = abs(ed.observables(ed.trotter_rho(0.01, 200))["Ener"] - ex)
    assert d2 < 0.2 * d1 and d2 < 1e-3


def test_ed_energy_matches_reference_known_answer():
    """The reference's own exact-diagonalisation anchor (testsuite/test_vs_ed/test_specs.yaml:25,
    ed_energy = -1.47261997 for the 4-site Hubbard ring at U = 4, beta = 2, t = 1): the 2 x 2
    cluster of this build has the ring's spectrum, and <H_T> + U sum <n_up n_dn> agrees."""
    from oracle.ed import HubbardED
    ham = hubbard(L1=2, L2=2, U=4.0, beta=2.0)
    assert np.allclose(np.sort(np.linalg.eigvalsh(ham.hop_matrix)), [-2, 0, 0, 2])
    ed = HubbardED(ham.hop_matrix, 4.0)
    assert abs(ed.observables(ed.thermal_rho(2.0))["Ener"].real - (-1.47261997)) < 5e-9
