"""Time-displaced Green's functions and measurements in the oracle (TAU_M, CGR2_2, ObserT),
pinned by what the mathematics fixes: the block inverse of [[1, B(beta,tau)], [-B(tau,0), 1]],
mpmath brute force of G(tau, 0), G(0, tau), G(tau, tau), the free-fermion closed forms
e^{-tau T} (1 + e^{-beta T})^-1, the boundary values at tau = 0 and tau = beta, and the
equal-time measurement at tau = 0.  No device needed."""
import numpy as np
import pytest
from scipy.linalg import expm

import oracle.dqmc_oracle as O
from helpers import fields_for, hubbard, oracle, rel, tape
from oracle.observables import HubbardObs, HubbardTauObs


def _unitary(rng, n):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return q


def _udv(rng, n, lo, hi, side):
    U = _unitary(rng, n)
    D = np.exp(np.linspace(hi, lo, n)).astype(complex)
    V = np.eye(n) + 0.3 * np.triu(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)), 1)
    return O.UDV(U, D, V, side)


@pytest.mark.parametrize("left_larger", [True, False])
def test_cgr2_2_inverts_the_block_matrix(left_larger):
    """O = [[1, B2], [-B1, 1]] with B1 = U D V (right), B2 = V D U^H (left): its inverse is
    [[G00, G0T], [GT0, GTT]]; both block orders of CGR2_2 (D1(1) > D2(1) or not)."""
    rng = np.random.default_rng(4)
    n = 10
    r = _udv(rng, n, -6.0, 4.0 if left_larger else 7.0, "R")
    l = _udv(rng, n, -5.0, 7.0 if left_larger else 4.0, "L")
    B1 = (r.U * r.D) @ r.V
    B2 = l.V @ (l.D[:, None] * l.U.conj().T)
    Om = np.block([[np.eye(n), B2], [-B1, np.eye(n)]])
    inv = np.linalg.inv(Om)
    GT0, G00, GTT, G0T = O.cgr2_2(r, l)
    assert (l.D[0].real > r.D[0].real) == left_larger
    for got, want in ((G00, inv[:n, :n]), (G0T, inv[:n, n:]), (GT0, inv[n:, :n]), (GTT, inv[n:, n:])):
        assert rel(got, want) < 1e-9
    # the identities the blocks obey
    assert rel(GT0, B1 @ G00) < 1e-9
    assert rel(GTT, np.linalg.inv(np.eye(n) + B1 @ B2)) < 1e-9


@pytest.mark.parametrize("cb", [False, True])
def test_tau_m_matches_brute_force(cb):
    """Propagated and CGR2_2-stabilised G(tau, 0), G(0, tau), G(tau, tau), G(0, 0) after a
    sweep equal the mpmath brute force at every tau (dense and checkerboard hopping)."""
    ham = hubbard(L1=4, L2=4, beta=1.0, nwrap=3, U=4.0, checkerboard=cb)
    f = fields_for(ham, 2)
    s = oracle(ham, f, tape(100000))
    s.init()
    s.sweep()
    seen = {}
    s.tau_m(hook=lambda nt, a, b, c, d: seen.__setitem__(nt, (a.copy(), b.copy(), c.copy(), d.copy())))
    assert sorted(seen) == list(range(ham.Ltrot + 1))
    for nt in (0, 2, 3, 5, 9, 10):
        want = O.brute_force_tau(s.m, s.f, nt, dps=30)
        for got, w in zip(seen[nt], want):
            assert rel(got, w) < 1e-9, nt
    assert s.ctrl.NCG_tau == 4 * len([t for t in s.stab[1:] if t >= 1])
    assert s.ctrl.XMAX_tau < 1e-9


def test_tau_m_free_fermions_closed_form():
    """U = 0: B(tau, 0) = e^{-tau T} exactly, so G(tau, 0) = e^{-tau T} (1 + e^{-beta T})^-1,
    G(tau, tau) = G(0, 0), G(beta, 0) = 1 - G(0, 0) and G(0, tau) = -(1 - G00) e^{tau T}."""
    ham = hubbard(L1=4, L2=4, beta=2.0, nwrap=5, U=0.0, chem=0.3)
    assert ham.n_op == 0
    s = oracle(ham, np.zeros((0, ham.Ltrot), dtype=np.int64), tape(10))
    s.init()
    T = ham.hop_matrix
    n, dt = ham.Ndim, ham.params.Dtau
    G0 = np.linalg.inv(np.eye(n) + expm(-2.0 * T))
    assert rel(s.G, G0) < 1e-11
    seen = {}
    s.tau_m(hook=lambda nt, a, b, c, d: seen.__setitem__(nt, (a.copy(), b.copy(), c.copy(), d.copy())))
    for nt, (GT0, G0T, G00, GTT) in seen.items():
        tau = nt * dt
        assert rel(GT0, expm(-tau * T) @ G0) < 1e-10
        assert rel(G0T, -(np.eye(n) - G0) @ expm(tau * T)) < 1e-10
        assert rel(GTT, G0) < 1e-10 and rel(G00, G0) < 1e-12
    assert rel(seen[ham.Ltrot][0], np.eye(n) - G0) < 1e-10


def _tau_obs(ham):
    from alf_amd.host import lattices as LT
    return HubbardTauObs(ham.Ltrot, ham.N_SUN, LT.distance_classes(ham.latt), ham.List[:, 0], ham.List[:, 1],
                         ham.unit.Norb)


@pytest.mark.parametrize("lattice,L,nphi", [("Square", 4, 0), ("Honeycomb", 3, 1)])
def test_tau_measurements_at_tau_zero_equal_the_equal_time_ones(lattice, L, nphi):
    """At tau = 0, G0T = -(1 - G): SpinZ(tau = 0) and Den(tau = 0) are the equal-time SpinZ and
    Den of G(0); Green(tau = beta) sums 1 - G(0); the Den background sums Z_I of G(tau, tau)
    over every tau; one TAU_M call counts once."""
    from alf_amd.host import lattices as LT
    ham = hubbard(L1=L, L2=L, lattice=lattice, n_phi=nphi, beta=1.0, nwrap=5)
    f = fields_for(ham, 3)
    s = oracle(ham, f, tape(200000))
    s.init()
    s.sweep()
    s.tau_obs = _tau_obs(ham)
    gtt = []
    s.tau_m(hook=lambda nt, a, b, c, d: gtt.append(np.diag(d).copy()))
    eq = HubbardObs(ham.hop_matrix, ham.params.Ham_U, ham.N_SUN, LT.distance_classes(ham.latt), ham.List[:, 0],
                    ham.List[:, 1], ham.unit.Norb)
    eq.measure(s.G, s.phase)
    t = s.tau_obs
    assert t.n == 1 and abs(t.sign - np.sign(s.phase.real)) == 0
    assert rel(t.latt[1, 0], eq.latt[1]) < 1e-12
    assert rel(t.latt[2, 0], eq.latt[2]) < 1e-12
    # the background is accumulated at every tau from G(tau, tau)
    w0 = s.phase / s.phase.real * np.sign(s.phase.real)
    want0 = np.zeros(ham.unit.Norb, dtype=complex)
    for d in gtt:
        np.add.at(want0, ham.List[:, 1], ham.N_SUN * (1.0 - d) * w0)
    assert len(gtt) == ham.Ltrot + 1 and rel(t.den0, want0) < 1e-12
    # Green(tau = beta) is the sum over pairs of 1 - G00
    n = ham.Ndim
    dist = LT.distance_classes(ham.latt)[ham.List[:, 0][:, None], ham.List[:, 0][None, :]]
    w = s.phase / s.phase.real * np.sign(s.phase.real)
    want = np.zeros_like(t.latt[0, 0])
    np.add.at(want, (dist, np.broadcast_to(ham.List[:, 1][:, None], (n, n)),
                     np.broadcast_to(ham.List[:, 1][None, :], (n, n))), (np.eye(n) - s.G) * w)
    assert rel(t.latt[0, ham.Ltrot], want) < 1e-9


def test_ed_time_displaced_correlators():
    """oracle/ed.py's Trotterised <A(t) B>: at U = 0 the one-flavour Green's function is
    e^{-t dtau T} (1 + e^{-beta T})^-1 and the density correlation follows from Wick's theorem;
    at t = 0 the spin and density correlators equal the equal-time ones."""
    from oracle.ed import HubbardED
    ham = hubbard(L1=2, L2=2, U=0.0, beta=1.0, dtau=0.1, chem=0.2)
    T = ham.hop_matrix
    ed = HubbardED(T, 0.0)
    tau = ed.trotter_tau(0.1, ham.Ltrot)
    n = ham.Ndim
    G0 = np.linalg.inv(np.eye(n) + expm(-1.0 * T))
    for t in range(ham.Ltrot + 1):
        Gt = expm(-0.1 * t * T) @ G0
        assert rel(tau["Green"][t], Gt) < 1e-10
        # <n_i(t) n_j> = <n_i><n_j> + 2 <c_i^+(t) c_j> <c_i(t) c_j^+> for two free flavours
        G0t = -(np.eye(n) - G0) @ expm(0.1 * t * T)               # G(0, t)
        dens = 2 * (1 - np.diag(G0))
        want = dens[:, None] * dens[None, :] - 2 * G0t.T * Gt
        assert rel(tau["Den"][t], want) < 1e-10
    ed4 = HubbardED(T, 4.0)
    tau4 = ed4.trotter_tau(0.1, ham.Ltrot)
    eq = ed4.observables(ed4.trotter_rho(0.1, ham.Ltrot))
    assert rel(tau4["SpinZ"][0], eq["SpinZ"]) < 1e-12 and rel(tau4["Den"][0], eq["Den"]) < 1e-12
