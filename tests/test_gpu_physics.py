"""End to end: sweeps and measurements on the GPU against exact diagonalisation.

On a 2 x 2 Hubbard cluster (4 sites, 256 states) the expectation of every equal-time
estimator of the sweep is exactly Tr[O B^L] / Tr[B^L] with B = exp(-dtau H_V) exp(-dtau H_T)
(oracle/ed.py).  A batch of independent chains gives independent bins; the Monte Carlo
averages must lie within 4 jackknife errors (plus the O(dtau^4) per-slice error of the
4-state field decoupling) of the exact values: Kin, Pot, Part, Ener and the Green, spin
and density correlations by distance class."""
import numpy as np
import pytest

from helpers import hubbard

pytestmark = pytest.mark.gpu


def test_monte_carlo_matches_exact_diagonalisation():
    from alf_amd import runtime as rt
    from alf_amd.host import lattices as LT
    from alf_amd.obs_io import jackknife
    from oracle.ed import HubbardED

    U, beta, dtau, nwrap, chains = 4.0, 2.0, 0.1, 10, 256
    ham = hubbard(L1=2, L2=2, U=U, beta=beta, dtau=dtau, nwrap=nwrap)
    b = rt.BatchEngine([ham] * chains, delay=8)
    for z in range(chains):
        ham.nsigma.set_uniform_four_state(500 + z)
        b.chain(z).set_fields(ham.nsigma.f.copy())
        b.chain(z).set_philox(77, z)
    b.obs_setup()
    b.init()
    b.sweep(20)                      # thermalise
    b.obs_enable(True)
    b.obs_reset()
    b.sweep(100)
    per = []
    for z in range(chains):
        b.chain(z)
        s = b.obs_get()
        per.append(s)
        assert s["n_scal"] == 100 * (2 * ham.Ltrot - 1) and s["sign_scal"] == s["n_scal"]
    b.close()

    ed = HubbardED(ham.hop_matrix, U)
    ex = ed.observables(ed.trotter_rho(dtau, ham.Ltrot))
    slack = 2e-3
    names = ("Kin", "Pot", "Part", "Ener")
    for q, name in enumerate(names):
        data = np.array([[s["scal"][q] / s["n_scal"] for s in per], [1.0 + 0j] * chains])
        best, err = jackknife(lambda X: X[0] / X[1], data)
        assert abs(best.real - ex[name].real) < 4 * err.real + slack, (name, best, err, ex[name])
    imj = LT.distance_classes(ham.latt)
    site = ham.Invlist[:, 0]
    for q, name in enumerate(("Green", "SpinZ", "Den")):
        for d in range(ham.latt.N):
            want = sum(ex[name][site[I], site[J]] for I in range(4) for J in range(4) if imj[I, J] == d)
            data = np.array([[s["latt"][q, d, 0, 0] / s["n_eq"] for s in per], [1.0 + 0j] * chains])
            best, err = jackknife(lambda X: X[0] / X[1], data)
            assert abs(best.real - want.real) < 4 * err.real + slack, (name, d, best, err, want)


def test_time_displaced_monte_carlo_matches_exact_diagonalisation():
    """Ltau = 1 on the same cluster: the Monte Carlo averages of Green(tau), SpinZ(tau) and
    Den(tau) by distance class lie within 4 jackknife errors (plus the decoupling's slack) of
    the Trotterised <A(tau) B> = Tr[B^(L - t) A B^t B'] / Tr[B^L] (oracle/ed.py), at every
    fourth tau."""
    from alf_amd import runtime as rt
    from alf_amd.host import lattices as LT
    from alf_amd.obs_io import jackknife
    from oracle.ed import HubbardED

    U, beta, dtau, nwrap, chains = 4.0, 2.0, 0.1, 10, 256
    ham = hubbard(L1=2, L2=2, U=U, beta=beta, dtau=dtau, nwrap=nwrap)
    b = rt.BatchEngine([ham] * chains, delay=8)
    for z in range(chains):
        ham.nsigma.set_uniform_four_state(900 + z)
        b.chain(z).set_fields(ham.nsigma.f.copy())
        b.chain(z).set_philox(91, z)
    b.obs_setup()
    b.init()
    b.sweep(100)                     # the local moment thermalises slower than the energy
    b.tau_enable(True)
    b.obs_reset()
    b.sweep(100)
    per = []
    for z in range(chains):
        b.chain(z)
        s = b.tau_get()
        assert s["n"] == 100 and s["sign"] == 100.0
        assert s["xmax_tau"] < 1e-8
        per.append(s)
    b.close()

    ex = HubbardED(ham.hop_matrix, U).trotter_tau(dtau, ham.Ltrot)
    imj = LT.distance_classes(ham.latt)
    site = ham.Invlist[:, 0]
    slack = 2e-3
    for q, name in enumerate(("Green", "SpinZ", "Den")):
        for t in range(0, ham.Ltrot + 1, 4):
            for d in range(ham.latt.N):
                want = sum(ex[name][t, site[I], site[J]] for I in range(4) for J in range(4) if imj[I, J] == d)
                data = np.array([[s["latt"][q, t, d, 0, 0] / s["n"] for s in per], [1.0 + 0j] * chains])
                best, err = jackknife(lambda X: X[0] / X[1], data)
                assert abs(best.real - want.real) < 4 * err.real + slack, (name, t, d, best, err, want)
