"""Time-displaced measurements (TAU_M, CGR2_2, ObserT) on the GPU against the oracle,
through the C ABI.

The engine and the oracle run the same sweeps on the same random tape with Ltau = 1; the
time-displaced sums of Green, SpinZ and Den for every tau and distance class, the Den
background, the counts and the sign sums must agree (1e-9 relative to the largest entry:
the device's CGR2_2 factorises with presorted columns, the oracle with ZGEQP3), and
Precision tau counts the same number of comparisons with a small maximum."""
import numpy as np
import pytest

from helpers import fields_for, hubbard, oracle, rel, tape
from oracle.observables import HubbardTauObs

pytestmark = pytest.mark.gpu

TOL = 1e-9


@pytest.fixture(scope="module")
def rt():
    from alf_amd import runtime
    runtime.load_library()
    return runtime


def tau_oracle(ham):
    from alf_amd.host import lattices as LT
    return HubbardTauObs(ham.Ltrot, ham.N_SUN, LT.distance_classes(ham.latt), ham.List[:, 0], ham.List[:, 1],
                         ham.unit.Norb)


@pytest.mark.parametrize("lattice,L,beta,nwrap,n_phi,chem,cb", [
    ("Square", 4, 1.0, 5, 0, 0.0, False),
    ("Square", 6, 2.0, 10, 0, 0.3, False),
    ("Honeycomb", 3, 1.0, 4, 1, 0.0, False),
    ("Square", 4, 1.5, 5, 0, 0.2, True),
    ("Honeycomb", 3, 1.0, 5, 1, 0.1, True),
    # Ndim 144: CGR2_2 factorises a 288-row matrix, its first panel on the LDS panel kernel
    # and the later ones (at most 256 rows) on the register panel
    ("Square", 12, 0.5, 5, 0, 0.0, False),
])
def test_tau_measurements_match_oracle(rt, lattice, L, beta, nwrap, n_phi, chem, cb):
    ham = hubbard(L1=L, L2=L, beta=beta, nwrap=nwrap, lattice=lattice, n_phi=n_phi, chem=chem, checkerboard=cb)
    f = fields_for(ham, 7)
    u = tape(2_000_000)
    s = oracle(ham, f, u)
    s.tau_obs = tau_oracle(ham)
    e = rt.Engine(ham, delay=8)
    e.set_fields(f)
    e.set_tape(u)
    e.obs_setup()
    e.tau_enable(True)
    s.init()
    e.init()
    for _ in range(2):
        s.sweep()
        e.sweep(1)
    assert np.array_equal(e.get_fields(), s.f)
    got = e.tau_get()
    t = s.tau_obs
    assert got["n"] == t.n == 2 and abs(got["sign"] - t.sign) < 1e-12
    assert rel(got["latt"], t.latt) < TOL
    assert rel(got["den0"], t.den0) < TOL
    assert got["ncg_tau"] == s.ctrl.NCG_tau
    assert got["xmax_tau"] < 1e-8 and s.ctrl.XMAX_tau < 1e-8
    # bin boundary clears the sums; TAU_M can also be called on its own
    e.obs_reset()
    z = e.tau_get()
    assert z["n"] == 0 and not np.any(z["latt"])
    e.tau_m()
    s.tau_obs.reset()
    s.tau_m()
    one = e.tau_get()
    assert one["n"] == 1 and rel(one["latt"], s.tau_obs.latt) < TOL
    e.close()


def test_tau_batch_matches_single_chains(rt):
    """Chains of a batch with their own couplings measure the same time-displaced sums as
    the same chains alone (graph-replayed sweeps included)."""
    hams = [hubbard(L1=4, L2=4, beta=1.0, nwrap=5, U=U) for U in (4.0, 2.0)]
    fs = [fields_for(h, 60 + z) for z, h in enumerate(hams)]
    b = rt.BatchEngine(hams, delay=8)
    for z, f in enumerate(fs):
        b.chain(z).set_fields(f)
        b.chain(z).set_philox(3, z)
    b.obs_setup()
    b.tau_enable(True)
    b.init()
    b.sweep(3)
    for z, (h, f) in enumerate(zip(hams, fs)):
        e = rt.Engine(h, delay=8)
        e.set_fields(f)
        e.set_philox(3, z)
        e.obs_setup()
        e.tau_enable(True)
        e.init()
        e.sweep(3)
        one = e.tau_get()
        b.chain(z)
        bat = b.tau_get()
        assert one["n"] == bat["n"] == 3
        assert rel(bat["latt"], one["latt"]) < 1e-12 and rel(bat["den0"], one["den0"]) < 1e-12
        e.close()
    b.close()
