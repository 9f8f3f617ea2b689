"""Equal-time measurements on the GPU against the oracle, through the C ABI.

The engine and the oracle run the same sweeps on the same random tape and measure on the
same slices (Obser of the Hubbard plugin, N_FL = 1); the sums of every observable, the
counts and the sign sums must agree.  Tolerance 1e-10 relative (FP64 path)."""
import numpy as np
import pytest

from helpers import fields_for, hubbard, oracle, rel, tape
from oracle.observables import HubbardObs

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def rt():
    from alf_amd import runtime
    runtime.load_library()
    return runtime


def obs_oracle(ham):
    from alf_amd.host import lattices as LT
    return HubbardObs(ham.hop_matrix, ham.params.Ham_U, ham.N_SUN, LT.distance_classes(ham.latt),
                      ham.List[:, 0], ham.List[:, 1], ham.unit.Norb)


def check_sums(got, ob):
    assert got["n_scal"] == ob.n_scal and got["n_eq"] == ob.n_eq
    assert abs(got["sign_scal"] - ob.sign_scal) < 1e-12 and abs(got["sign_eq"] - ob.sign_eq) < 1e-12
    assert rel(got["scal"], ob.scal) < TOL
    assert rel(got["latt"], ob.latt) < TOL
    assert rel(got["den0"], ob.den0) < TOL


@pytest.mark.parametrize("lattice,L,beta,nwrap,n_phi,chem,lobs", [
    ("Square", 4, 2.0, 10, 0, 0.0, None),
    ("Square", 6, 1.5, 5, 0, 0.3, (3, 12)),
    ("Honeycomb", 3, 1.0, 8, 1, 0.0, None),
])
def test_measurements_match_oracle(rt, lattice, L, beta, nwrap, n_phi, chem, lobs):
    ham = hubbard(L1=L, L2=L, beta=beta, nwrap=nwrap, lattice=lattice, n_phi=n_phi, chem=chem)
    f = fields_for(ham, 5)
    u = tape(2_000_000)
    s = oracle(ham, f, u)
    s.obs = obs_oracle(ham)
    s.lobs = lobs
    e = rt.Engine(ham, delay=8)
    e.set_fields(f)
    e.set_tape(u)
    e.obs_setup(lobs=lobs)
    e.obs_enable(True)
    s.init()
    e.init()
    for _ in range(2):
        s.sweep()
        e.sweep(1)
    assert np.array_equal(e.get_fields(), s.f)
    got = e.obs_get()
    lo, hi = lobs if lobs else (1, ham.Ltrot)
    per_sweep = sum(1 for t in range(1, ham.Ltrot + 1) if lo <= t <= hi) + \
        sum(1 for t in range(0, ham.Ltrot) if lo <= t <= hi)
    assert got["n_scal"] == 2 * per_sweep
    check_sums(got, s.obs)
    e.obs_reset()
    z = e.obs_get()
    assert z["n_scal"] == 0 and z["n_eq"] == 0 and not np.any(z["latt"]) and not np.any(z["scal"])
    e.sweep(1)
    assert e.obs_get()["n_scal"] == per_sweep
    e.obs_enable(False)
    e.sweep(1)
    assert e.obs_get()["n_scal"] == per_sweep
    e.close()


def test_batched_measurements_match_single_chains(rt):
    """Every chain of a batch measures with its own couplings (U, hopping)."""
    hams = [hubbard(L1=4, L2=4, beta=1.5, nwrap=5, U=U) for U in (4.0, 2.0)]
    fs = [fields_for(h, 40 + z) for z, h in enumerate(hams)]
    b = rt.BatchEngine(hams, delay=8)
    for z, f in enumerate(fs):
        b.chain(z).set_fields(f)
        b.chain(z).set_philox(9, z)
    b.obs_setup()
    b.obs_enable(True)
    b.init()
    b.sweep(2)
    for z, (h, f) in enumerate(zip(hams, fs)):
        e = rt.Engine(h, delay=8)
        e.set_fields(f)
        e.set_philox(9, z)
        e.obs_setup()
        e.obs_enable(True)
        e.init()
        e.sweep(2)
        one = e.obs_get()
        b.chain(z)
        bat = b.obs_get()
        assert one["n_scal"] == bat["n_scal"] > 0
        assert rel(bat["scal"], one["scal"]) < TOL and rel(bat["latt"], one["latt"]) < TOL
        e.close()
    b.close()


@pytest.mark.parametrize("lattice,L,n_phi,checkerboard", [
    ("Square", 4, 0, False), ("Square", 6, 0, True), ("Honeycomb", 3, 1, False), ("Honeycomb", 4, 2, True),
])
def test_symm_measurements_match_oracle(rt, lattice, L, n_phi, checkerboard):
    """Symm = .T.: the measurements see Hop_mod_Symm(G) = exp(-dtau T/2) G exp(dtau T/2), dense
    or bond by bond over the symmetrised checkerboard families; sums equal the oracle's."""
    ham = hubbard(L1=L, L2=L, beta=1.0, nwrap=5, lattice=lattice, n_phi=n_phi, checkerboard=checkerboard,
                  symm=True)
    f = fields_for(ham, 6)
    u = tape(2_000_000, seed=3)
    s = oracle(ham, f, u)
    assert s.m.symm
    s.obs = obs_oracle(ham)
    e = rt.Engine(ham, delay=8)
    e.set_fields(f)
    e.set_tape(u)
    e.obs_setup()
    e.obs_enable(True)
    s.init()
    e.init()
    for _ in range(2):
        s.sweep()
        e.sweep(1)
    assert np.array_equal(e.get_fields(), s.f)
    check_sums(e.obs_get(), s.obs)
    with pytest.raises(rt.AlfqError):
        e.tau_enable(True)
    e.close()
