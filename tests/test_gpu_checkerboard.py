"""Checkerboard Op_T on the GPU against the oracle, through the C ABI.

The oracle applies the 2 x 2 bond exponentials one by one in the reference's order
(Hop_mod.F90:143-250); the device applies a family at a time from LDS strips.  Tolerances
as for the dense path: G and PHASE within 1e-10 relative, fields, acceptance counters and
random-stream positions bit-exact on the same random tape."""
import numpy as np
import pytest

from helpers import fields_for, hubbard, oracle, rel, tape
from oracle import dqmc_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def rt():
    from alf_amd import runtime
    runtime.load_library()
    return runtime


def engine(rt, ham, f, u, delay=8):
    e = rt.Engine(ham, delay=delay)
    e.set_fields(f)
    e.set_tape(u)
    return e


@pytest.mark.parametrize("left", [False, True])
@pytest.mark.parametrize("lattice,L,nphi", [("Square", 6, 0), ("Honeycomb", 3, 1)])
def test_checkerboard_udv_chains(rt, lattice, L, nphi, left):
    """WRAPUR (D P U, bonds Ncheck..1) and WRAPUL (P^+ conj(D) U, Hop_mod_mmthlc)."""
    ham = hubbard(L1=L, L2=L, lattice=lattice, n_phi=nphi, beta=3.0, nwrap=10, checkerboard=True)
    f = fields_for(ham, 2)
    s = oracle(ham, f, tape(10), presort=True)
    e = engine(rt, ham, f, tape(10))
    n = ham.Ndim
    if left:
        e.udv_reset(rt.ALFQ_UDVL, "L")
        e.wrapul(30, 20)
        e.wrapul(20, 10)
        U, D, V = e.get_udv(rt.ALFQ_UDVL)
        uo = s.wrapul(20, 10, s.wrapul(30, 20, O.UDV.identity(n, "L")))
        got, want = (U * D) @ V.conj().T, (uo.U * uo.D) @ uo.V.conj().T
    else:
        e.udv_reset(rt.ALFQ_UDVR, "R")
        e.wrapur(0, 10)
        e.wrapur(10, 20)
        U, D, V = e.get_udv(rt.ALFQ_UDVR)
        uo = s.wrapur(10, 20, s.wrapur(0, 10, O.UDV.identity(n, "R")))
        got, want = (U * D) @ V, (uo.U * uo.D) @ uo.V
    assert rel(got, want) < 1e-12
    e.close()


@pytest.mark.parametrize("lattice,L,nphi", [("Square", 4, 0), ("Honeycomb", 3, 1)])
def test_checkerboard_wrapgr_single_slices(rt, lattice, L, nphi):
    ham = hubbard(L1=L, L2=L, lattice=lattice, n_phi=nphi, beta=2.0, nwrap=10, checkerboard=True)
    f = fields_for(ham, 3)
    u = tape(100000)
    s = oracle(ham, f, u)
    s.init()
    e = engine(rt, ham, f, u, delay=5)
    e.set_green(s.G)
    e.set_phase(s.phase)
    for ntau in range(0, 4):
        s.G = s.wrapgrup(s.G, ntau)
        ph = e.wrapgrup(ntau, s.phase if ntau == 0 else e.get_phase())
        assert rel(e.get_green(), s.G) < TOL
        assert abs(ph - s.phase) < TOL
    for ntau in range(4, 1, -1):
        s.G = s.wrapgrdo(s.G, ntau)
        ph = e.wrapgrdo(ntau, e.get_phase())
        assert rel(e.get_green(), s.G) < TOL
    assert np.array_equal(e.get_fields(), s.f)
    e.close()


@pytest.mark.parametrize("lattice,L,nphi,chem,beta,nwrap,delay", [
    ("Square", 4, 0, 0.0, 4.0, 10, 32),
    ("Square", 6, 0, 0.3, 3.0, 10, 7),
    ("Honeycomb", 3, 1, 0.0, 2.0, 8, 6),
    ("Honeycomb", 4, 2, 0.2, 2.0, 5, 16),
])
def test_checkerboard_sweeps_match_oracle(rt, lattice, L, nphi, chem, beta, nwrap, delay):
    ham = hubbard(L1=L, L2=L, lattice=lattice, n_phi=nphi, chem=chem, beta=beta, nwrap=nwrap, U=4.0,
                  checkerboard=True)
    f = fields_for(ham, 0)
    u = tape(2_000_000)
    s = oracle(ham, f, u)
    e = engine(rt, ham, f, u, delay=delay)
    s.init()
    e.init()
    assert rel(e.get_green(), s.G) < TOL
    for _ in range(2):
        s.sweep()
        e.sweep(1)
        assert np.array_equal(e.get_fields(), s.f)
        assert rel(e.get_green(), s.G) < TOL
        assert abs(e.get_phase() - s.phase) < TOL
        assert e.rng_position() == s.rng.pos
    c = e.control()
    assert c["NC_up"] == s.ctrl.NC_up and c["ACC_up"] == s.ctrl.ACC_up
    assert c["XMAXG"] < 1e-9 and c["XMAXP"] < 1e-9
    e.close()


def test_checkerboard_batch_matches_single_chains(rt):
    """Chains of one batch with their own bond exponentials (different chemical potential,
    hence different diagonals) follow the trajectories of the same chains run alone; the
    24 x 24 lattice (Ndim 576) uses narrower LDS strips than the 16 x 16 one."""
    for L in (16, 24):
        hams = [hubbard(L1=L, L2=L, beta=0.5, nwrap=5, U=U, chem=mu, checkerboard=True)
                for U, mu in ((3.0, 0.0), (4.0, 0.2))]
        fs = [fields_for(h, 50 + z) for z, h in enumerate(hams)]
        b = rt.BatchEngine(hams, delay=16)
        for z, f in enumerate(fs):
            b.chain(z).set_fields(f)
            b.chain(z).set_philox(7, z)
        b.init()
        b.sweep(1)
        for z, (h, f) in enumerate(zip(hams, fs)):
            e = rt.Engine(h, delay=16)
            e.set_fields(f)
            e.set_philox(7, z)
            e.init()
            e.sweep(1)
            b.chain(z)
            assert np.array_equal(b.get_fields(), e.get_fields())
            assert rel(b.get_green(), e.get_green()) < TOL
            assert b.control()["XMAXG"] < 1e-8
            e.close()
        b.close()


def test_checkerboard_large_lattice_matches_oracle(rt):
    """Square 16 x 16 (Ndim 256, the benchmark lattice) for one sweep."""
    ham = hubbard(L1=16, L2=16, beta=1.0, nwrap=10, U=4.0, checkerboard=True)
    f = fields_for(ham, 12)
    u = tape(2_000_000)
    s = oracle(ham, f, u)
    e = engine(rt, ham, f, u, delay=32)
    s.init()
    e.init()
    s.sweep()
    e.sweep(1)
    assert np.array_equal(e.get_fields(), s.f)
    assert rel(e.get_green(), s.G) < TOL
    assert abs(e.get_phase() - s.phase) < TOL
    e.close()


def test_checkerboard_desc_is_validated(rt):
    """Bonds of one family that share a site, sites outside 1..Ndim and family lengths
    that do not add up are refused before anything runs (ERROR_HAMILTONIAN)."""
    import ctypes as C
    ham = hubbard(L1=4, L2=4, beta=1.0, checkerboard=True)
    e = rt.Engine(ham)
    d, keep = rt.checkerboard_desc(ham)
    p = keep["p"].copy()
    p[1] = p[0]                                   # bond 2 shares bond 1's sites, same family
    bad = [("p", p), ("fam", keep["fam"] + np.array([1] + [0] * (len(keep["fam"]) - 1), dtype=np.int32))]
    q = keep["p"].copy()
    q[0, 0] = ham.Ndim + 1
    bad.append(("p", q))
    for key, arr in bad:
        k2 = dict(keep)
        k2[key] = np.ascontiguousarray(arr, dtype=np.int32)
        d2 = rt.CheckerboardDesc(d.nbond, d.nfam, rt._iptr(k2["fam"]), rt._iptr(k2["p"]), rt._dptr(k2["E"]),
                                 rt._dptr(k2["Ei"]))
        with pytest.raises(rt.AlfqError) as ex:
            e._c(e.lib.alfq_set_checkerboard(e.ctx, C.byref(d2)))
        assert ex.value.code == 6
    e.close()
