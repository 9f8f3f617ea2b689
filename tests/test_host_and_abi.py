"""CPU checks of the host-side plugin mirror and of the C ABI library (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import hubbard

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_square_lattice_sites_and_neighbours():
    ham = hubbard(L1=4, L2=6, beta=1.0)
    latt = ham.latt
    assert ham.Ndim == 24 and latt.N == 24
    # every cell has distinct +-a1, +-a2 neighbours and the relation is symmetric
    for I in range(latt.N):
        nn = {int(latt.nnlist[I, 2, 1]), int(latt.nnlist[I, 0, 1]), int(latt.nnlist[I, 1, 2]), int(latt.nnlist[I, 1, 0])}
        assert len(nn) == 4 and I not in nn
        assert latt.nnlist[int(latt.nnlist[I, 2, 1]), 0, 1] == I
    T = ham.hop_matrix
    assert np.allclose(T, T.T)
    assert np.all(np.count_nonzero(T, axis=1) == 4)     # z = 4, zero chemical potential


def test_honeycomb_sites():
    ham = hubbard(L1=3, L2=3, beta=1.0, lattice="Honeycomb")
    assert ham.Ndim == 18 and ham.unit.Norb == 2
    T = ham.hop_matrix
    assert np.all(np.count_nonzero(np.abs(T) > 1e-12, axis=1) == 3)   # z = 3


def test_hubbard_plugin_shapes():
    ham = hubbard(L1=4, L2=4, beta=5.0, nwrap=10)
    assert ham.Ltrot == 50 and len(ham.Op_V) == 16 and ham.nsigma.f.shape == (16, 50)
    assert sorted(int(o.P[0]) for o in ham.Op_V) == list(range(16))
    assert ham.S0(0, 0, 1) == 1.0


def test_fields_conf_roundtrip(tmp_path):
    ham = hubbard(L1=2, L2=2, beta=0.5)
    ham.nsigma.set_uniform_four_state(3)
    f = ham.nsigma.f.copy()
    p = tmp_path / "confin_0"
    ham.nsigma.write_conf(str(p), [8752143])
    ham.nsigma.f[:] = 1
    seeds = ham.nsigma.read_conf(str(p), 1)
    assert seeds == [8752143] and np.array_equal(ham.nsigma.f, f)
    assert set(np.unique(f)) <= {-2, -1, 1, 2}


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "alfq.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(alfq_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    from alf_amd import runtime
    lib = runtime.load_library()
    declared = _header_symbols()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(runtime.EXPORTS)
    assert lib.alfq_version().startswith(b"alfq")


def test_abi_rejects_bad_descriptors_without_touching_the_device():
    from alf_amd import runtime
    lib = runtime.load_library()
    h = ctypes.c_void_p()
    assert lib.alfq_ctx_create(None, ctypes.byref(h)) == 1
    d = runtime.ModelDesc()
    d.ndim, d.ltrot, d.nwrap, d.n_sun, d.n_fl, d.n_op = 4, 10, 5, 2, 2, 0
    assert lib.alfq_ctx_create(ctypes.byref(d), ctypes.byref(h)) == 6      # ERROR_HAMILTONIAN
    assert b"N_FL" in lib.alfq_last_error()
    d.n_fl = 1
    d.ndim = 0
    assert lib.alfq_ctx_create(ctypes.byref(d), ctypes.byref(h)) == 6
    d.ndim = 4
    P = (ctypes.c_int32 * 1)(7)
    T = (ctypes.c_int32 * 1)(2)
    d.n_op, d.op_P, d.op_type = 1, P, T
    assert lib.alfq_ctx_create(ctypes.byref(d), ctypes.byref(h)) == 6      # site out of range
    P[0], T[0] = 1, 3
    assert lib.alfq_ctx_create(ctypes.byref(d), ctypes.byref(h)) == 6      # continuous field type
    assert lib.alfq_set_delay(None, 4) == 1
    assert lib.alfq_sweep(None, 1) == 1
