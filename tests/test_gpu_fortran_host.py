"""The Fortran host (ISO_C_BINDING module alfq_mod + alfq_ring_demo) against the oracle.

The demo builds a 4-site Hubbard ring entirely in Fortran (closed-form
exp(-dtau T)), drives the engine through the C ABI with a random tape read from
a file, and writes the final fields, G, PHASE and counters.  The oracle runs
the same model and tape.  A second run of the demo sets a checkerboard Op_T with
bonds of unequal amplitude, the ring's measurement tables and TAU_M from Fortran
and writes the time-displaced sums as well.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from helpers import rel
from oracle import dqmc_oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "alf_amd", "fortran", "alfq_ring_demo")


def ring_model():
    n, dtau, U = 4, 0.1, 4.0
    e = np.zeros((n, n))
    ei = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            d = (i - j) % n
            e[i, j] = (math.exp(2 * dtau) + 2 * math.cos(math.pi * d / 2) + math.exp(-2 * dtau) * (-1) ** d) / 4
            ei[i, j] = (math.exp(-2 * dtau) + 2 * math.cos(math.pi * d / 2) + math.exp(2 * dtau) * (-1) ** d) / 4
    g = np.sqrt(complex(-dtau * U / 2.0))
    return O.OracleModel(ndim=4, ltrot=20, nwrap=5, n_sun=2, mat=e.astype(complex), invmat=ei.astype(complex),
                         op_P=np.arange(4), op_E=np.ones(4), op_g=np.full(4, g), op_alpha=np.full(4, -0.5 + 0j),
                         op_type=np.full(4, 2))


def demo_fields():
    f = np.zeros((4, 20), dtype=np.int64)
    for nt in range(1, 21):
        for i in range(1, 5):
            f[i - 1, nt - 1] = (1 if (i + nt) % 2 == 0 else -1) * (2 if (i * nt) % 3 == 0 else 1)
    return f


def test_fortran_host_sweeps_match_oracle(tmp_path):
    if not os.path.exists(DEMO):
        pytest.fail("alf_amd/fortran/alfq_ring_demo is missing: run __graft_entry__.build()")
    u = np.random.default_rng(21).random(5000)
    tape_file = tmp_path / "tape.txt"
    with open(tape_file, "w") as fh:
        fh.write(f"{u.size}\n")
        fh.write("\n".join(repr(float(x)) for x in u) + "\n")
    out_file = tmp_path / "out.txt"
    r = subprocess.run([DEMO, str(tape_file), str(out_file)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = open(out_file).read().split()
    f_gpu = np.array([int(x) for x in vals[:80]]).reshape(20, 4).T
    rest = [float(x) for x in vals[80:80 + 34]]
    G_gpu = np.array([complex(rest[2 * k], rest[2 * k + 1]) for k in range(16)]).reshape(4, 4, order="F")
    phase_gpu = complex(rest[32], rest[33])
    nc, acc = int(vals[114]), int(vals[115])

    m = ring_model()
    s = O.OracleSweep(m, demo_fields(), O.Tape(u))
    s.init()
    s.sweep()
    s.sweep()
    assert np.array_equal(f_gpu, s.f)
    assert nc == s.ctrl.NC_up and acc == s.ctrl.ACC_up
    assert rel(G_gpu, s.G) < 1e-10
    assert abs(phase_gpu - s.phase) < 1e-10
    Gb, _ = O.brute_force_green(m, s.f, 0, dps=30)
    assert rel(G_gpu, Gb) < 1e-10


def ring_checkerboard_model():
    """The ring with the checkerboard Op_T of the demo: bonds (1,2) (3,4) | (2,3) (4,1) (0-based
    here) with amplitudes t_b, E_b = exp(dtau t_b sigma_x); the dense matrices are
    P = E_1 E_2 E_3 E_4 and P^-1."""
    dtau = 0.1
    p = np.array([[0, 1], [2, 3], [1, 2], [3, 0]])
    tb = (1.0, 0.8, 1.2, 0.9)
    E = np.array([[[math.cosh(dtau * t), math.sinh(dtau * t)], [math.sinh(dtau * t), math.cosh(dtau * t)]]
                  for t in tb], dtype=complex)
    Ei = np.array([np.linalg.inv(e) for e in E])

    def dense(mats, order):
        M = np.eye(4, dtype=complex)
        for b in order:
            e = np.eye(4, dtype=complex)
            e[np.ix_(p[b], p[b])] = mats[b]
            M = M @ e
        return M
    m = ring_model()
    m.mat, m.invmat = dense(E, range(4)), dense(Ei, range(3, -1, -1))
    m.cb_p, m.cb_E, m.cb_Einv = p, E, Ei
    return m


def test_fortran_host_checkerboard_tau_matches_oracle(tmp_path):
    """alfq_set_checkerboard, alfq_obs_setup, alfq_tau_enable and alfq_tau_get called from
    Fortran: fields, G, PHASE and the time-displaced sums after two sweeps equal the oracle's
    with the same bonds, tape and measurement tables."""
    from oracle.observables import HubbardTauObs
    if not os.path.exists(DEMO):
        pytest.fail("alf_amd/fortran/alfq_ring_demo is missing: run __graft_entry__.build()")
    u = np.random.default_rng(22).random(5000)
    tape_file = tmp_path / "tape.txt"
    with open(tape_file, "w") as fh:
        fh.write(f"{u.size}\n")
        fh.write("\n".join(repr(float(x)) for x in u) + "\n")
    out_file, cb_file = tmp_path / "out.txt", tmp_path / "cb.txt"
    r = subprocess.run([DEMO, str(tape_file), str(out_file), str(cb_file)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = open(cb_file).read().split()
    f_gpu = np.array([int(x) for x in vals[:80]]).reshape(20, 4).T
    rest = [float(x) for x in vals[80:114]]
    G_gpu = np.array([complex(rest[2 * k], rest[2 * k + 1]) for k in range(16)]).reshape(4, 4, order="F")
    phase_gpu = complex(rest[32], rest[33])
    nc, acc, ntau_calls, ncg_tau = (int(x) for x in vals[114:118])
    sign, xmax_tau = float(vals[118]), float(vals[119])
    z = [float(x) for x in vals[120:]]
    cz = np.array([complex(z[2 * k], z[2 * k + 1]) for k in range(len(z) // 2)])
    assert cz.size == 3 * 21 * 4 + 1
    latt_gpu, den0_gpu = cz[:-1].reshape(3, 21, 4), cz[-1]

    m = ring_checkerboard_model()
    s = O.OracleSweep(m, demo_fields(), O.Tape(u))
    imj = (np.arange(4)[:, None] - np.arange(4)[None, :]) % 4
    s.tau_obs = HubbardTauObs(20, 2, imj, np.arange(4), np.zeros(4, dtype=int), 1)
    s.init()
    s.sweep()
    s.sweep()
    assert np.array_equal(f_gpu, s.f)
    assert nc == s.ctrl.NC_up and acc == s.ctrl.ACC_up
    assert rel(G_gpu, s.G) < 1e-10
    assert abs(phase_gpu - s.phase) < 1e-10
    t = s.tau_obs
    assert ntau_calls == t.n == 2 and abs(sign - t.sign) < 1e-12
    assert ncg_tau == s.ctrl.NCG_tau and xmax_tau < 1e-8
    assert rel(latt_gpu, t.latt[:, :, :, 0, 0]) < 1e-9
    assert abs(den0_gpu - t.den0[0]) < 1e-9 * abs(t.den0[0])
    Gb, _ = O.brute_force_green(m, s.f, 0, dps=30)
    assert rel(G_gpu, Gb) < 1e-10
    # the families do not commute, so the bond order shows in G: the reverse order is
    # another model, far outside the tolerance above
    m2 = ring_checkerboard_model()
    m2.cb_p, m2.cb_E, m2.cb_Einv = m2.cb_p[::-1].copy(), m2.cb_E[::-1].copy(), m2.cb_Einv[::-1].copy()
    g_fwd, g_rev = O.OracleSweep(m, s.f, O.Tape(u)), O.OracleSweep(m2, s.f, O.Tape(u))
    g_fwd.init()
    g_rev.init()
    assert rel(g_fwd.G, G_gpu) < 1e-10 and rel(g_rev.G, G_gpu) > 1e-6
