"""The Fortran host (ISO_C_BINDING module alfq_mod + alfq_ring_demo) against the oracle.

The demo builds a 4-site Hubbard ring entirely in Fortran (closed-form
exp(-dtau T)), drives the engine through the C ABI with a random tape read from
a file, and writes the final fields, G, PHASE and counters.  The oracle runs
the same model and tape.
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
