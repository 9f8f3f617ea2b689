"""The run loop on the GPU: bins, configurations and restarts."""
import os
import shutil

import numpy as np
import pytest

from alf_amd.host.hamiltonian import HubbardParams
from alf_amd.obs_io import jackknife_scal, read_scal
from alf_amd.run import read_conf, simulate

pytestmark = pytest.mark.gpu


def test_restart_continues_the_same_chain(tmp_path):
    """Two bins in one run equal one bin, a stop, a restart from confin and one more bin:
    the same fields and random-stream position, and the same bin values to roundoff (a
    restart rebuilds G from the fields, the running sweep had it from its last CGR)."""
    p = HubbardParams(L1=4, L2=4, Beta=2.0, Ham_U=4.0, Nwrap=10)
    a, b = tmp_path / "a", tmp_path / "b"
    ia = simulate(p, nsweep=3, nbin=2, directory=str(a), seed=7)
    assert ia["bins"] == 2 and not (a / "RUNNING").exists() and (a / "info").exists()
    simulate(p, nsweep=3, nbin=1, directory=str(b), seed=7)
    shutil.copy(b / "confout_0", b / "confin_0")
    simulate(p, nsweep=3, nbin=1, directory=str(b), seed=7)
    sa, fa = read_conf(str(a / "confout_0"), 16, 20)
    sb, fb = read_conf(str(b / "confout_0"), 16, 20)
    assert sa == sb and np.array_equal(fa, fb)
    for name in ("Kin", "Pot", "Part", "Ener"):
        va, _ = read_scal(str(a / f"{name}_scal"))
        vb, _ = read_scal(str(b / f"{name}_scal"))
        assert va.shape == vb.shape == (2, 1)
        assert np.allclose(va, vb, rtol=1e-9, atol=1e-12)
    (best, err), = jackknife_scal(str(a / "Ener_scal"))
    assert np.isfinite(best.real) and err.real >= 0
    for name in ("Green", "SpinZ", "Den"):
        assert (a / f"{name}_eq").exists() and (a / f"{name}_eq_info").exists()


def test_run_with_time_displaced_bins(tmp_path):
    """Ltau = 1: every bin appends Green_tau, SpinZ_tau and Den_tau (Ntau = Ltrot + 1) with
    their info files next to the equal-time ones."""
    p = HubbardParams(L1=4, L2=4, Beta=1.0, Ham_U=4.0, Nwrap=5, Checkerboard=True)
    info = simulate(p, nsweep=2, nbin=2, directory=str(tmp_path), seed=3, ltau=1)
    assert info["bins"] == 2
    nc, no, nt = 16, 1, 11
    for name in ("Green", "SpinZ", "Den"):
        lines = open(tmp_path / f"{name}_tau").read().splitlines()
        assert len(lines) == 2 * (1 + no + nc * (1 + nt * no * no))
        head = lines[0].split()
        assert [int(v) for v in head[1:4]] == [no, nc, nt]
        assert (tmp_path / f"{name}_tau_info").exists()
