"""Configuration files and the run lock of the run loop (no device needed)."""
import numpy as np
import pytest

from alf_amd.host.hamiltonian import HubbardParams
from alf_amd.run import read_conf, simulate, write_conf


def test_configuration_file_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    f = rng.choice([-2, -1, 1, 2], size=(16, 20))
    p = str(tmp_path / "confout_0")
    write_conf(p, f, (12345, 3, 987654321))
    lines = open(p).read().splitlines()
    assert len(lines) == 1 + 16 * 20
    assert lines[1].strip() == str(f[0, 0]) and lines[2].strip() == str(f[1, 0])    # operators inner
    assert lines[17].strip() == str(f[0, 1])                                       # then the next slice
    seed, g = read_conf(p, 16, 20)
    assert seed == [12345, 3, 987654321] and np.array_equal(g, f)
    with pytest.raises(ValueError):
        read_conf(p, 16, 21)


def test_running_lock_refuses_a_second_run(tmp_path):
    (tmp_path / "RUNNING").write_text("ALF is running\n")
    with pytest.raises(RuntimeError, match="already running"):
        simulate(HubbardParams(L1=4, L2=4, Beta=1.0), nsweep=1, nbin=1, directory=str(tmp_path))
    assert (tmp_path / "RUNNING").exists()        # the lock of the other run is left alone


def test_failed_run_keeps_the_lock(tmp_path, monkeypatch):
    """A run that raises leaves RUNNING in place (main.F90:961-965 removes it only after a
    finished run), so the next run is refused until its files are checked."""
    import alf_amd.run as run

    class FailingEngine:
        def __init__(self, *a, **k):
            raise RuntimeError("device error during the run")

    monkeypatch.setattr(run, "Engine", FailingEngine)
    with pytest.raises(RuntimeError, match="device error"):
        simulate(HubbardParams(L1=4, L2=4, Beta=1.0), nsweep=1, nbin=1, directory=str(tmp_path))
    assert (tmp_path / "RUNNING").exists()
    with pytest.raises(RuntimeError, match="already running"):
        simulate(HubbardParams(L1=4, L2=4, Beta=1.0), nsweep=1, nbin=1, directory=str(tmp_path))
