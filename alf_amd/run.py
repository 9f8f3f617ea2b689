"""The reference's run loop (Prog/main.F90) for the Hubbard plugin on the device.

``simulate`` does what ALF.out does for a finite-temperature Hubbard run with
-DOBS_LEGACY, minus MPI (one chain per call; ``rank`` names the files):

* a ``RUNNING`` lock in the run directory, refused if present (main.F90:211-239);
* the start from ``confin_<rank>`` if it exists, else a configuration drawn from the seed
  (Fields_in, Prog/Fields_mod.F90:337-506);
* NBin bins of NSweep sweeps with the equal-time measurements on every slice of
  LOBS_ST..LOBS_EN (main.F90:642-887) and, with Ltau = 1, the time-displaced ones after
  every sweep (main.F90:874-876), the bin files written at every bin end
  (``alf_amd.obs_io``) and the configuration saved to ``confout_<rank>``
  (Fields_write_conf, Fields_mod.F90:739-760: a seed line, then one field per line,
  operators inner, slices outer);
* the controlled stop when less than 1.5 bins fit into CPU_MAX hours
  (make_truncation, Prog/control_mod.F90:511-562);
* an ``info`` file with the precision and acceptance statistics (Control_Print,
  Prog/control_mod.F90:440-505).

For this engine the seed line holds the Philox key, the stream and the stream position, so a
restarted run continues the same random stream.
"""
from __future__ import annotations

import os
import time

import numpy as np

from alf_amd.host.hamiltonian import HubbardParams, hubbard_ham_set
from alf_amd.obs_io import BinWriter
from alf_amd.runtime import Engine

TRUNCATION_BINS = 1.5


def write_conf(path: str, fields: np.ndarray, seed_line) -> None:
    """Fields_write_conf for discrete fields (types 1 and 2): integers, slice-outer."""
    f = np.asarray(fields)
    with open(path, "w") as fh:
        fh.write(" " + " ".join(str(int(v)) for v in seed_line) + "\n")
        for nt in range(f.shape[1]):
            for i in range(f.shape[0]):
                fh.write(f" {int(f[i, nt]):d}\n")


def read_conf(path: str, n_op: int, ltrot: int):
    """Fields_read_conf: (seed line, fields[n_op, ltrot])."""
    with open(path) as fh:
        seed = [int(v) for v in fh.readline().split()]
        vals = [int(line.split()[0]) for line in fh if line.strip()]
    if len(vals) != n_op * ltrot:
        raise ValueError(f"{path}: {len(vals)} fields for {n_op} x {ltrot}")
    return seed, np.array(vals, dtype=np.int64).reshape(ltrot, n_op).T.copy()


def simulate(params: HubbardParams, nsweep: int, nbin: int, directory: str = ".", seed: int = 0, rank: int = 0,
             cpu_max_hours: float = 0.0, delay: int = 32, lobs=None, ltau: int = 0) -> dict:
    os.makedirs(directory, exist_ok=True)
    lock = os.path.join(directory, "RUNNING")
    if os.path.exists(lock):
        raise RuntimeError("ALF is already running or the previous run failed: remove RUNNING after checking "
                           "the data and configuration files")
    with open(lock, "w") as fh:
        fh.write(" ALF is running\n")
    t_start = time.time()
    # RUNNING stays in place when the run fails, as in the reference, which removes it only
    # after a finished run (main.F90:961-965): the bins and confout of a failed run need checking
    # before the next run appends to them.
    ham = hubbard_ham_set(params)
    eng = Engine(ham, delay=delay)
    confin = os.path.join(directory, f"confin_{rank}")
    stream = rank
    if os.path.exists(confin):
        line, f = read_conf(confin, ham.n_op, ham.Ltrot)
        seed, stream, pos = line
        eng.set_fields(f)
        eng.set_philox(seed, stream)
        eng.set_rng_position(pos)
    else:
        ham.nsigma.set_uniform_four_state(seed)
        eng.set_fields(ham.nsigma.f.copy())
        eng.set_philox(seed, stream)
    eng.obs_setup(lobs=lobs)
    eng.obs_enable(True)
    if ltau == 1:
        eng.tau_enable(True)
    eng.init()
    writer = BinWriter(directory, ham)
    done = 0
    truncated = False
    for _ in range(nbin):
        t_bin = time.time()
        eng.obs_reset()
        eng.sweep(nsweep)
        writer.write(eng.obs_get())
        if ltau == 1:
            writer.write_tau(eng.tau_get())
        write_conf(os.path.join(directory, f"confout_{rank}"), eng.get_fields(),
                   (seed, stream, eng.rng_position()))
        done += 1
        if cpu_max_hours > 0:
            t_end = time.time()
            remain = t_start + cpu_max_hours * 3600.0 - t_end
            if remain / max(t_end - t_bin, 1e-9) < TRUNCATION_BINS:
                truncated = True
                break
    c = eng.control()
    eng.close()
    elapsed = time.time() - t_start
    info = {"bins": done, "truncated": truncated, "precision_green_mean": c["precision_green_mean"],
            "precision_green_max": c["XMAXG"], "precision_phase_max": c["XMAXP"],
            "acceptance": c["ACC_up"] / max(c["NC_up"], 1), "cpu_time_s": elapsed}
    with open(os.path.join(directory, "info"), "a") as fh:
        fh.write(f" Precision Green  Mean, Max :  {info['precision_green_mean']:.16E}  "
                 f"{info['precision_green_max']:.16E}\n")
        fh.write(f" Precision Phase, Max       :  {info['precision_phase_max']:.16E}\n")
        fh.write(f" Acceptance                 :  {info['acceptance']:.16E}\n")
        fh.write(f" CPU Time                   :  {elapsed:.16E}\n")
    os.remove(lock)
    return info
