"""Host-core baseline of the sweep: the FP64 oracle timed on the CPU.

TEST INFRASTRUCTURE ONLY.  Only the ``cpu_baseline`` leg of ``bench.py`` runs this
module (as child processes); nothing in the shipped path imports it.

Three modes, as the reference is deployed (Documentation/performance.tex:29-40):

* throughput: one single-threaded chain per host core, all started together (the
  reference's recommended layout, one MPI walker per core with OMP_NUM_THREADS = 1);
  aggregate sweeps/s = total timed sweeps / wall time of the timed window;
* single core: one single-threaded chain alone;
* latency: one chain with the BLAS threads spread over all the cores used.

Each chain is the 16 x 16 U = 4 beta = 10 Hubbard workload (or the one the arguments
name) with its own uniform four-state field (seed i) and its own random tape; one sweep
is main.F90:714-887 (up pass, down pass, slice-0 refresh), timed after the set-up and
the initial CGR.  The median over the timed sweeps is reported.

    python3 -m oracle.cpu_baseline --procs 16 --sweeps 2 --latency-sweeps 1
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
THREAD_VARS = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "BLIS_NUM_THREADS")


def host_info() -> dict:
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    lapack = []
    try:
        from threadpoolctl import threadpool_info
        import numpy  # noqa: F401  (loads the BLAS the oracle calls)
        import scipy.linalg  # noqa: F401
        for p in threadpool_info():
            if p.get("user_api") == "blas":
                lapack.append(f"{p.get('internal_api')} {p.get('version')} ({os.path.basename(p.get('filepath', ''))})")
    except Exception as e:       # reported, not fatal
        lapack.append(f"unknown ({e})")
    return {"nproc": os.cpu_count(), "cpu_model": model, "lapack": sorted(set(lapack))}


def worker(args) -> None:
    """One chain: set up, report ready on stdout, wait for 'go' on stdin, time the sweeps."""
    import numpy as np

    sys.path.insert(0, ROOT)
    from alf_amd.host.hamiltonian import HubbardParams, hubbard_ham_set
    from oracle.dqmc_oracle import OracleSweep, Tape, model_from_ham

    ham = hubbard_ham_set(HubbardParams(Lattice_type="Square", L1=args.L, L2=args.L, Ham_U=args.U,
                                        Beta=args.beta, Dtau=0.1, Nwrap=10))
    ham.nsigma.set_uniform_four_state(args.seed)
    m = model_from_ham(ham)
    visits = 2 * m.n_op * m.ltrot
    tape = Tape(np.random.default_rng(1000 + args.seed).random(3 * visits * args.sweeps + 16))
    s = OracleSweep(m, ham.nsigma.f.copy(), tape)
    s.init()
    print("ready", flush=True)
    if sys.stdin.readline().strip() != "go":
        return
    t0 = time.time()
    times = []
    for _ in range(args.sweeps):
        t = time.perf_counter()
        s.sweep()
        times.append(time.perf_counter() - t)
    print(json.dumps({"t0": t0, "t1": time.time(), "sweep_s": times, "acc": s.ctrl.ACC_up / max(s.ctrl.NC_up, 1)}),
          flush=True)


def run_group(procs: int, threads: int, args) -> dict:
    """procs chains with `threads` BLAS threads each, started together after their set-up."""
    env = dict(os.environ)
    for k in THREAD_VARS:
        env[k] = str(threads)
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--worker", "--L", str(args.L), "--beta", str(args.beta),
           "--U", str(args.U), "--sweeps", str(args.sweeps)]
    ps = [subprocess.Popen(cmd + ["--seed", str(i)], cwd=ROOT, env=env, stdin=subprocess.PIPE,
                           stdout=subprocess.PIPE, text=True) for i in range(procs)]
    try:
        for p in ps:
            if p.stdout.readline().strip() != "ready":
                raise RuntimeError("cpu baseline worker failed during set-up")
        go = time.time()
        for p in ps:
            p.stdin.write("go\n")
            p.stdin.flush()
        outs = [json.loads(p.stdout.readline()) for p in ps]
    finally:
        for p in ps:
            try:
                p.stdin.close()
            except OSError:
                pass
            p.wait(timeout=600)
    wall = max(o["t1"] for o in outs) - go
    sweeps = [t for o in outs for t in o["sweep_s"]]
    return {"chains": procs, "blas_threads": threads, "sweeps": len(sweeps), "wall_s": wall,
            "aggregate_sweeps_per_s": len(sweeps) / wall, "median_sweep_s": statistics.median(sweeps),
            "acceptance": statistics.mean(o["acc"] for o in outs)}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--worker", action="store_true")
    ap.add_argument("--procs", type=int, default=16)
    ap.add_argument("--sweeps", type=int, default=2, help="timed sweeps per chain (throughput, single core)")
    ap.add_argument("--latency-sweeps", type=int, default=1, help="timed sweeps of the threaded chain (0: skip)")
    ap.add_argument("--L", type=int, default=16)
    ap.add_argument("--beta", type=float, default=10.0)
    ap.add_argument("--U", type=float, default=4.0)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    if args.worker:
        worker(args)
        return
    t = time.perf_counter()
    info = host_info()
    out = dict(info, workload=f"Hubbard {args.L}x{args.L} square, U={args.U}, beta={args.beta}, dtau=0.1, Nwrap=10")
    # (name, chains, BLAS threads per chain, sweeps per chain)
    modes = [("throughput", args.procs, 1, args.sweeps), ("single_core", 1, 1, args.sweeps),
             ("latency", 1, args.procs, args.latency_sweeps)]
    for name, procs, threads, sweeps in modes:
        if sweeps <= 0:
            continue
        print(f"[cpu baseline] {name}: {procs} chain(s) x {sweeps} sweeps, {threads} BLAS thread(s) each",
              file=sys.stderr, flush=True)
        args.sweeps = sweeps
        out[name] = run_group(procs, threads, args)
    out["elapsed_s"] = time.perf_counter() - t
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
