"""Throughput of the gfx950 DQMC sweep engine on one MI355X.

Workload (the single-chain, real-size case): Hubbard model on a 16x16 square
lattice (Ndim = 256), t = 1, U = 4, beta = 10, dtau = 0.1 (Ltrot = 100),
Nwrap = 10, N_SUN = 2, N_FL = 1, dense exp(-dtau T), complex128 Green's
function, discrete four-state HS field drawn uniformly from seed 0, Philox
random stream.  One step = one sweep of ALF's sequential finite-temperature
update (up pass, down pass, slice-0 refresh), all on the device.

Prints one JSON line (rank 0).  The CPU baseline is the FP64 oracle (numpy /
LAPACK restatement of the same sweep) on the host cores: one single-threaded chain per
core (at most 16), all at once, timed before the GPU is touched; the one-core and
threaded-BLAS single-chain rates stand beside it.
Beside the single chain, the line reports the aggregate rate of a batch of
independent chains of the same workload (distinct seeds) advancing in one
context, every kernel launch covering all of them.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

def note(msg):
    """A progress line on stderr (stdout carries only the result line)."""
    print(f"[bench {time.perf_counter() - T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


T0 = time.perf_counter()
WORKLOAD = dict(Lattice_type="Square", L1=16, L2=16, ham_T=1.0, Ham_U=4.0, Beta=10.0, Dtau=0.1, Nwrap=10)


def algorithmic_flops_per_sweep(n, L, S, acc, complex_t=False):
    """F = n^3 (6 L c_T + 117.3 S) + 16 acc L n^3 with c_T = 4 (real T) or 8."""
    cT = 8 if complex_t else 4
    return n ** 3 * (6 * L * cT + 117.3 * S) + 16 * acc * L * n ** 3


def attainable_sweep_seconds(n, L, S, acc, complex_t=False, peak=78.6e12, bw=8.0e12, delay=1):
    """Time-weighted roofline of one sweep: sum over the kernel classes of max(F_k / P, B_k / BW)
    with the per-sweep work model of the reference's algorithm: wraps F = 6 L c_T n^3,
    B = 2 L (3 (T_B + 32) + 64) n^2; rank-1 updates F = 16 acc L n^3, B = 64 acc L n^3 (G streamed
    once per update) divided by `delay` for updates delayed to rank-`delay` flushes (G streamed
    once per flush); QR and solves F = 117.3 S n^3."""
    cT, TB = (8, 16) if complex_t else (4, 8)
    t_wrap = max(6 * L * cT * n ** 3 / peak, 2 * L * (3 * (TB + 32) + 64) * n ** 2 / bw)
    t_rank1 = max(16 * acc * L * n ** 3 / peak, 64 * acc * L * n ** 3 / (delay * bw))
    t_qr = 117.3 * S * n ** 3 / peak
    return t_wrap + t_rank1 + t_qr


def cpu_baseline_run(procs, sweeps=2, latency_sweeps=1):
    """The FP64 oracle on the host cores, in child processes started before this process
    touches the GPU and finished before it does: throughput mode (one single-threaded chain of
    the workload per core, all at once), one single-threaded chain alone, and one chain with
    threaded BLAS on all of them.  Returns the parsed report or an error note."""
    import subprocess
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--procs", str(procs), "--sweeps", str(sweeps),
           "--latency-sweeps", str(latency_sweeps)]
    r = subprocess.run(cmd, cwd=HERE, stdout=subprocess.PIPE, text=True, timeout=900)
    if r.returncode != 0:
        return {"error": f"oracle.cpu_baseline exited with {r.returncode}"}
    rep = json.loads(r.stdout.strip().splitlines()[-1])
    thr = rep["throughput"]
    out = {"value": thr["aggregate_sweeps_per_s"], "unit": "sweeps/s", "cores": procs, "kind": "port",
           "sample": (f"throughput mode: {procs} single-threaded chains of the 16x16 U=4 beta=10 workload "
                      f"(seeds 0..{procs - 1}) x {sweeps} sweeps each, started together on {procs} host cores, "
                      "run by the numpy/LAPACK FP64 oracle (ZGEQP3 pivoting, ZGERU rank-1 updates, real "
                      "exp(-dtau T)); value = total sweeps / wall time of the timed window"),
           "median_sweep_s_throughput": thr["median_sweep_s"],
           "nproc": rep["nproc"], "cpu_model": rep["cpu_model"], "lapack": rep["lapack"]}
    if "single_core" in rep:
        out["single_core_sweeps_per_s"] = 1.0 / rep["single_core"]["median_sweep_s"]
    if "latency" in rep:
        out["latency_mode_sweeps_per_s"] = 1.0 / rep["latency"]["median_sweep_s"]
        out["latency_mode_threads"] = procs
    return out


def cpu_oracle_rate(ham, fields, sweeps=3):
    """Median sweeps/s of the oracle on one host thread for a small model (the 4x4 case)."""
    from threadpoolctl import threadpool_limits

    from oracle.dqmc_oracle import OracleSweep, Tape, model_from_ham
    m = model_from_ham(ham)
    rng = Tape(np.random.default_rng(12345).random(3 * m.n_op * m.ltrot * 2 * sweeps + 16))
    times = []
    with threadpool_limits(1):
        s = OracleSweep(m, fields, rng)
        s.init()
        for _ in range(sweeps):
            t0 = time.perf_counter()
            s.sweep()
            times.append(time.perf_counter() - t0)
    return 1.0 / float(np.median(times))


BATCH_THERMALISATION = 15


def batched_rate(chains, steps, prec=0, checkerboard=False, params=None):
    """Aggregate sweeps/s of `chains` independent chains of the workload (or of the model
    `params` names) in one batched context."""
    from alf_amd.host.hamiltonian import HubbardParams, hubbard_ham_set
    from alf_amd.runtime import BatchEngine
    ham = hubbard_ham_set(HubbardParams(**(WORKLOAD if params is None else params), Checkerboard=checkerboard))
    b = BatchEngine([ham] * chains, delay=32)
    b.set_precision(prec)
    for z in range(chains):
        ham.nsigma.set_uniform_four_state(z)
        b.chain(z).set_fields(ham.nsigma.f.copy())
        b.chain(z).set_philox(seed=z, stream=0)
    b.init()
    # thermalisation from the random start: the first sweep runs eagerly and the next two
    # capture both graph parities; the acceptance, and with it the number of delayed-update
    # rounds per slice (set by the busiest slice of any chain), settles within ~15 sweeps
    b.sweep(BATCH_THERMALISATION)
    ms = b.sweep_timed(steps)
    ctl = [b.chain(z).control() for z in range(chains)]
    b.close()
    return {"chains": chains, "steps": steps, "thermalisation_sweeps": BATCH_THERMALISATION,
            "ms_per_batched_sweep": ms / steps,
            "aggregate_sweeps_per_s": chains * steps / (ms / 1e3),
            "precision_green_max": max(c["XMAXG"] for c in ctl),
            "precision_green_mean": float(np.mean([c["precision_green_mean"] for c in ctl])),
            "precision_green_mean_worst_chain": float(max(c["precision_green_mean"] for c in ctl)),
            "precision_phase_max": max(c["XMAXP"] for c in ctl)}


def serving_rate(nreq, nsweeps, rate, seed=0):
    """p50/p99 time-to-result and aggregate sweeps/s of a Poisson stream of Hubbard requests
    served by the continuous-batching loop (alf_amd.serving)."""
    from alf_amd.serving import Server, latency_summary, poisson_requests
    reqs = poisson_requests(nreq, rate, seed=seed, nsweeps=nsweeps)
    srv = Server(min_slots=2)
    res = srv.run(reqs)
    out = latency_summary(res, srv.wall)
    out.update({"sample": f"{nreq} requests x {nsweeps} sweeps (the serving configuration asks 500 sweeps each), "
                          f"Poisson arrivals at {rate}/s, square L in {{8,12,16,24,32}} weighted 4:3:2:1:1 "
                          "(Ndim 64..1024), beta in {4,8,16} (Ltrot 40..160), U in {2,4,6}, Nwrap 10, FP64",
                "peak_concurrent_batches": srv.peak_batches, "joined_running_lanes": srv.joined,
                "largest_batch": max(r.batch_size for r in res),
                "precision_green_max": max(r.control["XMAXG"] for r in res)})
    return out


def c4_mix_rate(chains, nsweeps, seed=0):
    """The batched-chains mix: `chains` independent Hubbard chains, square L uniform in
    {6, 8, ..., 20} (Ndim 36..400), U uniform in [2, 8], beta uniform in [2, 12] (Ltrot =
    nint(beta / 0.1)), Nwrap 10, chain i seeded with i, all at once: one batch of uneven Ltrot
    per lattice size (every kernel launch covers the size's chains; a chain sweeps its batch's
    largest Ltrot with identity slices beyond its own), the batches concurrent on their own
    streams.  The steady rate is the extra sweeps of a second run with more sweeps per chain
    over its extra wall time, so model set-up, context creation, init and download cancel."""
    from alf_amd.serving import Request, _HubbardBackend, run_uneven
    rng = np.random.default_rng(seed)
    draws = [(int(rng.choice(np.arange(6, 21, 2))), float(rng.uniform(2.0, 12.0)), float(rng.uniform(2.0, 8.0)))
             for _ in range(chains)]
    be = _HubbardBackend()
    for i, (L, b, U) in enumerate(draws):          # host model set-up once, outside the timed runs
        be.model(Request(rid=i, L=L, beta=b, U=U, seed=i, nsweeps=1))
        if i % 32 == 31:
            note(f"uneven mix: {i + 1} models set up")
    out = {"chains": chains, "distinct_shapes": len({(L, int(round(b / 0.1))) for L, b, _ in draws})}
    for label, ratio in (("one_batch_per_size", None), ("batches_cut_at_ltrot_ratio_2", 2.0)):
        walls, nb, worst = {}, 0, 0.0
        for sw in (nsweeps, 3 * nsweeps):
            reqs = [Request(rid=i, L=L, beta=b, U=U, seed=i, nsweeps=sw) for i, (L, b, U) in enumerate(draws)]
            res, wall, nb = run_uneven(reqs, max_ratio=ratio, backend=be)
            note(f"uneven mix, {label}: {sw} sweeps per chain in {wall:.1f} s")
            walls[sw] = wall
            worst = max(worst, max(r.control["XMAXG"] for r in res))
        steady = chains * 2 * nsweeps / (walls[3 * nsweeps] - walls[nsweeps])
        out[label] = {"batches": nb, "aggregate_sweeps_per_s": steady,
                      "wall_s": walls, "with_setup_sweeps_per_s": chains * nsweeps / walls[nsweeps],
                      "precision_green_max": worst}
    out["aggregate_sweeps_per_s"] = max(out[k]["aggregate_sweeps_per_s"] for k in
                                        ("one_batch_per_size", "batches_cut_at_ltrot_ratio_2"))
    return out


def measured_rate(ham, fields, delay, steps):
    """Sweeps/s of one chain with the equal-time measurements on (2 Ltrot - 1 per sweep) and
    the measured energy and double occupancy of those sweeps."""
    from alf_amd.runtime import Engine
    eng = Engine(ham, delay=delay)
    eng.set_fields(fields)
    eng.set_philox(seed=0, stream=0)
    eng.obs_setup()
    eng.init()
    eng.sweep(2)
    eng.obs_enable(True)
    eng.sweep(1)                  # captures the sweep with its measurements
    eng.obs_reset()
    ms = eng.sweep_timed(steps)
    s = eng.obs_get()
    eng.close()
    n, u = ham.Ndim, ham.params.Ham_U
    return {"sweeps_per_s": steps / (ms / 1e3), "measurements_per_sweep": 2 * ham.Ltrot - 1,
            "energy_per_site": float((s["scal"][3] / s["n_scal"]).real) / n,
            "double_occupancy": float((s["scal"][1] / s["n_scal"]).real) / (u * n),
            "note": f"{steps} sweeps after 3 from a random start (not a thermalised estimate)"}


def tau_rate(ham, fields, delay, steps):
    """Sweeps/s of one chain with the equal-time and the time-displaced measurements on
    (Ltau = 1: TAU_M with CGR2_2 and ObserT after every sweep) and Precision tau."""
    from alf_amd.runtime import Engine
    eng = Engine(ham, delay=delay)
    eng.set_fields(fields)
    eng.set_philox(seed=0, stream=0)
    eng.obs_setup()
    eng.obs_enable(True)
    eng.tau_enable(True)
    eng.init()
    eng.sweep(2)                  # the eager sweep, then the captured one
    eng.obs_reset()
    ms = eng.sweep_timed(steps)
    t = eng.tau_get()
    eng.close()
    return {"sweeps_per_s": steps / (ms / 1e3), "tau_slices_per_sweep": ham.Ltrot + 1,
            "precision_tau_max": t["xmax_tau"], "precision_tau_mean": t["xmean_tau_sum"] / max(t["ncg_tau"], 1)}


def config_rates(delay, cpu=True):
    """The other single-chain configurations: the 4x4 U=4 beta=5 case (with the oracle's
    rate on one host core beside it) and the 12x12 honeycomb with a pi/4 flux per plaquette
    (complex hopping), U=3, beta=12, Nwrap=8, each on the dense and the checkerboard hopping."""
    from alf_amd.host.hamiltonian import HubbardParams, hubbard_ham_set
    out = {}
    cases = {
        "square_4x4_U4_beta5": dict(Lattice_type="Square", L1=4, L2=4, Ham_U=4.0, Beta=5.0, Dtau=0.1, Nwrap=10),
        "honeycomb_12x12_flux_pi4_U3_beta12": dict(Lattice_type="Honeycomb", L1=12, L2=12, Ham_U=3.0, Beta=12.0,
                                                   Dtau=0.1, Nwrap=8, N_Phi=18),
    }
    for name, kw in cases.items():
        row = {}
        for cb in (False, True):
            ham = hubbard_ham_set(HubbardParams(**kw, Checkerboard=cb))
            ham.nsigma.set_uniform_four_state(0)
            f = ham.nsigma.f.copy()
            r = single_rate(ham, f, delay, 5 if ham.Ndim > 64 else 20, 2, 0)
            row["checkerboard" if cb else "dense"] = {k: r[k] for k in ("sweeps_per_s", "precision_green_max")}
            if cpu and not cb and ham.Ndim <= 16:
                row["cpu_oracle_sweeps_per_s"] = cpu_oracle_rate(ham, f)
        if kw["L1"] * kw["L2"] <= 16:
            # a lattice this small leaves a single chain latency-bound: many chains per launch
            r = batched_rate(256, 5, params=kw)
            row["dense_batch_256"] = {k: r[k] for k in ("aggregate_sweeps_per_s", "ms_per_batched_sweep",
                                                        "precision_green_max")}
        row["ndim"] = ham.Ndim
        row["ltrot"] = ham.Ltrot
        out[name] = row
    return out


def paired_error(ham, fields, delay, sweeps=1):
    """The split-product chain and the FP64 drop-in on the same random tape from the same fields:
    both rebuild the same G from the same fields at every stabilisation point, so there
    |G_split - G_FP64| <= (Precision Green of the split run) + (that of the FP64 run)."""
    from alf_amd.runtime import PREC_FP64, PREC_INT8_SPLIT, Engine
    u = np.random.default_rng(77).random(3 * 2 * ham.Ndim * ham.Ltrot * sweeps + 64)
    logs, fl = [], []
    for prec in (PREC_FP64, PREC_INT8_SPLIT):
        e = Engine(ham, delay=delay)
        e.set_precision(prec)
        e.set_fields(fields)
        e.set_tape(u)
        e.init()
        e.sweep(sweeps)
        logs.append(e.precision_log())
        fl.append(e.get_fields())
        e.close()
    a, b = logs
    return {"sweeps": sweeps, "stabilisations": int(len(b)), "same_fields": bool(np.array_equal(fl[0], fl[1])),
            "precision_green_max_fp64": float(a[:, 0].max()), "precision_green_max_split": float(b[:, 0].max()),
            "g_err_bound_max": float((a[:, 0] + b[:, 0]).max()) if len(a) == len(b) else float("inf")}


def single_rate(ham, fields, delay, steps, warmup, prec):
    """Sweeps/s of one chain of the workload and its Precision Green/Phase statistics."""
    from alf_amd.runtime import Engine
    eng = Engine(ham, delay=delay)
    eng.set_precision(prec)
    eng.set_fields(fields)
    eng.set_philox(seed=0, stream=0)
    eng.init()
    for _ in range(warmup):
        eng.sweep(1)
    eng.sync()
    eng.control_reset()
    ms = eng.sweep_timed(steps)
    ctl = eng.control()
    eng.close()
    return {"sweeps_per_s": steps / (ms / 1e3), "ms_per_sweep": ms / steps,
            "precision_green_mean": ctl["precision_green_mean"], "precision_green_max": ctl["XMAXG"],
            "precision_phase_max": ctl["XMAXP"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--delay", type=int, default=32)
    ap.add_argument("--batch", type=int, default=256, help="chains of the batched line (0: skip it)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32-wrap legs")
    ap.add_argument("--no-checkerboard", action="store_true", help="skip the checkerboard-hopping legs")
    ap.add_argument("--no-configs", action="store_true", help="skip the 4x4 and 12x12-honeycomb single chains")
    ap.add_argument("--c4-chains", type=int, default=256, help="chains of the uneven-size mix (0: skip it)")
    ap.add_argument("--c4-sweeps", type=int, default=4)
    ap.add_argument("--only-headline", action="store_true",
                    help="the single-chain line and its roofline only (for kernel profiles of that run)")
    ap.add_argument("--serve-requests", type=int, default=24, help="requests of the serving leg (0: skip it)")
    ap.add_argument("--serve-sweeps", type=int, default=20)
    args = ap.parse_args()
    # several batches and serving lanes run on their own streams: give them hardware queues
    # of their own (before the first GPU call; the single chain uses one stream either way)
    from alf_amd.serving import use_hw_queues
    use_hw_queues()
    if args.only_headline:
        args.batch, args.no_mixed, args.no_checkerboard, args.no_configs = 0, True, True, True
        args.serve_requests, args.no_cpu_baseline, args.c4_chains = 0, True, 0
    if args.gpus != 1:
        print("bench.py: this engine runs one Markov chain per GPU and this build measures one GPU only; "
              "independent chains shard across GPUs without communication, so use --gpus 1", file=sys.stderr)
        sys.exit(2)

    # the host-core baseline runs first, in child processes, before this process touches the GPU
    cpu = None if args.no_cpu_baseline else cpu_baseline_run(min(16, os.cpu_count() or 1))

    from alf_amd.host.hamiltonian import HubbardParams, hubbard_ham_set
    from alf_amd.runtime import Engine
    from alf_amd import kernelbench

    note("single chain (headline)")
    ham = hubbard_ham_set(HubbardParams(**WORKLOAD))
    ham.nsigma.set_uniform_four_state(0)
    fields = ham.nsigma.f.copy()
    eng = Engine(ham, delay=args.delay)
    eng.set_fields(fields)
    eng.set_philox(seed=0, stream=0)
    eng.init()
    for _ in range(args.warmup):
        eng.sweep(1)
    eng.sync()
    eng.control_reset()
    t0 = time.perf_counter()
    ms = eng.sweep_timed(args.steps)
    wall = time.perf_counter() - t0
    ctl = eng.control()
    n, L = ham.Ndim, ham.Ltrot
    S = (L + ham.Nwrap - 1) // ham.Nwrap
    acc = ctl["ACC_up"] / max(ctl["NC_up"], 1)
    sweeps_per_s = args.steps / (ms / 1e3)

    # dominant kernel: its average launch duration measured with HIP events on the engine's stream
    kb = kernelbench.measure_dominant(eng, sweeps=1)
    result = {
        "metric": "MC sweeps/s per GPU (FP64 drop-in path)",
        "value": sweeps_per_s,
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "Hubbard 16x16 square, Ndim=256, U=4, beta=10, dtau=0.1 (Ltrot=100), Nwrap=10, "
                               "complex128 G, 4-state HS field from seed 0, Philox stream",
                   "path": "FP64 drop-in", "delay_k": args.delay,
                   "acceptance": acc,
                   "precision_green_mean": ctl["precision_green_mean"], "precision_green_max": ctl["XMAXG"],
                   "precision_phase_max": ctl["XMAXP"],
                   "wall_s": wall,
                   "algorithmic_gflop_per_sweep": algorithmic_flops_per_sweep(n, L, S, acc) / 1e9,
                   "sweep_fraction_of_fp64_peak": algorithmic_flops_per_sweep(n, L, S, acc) * sweeps_per_s / 78.6e12,
                   "sweep_roofline_fraction": attainable_sweep_seconds(n, L, S, acc) * sweeps_per_s,
                   "sweep_roofline_fraction_delayed": attainable_sweep_seconds(n, L, S, acc, delay=args.delay) *
                                                      sweeps_per_s},
        "roofline": kb,
    }
    eng.close()
    if not args.only_headline:
        note("with measurements")
        result["with_measurements"] = measured_rate(ham, fields, args.delay, args.steps)
        result["with_time_displaced"] = tau_rate(ham, fields, args.delay, max(2, args.steps // 2))
    if args.batch > 0:
        note(f"batch of {args.batch}, FP64")
        result["batched"] = batched_rate(args.batch, max(2, min(args.steps, 5)))
        # priced for the rank-32 delayed flushes the engine runs (the per-flip ZGERU streaming of
        # the reference's form would put the batch above 1, so that form is not reported here)
        agg = result["batched"]["aggregate_sweeps_per_s"]
        result["batched"]["sweep_roofline_fraction_delayed"] = attainable_sweep_seconds(n, L, S, acc, delay=32) * agg
    # the chosen precision form: hopping wraps of G as exact int8 digit products on the matrix
    # cores (8 digits of 7 bits per operand), everything else FP64; FP32 wraps beside it
    from alf_amd.runtime import PREC_FP32_WRAPS, PREC_INT8_SPLIT
    mixed = None
    if not args.no_mixed:
        note("mixed precision forms")
        mixed = {"form": "hopping wraps of G (Hop_mod products in WRAPGRUP/WRAPGRDO) as int8 split products: each "
                         "operand row/column = 2^e x 8 base-128 int8 digits rounded to nearest, digit pairs with s+t <= 9 "
                         "multiplied exactly on v_mfma_i32_32x32x32_i8 (int32 sums), combined in FP64; UDV chain, "
                         "QR, CGR, Metropolis ratio and delayed updates FP64",
                 "single": single_rate(ham, fields, args.delay, args.steps, args.warmup, PREC_INT8_SPLIT),
                 "vs_fp64_same_fields": paired_error(ham, fields, args.delay)}
        if args.batch > 0:
            mixed["batched"] = batched_rate(args.batch, max(2, min(args.steps, 5)), PREC_INT8_SPLIT)
            fp64_b = result.get("batched")
            if fp64_b:
                mixed["batched"]["fp64_precision_green_max_same_seeds"] = fp64_b["precision_green_max"]
                mixed["batched"]["rate_vs_fp64"] = (mixed["batched"]["aggregate_sweeps_per_s"] /
                                                    fp64_b["aggregate_sweeps_per_s"])
        # G err of the mixed form against FP64 (the paired run's bound at every stabilisation)
        # and, for every chain of the batch, the mean Precision Green (the reference's criterion,
        # Documentation/stabilization.tex:225)
        mixed["criterion"] = ("|G_mixed - G_FP64| <= 1e-8 at every stabilisation of the paired single chain, "
                              "and mean Precision Green <= 1e-8 for every chain of the batch")
        mixed["meets_green_1e-8"] = bool(mixed["vs_fp64_same_fields"]["g_err_bound_max"] <= 1e-8 and
                                         mixed["single"]["precision_green_max"] <= 1e-8 and
                                         mixed.get("batched", {}).get("precision_green_mean_worst_chain", 0.0) <= 1e-8)
        fp32 = {"form": "hopping wraps of G with FP32 operands, FP32 MFMA and FP32 accumulation",
                "single": single_rate(ham, fields, args.delay, args.steps, args.warmup, PREC_FP32_WRAPS)}
        if args.batch > 0:
            fp32["batched"] = batched_rate(args.batch, max(2, min(args.steps, 5)), PREC_FP32_WRAPS)
        mixed["fp32_wraps"] = fp32
        result["mixed"] = mixed
    if not args.no_checkerboard:
        note("checkerboard")
        # the Hubbard plugin's default hopping: the checkerboard product of bond exponentials
        cb_ham = hubbard_ham_set(HubbardParams(**WORKLOAD, Checkerboard=True))
        cb = {"form": "Checkerboard=.T.: 4 families of 2x2 bond exponentials applied from LDS strips "
                      "(HBM-bound, 2 x 16 n^2 bytes per application) instead of the dense n^3 products",
              "single": single_rate(cb_ham, fields, args.delay, args.steps, args.warmup, 0)}
        if args.batch > 0:
            cb["batched"] = batched_rate(args.batch, max(2, min(args.steps, 5)), 0, checkerboard=True)
        result["checkerboard"] = cb
    if args.serve_requests > 0:
        note("serving")
        result["serving"] = serving_rate(args.serve_requests, args.serve_sweeps, 50.0)
    if not args.no_configs:
        note("other configurations")
        result["other_configs"] = config_rates(args.delay, cpu=not args.no_cpu_baseline)
    if args.c4_chains > 0:
        note("uneven mix")
        result["uneven_batch_mix"] = c4_mix_rate(args.c4_chains, args.c4_sweeps)
    # attainable issue rates on this device (published dense peaks: FP64 78.6, FP32 157.3 TF/s,
    # int8 5000 TOP/s)
    from alf_amd.runtime import mfma_rate
    if not args.only_headline:
        result["device_ceilings_tflops"] = {"fp64_mfma": mfma_rate(0), "fp32_mfma": mfma_rate(1),
                                            "fp64_vector_fma": mfma_rate(2), "int8_mfma_tops": mfma_rate(3)}
    if cpu is not None:
        result["cpu_baseline"] = cpu
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
