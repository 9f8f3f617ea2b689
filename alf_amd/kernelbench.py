"""Live roofline figure of the dominant kernel, measured inside the benchmark.

The dominant kernel of the sweep is the Metropolis block kernel
(chain_block_kernel).  Its launches are timed with start/stop events that each
launch records on the engine's stream (alfq_kernel_timing).  Algorithmic bytes:
every accepted flip is a rank-1 Sherman-Morrison update of the n x n complex128
Green's function, whose reference form (ZGERU, Prog/upgrade_mod.F90:269) reads and
writes G once: 32 n^2 bytes per accepted update.  achieved = accepted updates per
launch x 32 n^2 B / average launch duration; bound = HBM.  The kernel itself keeps
a 32 x 32 block of G in LDS and leaves the full-size products to the flush GEMMs,
so the figure is the rate at which it retires the reference's rank-1 updates.

``traffic`` (HBM bytes per launch) comes from hardware counters, which cannot be
read in the timed run: two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of the
benchmark, each scaled by what the same counter reads for a device-to-device copy
of known size (``python3 -m alf_amd.kernelbench copy``).  The summary of those runs
is committed beside the other profiles and read here.

    python3 -m alf_amd.kernelbench copy [--mib N]   # the calibration copy

The ``sweep`` mode runs sweeps of the 16 x 16 U = 4 beta = 10 Hubbard workload, one chain
and then a batch of chains, for per-kernel profiles of either hopping form:

    python3 -m alf_amd.kernelbench sweep [--checkerboard] [--tau] [--batch B] [--steps K] [--prec P]

The ``counters`` mode is the program for hardware-counter passes (rocprofv3 --pmc): one
chain and then a batch of chains of the same workload, each initialised and swept once with
every kernel launched from the host (no graph replay, no timing events), with a progress
line on stderr before and after every step.  A counter pass serialises the dispatches it
instruments, so the pass names the kernels to instrument (--kernel-include-regex) and the
program issues a bounded number of dispatches (about 5000 per sweep of one chain):

    python3 -m alf_amd.kernelbench counters [--batch B] [--checkerboard]

The ``traffic`` mode turns the two counter passes and the two calibration passes (their
rocprofv3 counter_collection.csv files) into the per-kernel table of bytes per dispatch:

    python3 -m alf_amd.kernelbench traffic FETCH.csv WRITE.csv COPY_FETCH.csv COPY_WRITE.csv

The ``shim`` mode times the up-pass slices of the same workload called one at a time, as
main.F90 calls WRAPGRUP: with G kept on the device, and with G uploaded and downloaded
around every call (the alfq_WRAPGRUP_shim form, whose rate includes PCIe):

    python3 -m alf_amd.kernelbench shim
"""
from __future__ import annotations

import ctypes as C
import os
import re
import sys

HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles",
                               "chain_block_kernel_16x16_U4_beta10_hbm_traffic.md")


def measured_traffic(path: str = TRAFFIC_SUMMARY):
    """HBM bytes per launch of chain_block_kernel from the committed counter summary, or None."""
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def measured_traffic_counters(path: str = TRAFFIC_SUMMARY):
    """Which counters the committed traffic figure sums ("FETCH_SIZE only" or both)."""
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"traffic_counters:\s*(.+)", text)
    return m.group(1).strip() if m else None


def measure_dominant(eng, sweeps: int = 1) -> dict:
    c0 = eng.control()
    eng.kernel_timing(True)
    eng.sweep(sweeps)
    total_ms, launches = eng.kernel_timing_get()
    eng.kernel_timing(False)
    c1 = eng.control()
    acc = c1["ACC_up"] - c0["ACC_up"]
    visits = c1["NC_up"] - c0["NC_up"]
    n = eng.n
    bytes_per_launch = acc * 32.0 * n * n / max(launches, 1)
    avg_s = total_ms * 1e-3 / max(launches, 1)
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    return {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": measured_traffic(),
        "traffic_counters": measured_traffic_counters(),
        "kernel": "chain_block_kernel (Metropolis decisions + delayed rank-1 updates)",
        "avg_launch_us": avg_s * 1e6,
        "launches": launches,
        "visits": visits,
        "accepted": acc,
        "algorithmic_bytes_per_launch": bytes_per_launch,
    }


def copy_calibration(nbytes: int, repeats: int = 3) -> None:
    """Device-to-device copies of nbytes (the HIP runtime's copy kernel), for scaling the
    FETCH_SIZE / WRITE_SIZE counters: each dispatch reads and writes exactly nbytes."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), nbytes) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(src, 1, nbytes) != 0 or hip.hipMemset(dst, 0, nbytes) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(repeats):
        if hip.hipMemcpy(dst, src, nbytes, 3) != 0:     # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    if hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"copied {nbytes} bytes device to device x {repeats}", flush=True)


def profile_sweeps(checkerboard: bool, batch: int, steps: int, tau: bool = False, prec: int = 0) -> None:
    """`steps` sweeps of one chain, then of `batch` chains in one batched context; with tau,
    the equal-time and time-displaced measurements on (TAU_M after every sweep); prec is the
    arithmetic of the hopping wraps (alfq_set_precision)."""
    from alf_amd.host.hamiltonian import HubbardParams, hubbard_ham_set
    from alf_amd.runtime import BatchEngine, Engine
    ham = hubbard_ham_set(HubbardParams(Lattice_type="Square", L1=16, L2=16, Ham_U=4.0, Beta=10.0, Dtau=0.1,
                                        Nwrap=10, Checkerboard=checkerboard))
    ham.nsigma.set_uniform_four_state(0)
    eng = Engine(ham, delay=32)
    eng.set_precision(prec)
    eng.set_fields(ham.nsigma.f.copy())
    eng.set_philox(seed=0, stream=0)
    if tau:
        eng.obs_setup()
        eng.obs_enable(True)
        eng.tau_enable(True)
    eng.init()
    ms = eng.sweep_timed(steps)
    eng.close()
    print(f"single chain: {steps} sweeps, {ms / steps:.2f} ms per sweep", flush=True)
    if batch > 0:
        b = BatchEngine([ham] * batch, delay=32)
        b.set_precision(prec)
        for z in range(batch):
            ham.nsigma.set_uniform_four_state(z)
            b.chain(z).set_fields(ham.nsigma.f.copy())
            b.chain(z).set_philox(seed=z, stream=0)
        if tau:
            b.obs_setup()
            b.obs_enable(True)
            b.tau_enable(True)
        b.init()
        ms = b.sweep_timed(steps)
        b.close()
        print(f"batch of {batch}: {steps} sweeps, {ms / steps:.2f} ms per batched sweep", flush=True)


def counter_run(batch: int, checkerboard: bool) -> None:
    """Init and one eagerly launched sweep of one chain, then of `batch` chains, with
    progress lines on stderr around every step."""
    import time

    from alf_amd.host.hamiltonian import HubbardParams, hubbard_ham_set
    from alf_amd.runtime import BatchEngine, Engine
    t0 = time.perf_counter()

    def say(msg):
        print(f"[counters {time.perf_counter() - t0:7.2f} s] {msg}", file=sys.stderr, flush=True)

    ham = hubbard_ham_set(HubbardParams(Lattice_type="Square", L1=16, L2=16, Ham_U=4.0, Beta=10.0, Dtau=0.1,
                                        Nwrap=10, Checkerboard=checkerboard))
    ham.nsigma.set_uniform_four_state(0)
    say("single chain: create")
    eng = Engine(ham, delay=32)
    eng.sweep_graphs(False)
    eng.set_fields(ham.nsigma.f.copy())
    eng.set_philox(seed=0, stream=0)
    say("single chain: init")
    eng.init()
    eng.sync()
    say("single chain: sweep")
    eng.sweep(1)
    eng.sync()
    c = eng.control()
    say(f"single chain: done, {c['ACC_up']} accepted of {c['NC_up']} visits")
    eng.close()
    if batch > 0:
        say(f"batch of {batch}: create")
        b = BatchEngine([ham] * batch, delay=32)
        b.sweep_graphs(False)
        for z in range(batch):
            ham.nsigma.set_uniform_four_state(z)
            b.chain(z).set_fields(ham.nsigma.f.copy())
            b.chain(z).set_philox(seed=z, stream=0)
        say(f"batch of {batch}: init")
        b.init()
        b.sync()
        say(f"batch of {batch}: sweep")
        b.sweep(1)
        b.sync()
        say(f"batch of {batch}: done")
        b.close()


def _counter_rows(path: str):
    import csv
    with open(path) as f:
        for row in csv.DictReader(f):
            yield row["Kernel_Name"], row["Counter_Name"], float(row["Counter_Value"])


def traffic_summary(fetch_csv: str, write_csv, copy_fetch_csv: str, copy_write_csv, copy_mib: int) -> str:
    """Per-kernel HBM bytes per dispatch from two counter passes (FETCH_SIZE, WRITE_SIZE, in KiB
    per dispatch), each scaled by what the same counter reads for a device-to-device copy of
    copy_mib MiB (which reads and writes exactly that many bytes).  write_csv / copy_write_csv
    may be "-" when only the FETCH_SIZE pass exists."""
    if write_csv == "-" or copy_write_csv == "-":
        write_csv = copy_write_csv = None
    def calib(path, counter):
        vals = [v for k, c, v in _counter_rows(path) if c == counter and "copyBuffer" in k]
        if not vals:
            raise RuntimeError(f"no copy dispatch in {path}")
        return (copy_mib << 20) / (max(vals) * 1024.0)     # bytes per counted KiB
    f_scale = calib(copy_fetch_csv, "FETCH_SIZE")
    w_scale = calib(copy_write_csv, "WRITE_SIZE") if write_csv else float("nan")
    agg = {}
    passes = [(fetch_csv, "FETCH_SIZE", f_scale)] + ([(write_csv, "WRITE_SIZE", w_scale)] if write_csv else [])
    for path, counter, scale in passes:
        for k, c, v in _counter_rows(path):
            if c != counter:
                continue
            name = k.replace("void ", "").replace("(anonymous namespace)::", "").split("(")[0].strip()
            a = agg.setdefault(name, {"FETCH_SIZE": [], "WRITE_SIZE": []})
            a[counter].append(v * 1024.0 * scale)
    lines = [f"copy calibration: FETCH_SIZE x {f_scale:.4f}, WRITE_SIZE x {w_scale:.4f} "
             f"(bytes per counted KiB, from a {copy_mib} MiB device-to-device copy)", "",
             "| Kernel | Dispatches | Read bytes / dispatch | Written bytes / dispatch |", "|---|---|---|---|"]
    for name, a in sorted(agg.items(), key=lambda kv: -sum(kv[1]["FETCH_SIZE"])):
        nf, nw = len(a["FETCH_SIZE"]), len(a["WRITE_SIZE"])
        rf = sum(a["FETCH_SIZE"]) / max(nf, 1)
        rw = sum(a["WRITE_SIZE"]) / nw if nw else float("nan")
        lines.append(f"| `{name}` | {max(nf, nw)} | {rf:.4g} | {rw:.4g} |")
        if name.startswith("chain_block_kernel"):
            lines.insert(0, f"traffic_bytes_per_launch: {rf + (rw if nw else 0.0):.6g}")
            lines.insert(1, f"traffic_counters: {'FETCH_SIZE + WRITE_SIZE' if nw else 'FETCH_SIZE only'}")
    return "\n".join(lines)


def mfma_summary(db_path: str, clock_ghz: float = 2.4) -> str:
    """Per-kernel matrix-core use from one rocprofv3 counter pass of SQ_VALU_MFMA_BUSY_CYCLES
    (summed over the SIMDs) and SQ_INSTS_VALU_MFMA_MOPS_F64 (FP64 MFMA operations / 512), read
    from the pass's database: the FP64 MFMA rate = 512 x MOPS over the dispatch's duration, and
    the MFMA-busy fraction of the SIMDs the dispatch can occupy (4 per CU, one CU per workgroup
    up to 256) at clock_ghz.  (GRBM_GUI_ACTIVE of a dispatch counts the whole collection window
    of a serialised dispatch, so the duration is the denominator.)"""
    import collections
    import sqlite3
    con = sqlite3.connect(db_path)
    rows = con.execute("select dispatch_id, kernel_name, grid_size, workgroup_size, counter_name, value, duration "
                       "from counters_collection").fetchall()
    con.close()
    disp = collections.defaultdict(dict)
    for did, name, gs, ws, cname, val, dur in rows:
        d = disp[did]
        d["name"] = name.replace("void ", "").replace("(anonymous namespace)::", "").split("(")[0].strip()
        d["wg"] = gs // max(ws, 1)
        d["ns"] = dur
        d[cname] = val
    agg = collections.defaultdict(lambda: [0, 0.0, 0.0, 0.0, 0.0])
    for d in disp.values():
        a = agg[d["name"]]
        a[0] += 1
        a[1] += d.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0)
        a[2] += 512.0 * d.get("SQ_INSTS_VALU_MFMA_MOPS_F64", 0.0)
        a[3] += d["ns"]
        a[4] += d["ns"] * min(d["wg"], 256) * 4 * clock_ghz
    total = sum(a[3] for a in agg.values()) or 1.0
    lines = ["| Kernel | Dispatches | Average us | FP64 MFMA TF/s | MFMA busy (occupied SIMDs) | Time share |",
             "|---|---|---|---|---|---|"]
    for name, a in sorted(agg.items(), key=lambda kv: -kv[1][3])[:16]:
        lines.append(f"| `{name}` | {a[0]} | {a[3] / a[0] / 1e3:.2f} | {a[2] / a[3] / 1e3:.2f} | "
                     f"{100 * a[1] / a[4]:.1f}% | {100 * a[3] / total:.1f}% |")
    return "\n".join(lines)


def shim_slices() -> None:
    """Per-slice time of alfq_wrapgrup with G on the device and with G crossing PCIe both
    ways around every call (PHASE is returned by every call in both forms)."""
    import time

    from alf_amd.host.hamiltonian import HubbardParams, hubbard_ham_set
    from alf_amd.runtime import Engine
    ham = hubbard_ham_set(HubbardParams(Lattice_type="Square", L1=16, L2=16, Ham_U=4.0, Beta=10.0, Dtau=0.1,
                                        Nwrap=10))
    ham.nsigma.set_uniform_four_state(0)
    eng = Engine(ham, delay=32)
    eng.set_fields(ham.nsigma.f.copy())
    eng.set_philox(seed=0, stream=0)
    eng.init()
    eng.sync()
    G, ph = eng.get_green(), eng.get_phase()
    L = ham.Ltrot
    t = time.perf_counter()
    for nt in range(L):
        ph = eng.wrapgrup(nt, ph)
    eng.sync()
    dev = (time.perf_counter() - t) / L
    eng.set_green(G)
    t = time.perf_counter()
    for nt in range(L):
        eng.set_green(G)
        ph = eng.wrapgrup(nt, ph)
        G = eng.get_green()
    eng.sync()
    host = (time.perf_counter() - t) / L
    eng.close()
    mb = 2 * 16 * ham.Ndim ** 2 / 1e6
    print(f"WRAPGRUP per slice, {L} slices: G on the device {dev * 1e6:.1f} us; G to and from the host "
          f"{host * 1e6:.1f} us ({mb:.2f} MB per call, {mb / 1e3 / max(host - dev, 1e-9):.1f} GB/s over the "
          f"difference)", flush=True)


if __name__ == "__main__":
    if len(sys.argv) >= 2 and sys.argv[1] == "copy":
        mib = int(sys.argv[sys.argv.index("--mib") + 1]) if "--mib" in sys.argv else 256
        copy_calibration(mib << 20)
    elif len(sys.argv) >= 2 and sys.argv[1] == "sweep":
        arg = lambda k, d: int(sys.argv[sys.argv.index(k) + 1]) if k in sys.argv else d
        profile_sweeps("--checkerboard" in sys.argv, arg("--batch", 256), arg("--steps", 3), "--tau" in sys.argv,
                       arg("--prec", 0))
    elif len(sys.argv) >= 2 and sys.argv[1] == "counters":
        arg = lambda k, d: int(sys.argv[sys.argv.index(k) + 1]) if k in sys.argv else d
        counter_run(arg("--batch", 64), "--checkerboard" in sys.argv)
    elif len(sys.argv) >= 2 and sys.argv[1] == "traffic":
        # traffic FETCH.csv WRITE.csv COPY_FETCH.csv COPY_WRITE.csv [--mib N]
        mib = int(sys.argv[sys.argv.index("--mib") + 1]) if "--mib" in sys.argv else 256
        print(traffic_summary(*sys.argv[2:6], copy_mib=mib))
    elif len(sys.argv) >= 2 and sys.argv[1] == "mfma":
        # mfma RESULTS.db [--clock-ghz F]
        print(mfma_summary(sys.argv[2], float(sys.argv[sys.argv.index("--clock-ghz") + 1]) if "--clock-ghz" in sys.argv
                           else 2.4))
    elif len(sys.argv) >= 2 and sys.argv[1] == "shim":
        shim_slices()
    else:
        print(__doc__)
