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
"""
from __future__ import annotations

HBM_PEAK_GBS = 8000.0


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
        "traffic": None,
        "kernel": "chain_block_kernel (Metropolis decisions + delayed rank-1 updates)",
        "avg_launch_us": avg_s * 1e6,
        "launches": launches,
        "visits": visits,
        "accepted": acc,
        "algorithmic_bytes_per_launch": bytes_per_launch,
    }
