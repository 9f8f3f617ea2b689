"""Serving loop: DQMC simulations requested over time, run as batched chains on one GPU.

ALF itself runs one parameter set per job and one Markov chain per MPI rank (no
counterpart in the reference).  Here a request asks for ``nsweeps`` sweeps of a Hubbard
model (lattice, linear size, beta, U, seed) and the loop serves a stream of them:

* Requests of the same shape (lattice, size, Ltrot, Nwrap, real or complex hopping) can
  share a batched context (one ``BatchEngine``: every kernel launch covers all its
  chains); couplings and seeds differ per chain.
* A queued request is admitted as soon as the device memory its chains need
  (``alfq_chain_bytes`` per chain) fits in the budget beside the batches already running.
  The queue is served in arrival order; the head request takes every queued request of its
  shape (up to ``max_batch``) into one new batch.  Requests that arrive while others run
  therefore start at once in a batch of their own, and their sweeps overlap on the device
  with the running batches: each batch is driven from a host thread of its own (model
  set-up, init, sweeps, download) on its own HIP stream, so batches progress independently.
* A batch whose sweeps are done returns each chain's fields, G, phase and precision
  statistics to the host and frees its memory.

Time-to-result of a request = its arrival -> its result on the host.  A chain's trajectory
does not depend on the batch it runs in (chains are independent and seeded per request), so
a request's result is the same alone and in any mix.
"""
from __future__ import annotations

import ctypes as C
import queue as _queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    L: int                     # linear size (L1 = L2)
    beta: float
    U: float
    seed: int
    nsweeps: int
    arrival: float = 0.0       # seconds after the start of the run
    lattice: str = "Square"
    nwrap: int = 10
    dtau: float = 0.1
    n_phi: int = 0

    @property
    def ltrot(self) -> int:
        return int(round(self.beta / self.dtau))

    def shape(self):
        return (self.lattice, self.L, self.ltrot, self.nwrap, self.n_phi != 0)


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    done: float
    nsweeps: int
    batch_size: int
    fields: np.ndarray = field(repr=False)
    green: np.ndarray = field(repr=False)
    phase: complex = 0j
    control: dict = field(default_factory=dict, repr=False)

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def device_free_bytes() -> int:
    """Free device memory of the current HIP device."""
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    if hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return int(free.value)


GRAPH_MIN_SWEEPS = 32        # requests with fewer sweeps run them eagerly


class _HubbardBackend:
    """Builds models and batched engines of the gfx950 sweep engine."""

    def __init__(self, delay: int = 32):
        from alf_amd.host.hamiltonian import HubbardParams, hubbard_ham_set
        from alf_amd import runtime
        self._params, self._ham_set, self.rt = HubbardParams, hubbard_ham_set, runtime
        self.delay = delay
        self._models = {}
        self._lock = threading.Lock()       # models are shared by the batches' threads

    def model(self, r: Request):
        key = (r.lattice, r.L, r.beta, r.U, r.nwrap, r.dtau, r.n_phi)
        with self._lock:
            m = self._models.get(key)
        if m is None:
            # built outside the lock, so that the batches' threads set up their models in parallel
            m = self._ham_set(self._params(Lattice_type=r.lattice, L1=r.L, L2=r.L, Beta=r.beta, Ham_U=r.U,
                                           Nwrap=r.nwrap, Dtau=r.dtau, N_Phi=r.n_phi))
            with self._lock:
                m = self._models.setdefault(key, m)
        return m

    def chain_bytes(self, r: Request) -> int:
        return self.rt.chain_bytes(self.model(r))

    def start(self, reqs):
        hams = [self.model(r) for r in reqs]
        eng = self.rt.BatchEngine(hams, delay=self.delay)
        for z, (r, h) in enumerate(zip(reqs, hams)):
            with self._lock:
                h.nsigma.set_uniform_four_state(r.seed)
                f = h.nsigma.f.copy()
            eng.chain(z).set_fields(f)
            eng.chain(z).set_philox(seed=r.seed, stream=0)
        eng.init()
        return eng

    def run(self, eng, nsweeps: int):
        # capturing a whole-sweep graph (two parities) costs more than a few eager sweeps
        # save, and captures of concurrent batches take turns
        if nsweeps < GRAPH_MIN_SWEEPS:
            eng.sweep_graphs(False)
        eng.sweep(nsweeps)
        eng.sync()

    def collect(self, eng, z: int):
        eng.chain(z)
        return eng.get_fields(), eng.get_green(), eng.get_phase(), eng.control()


@dataclass
class _Batch:
    reqs: list
    nbytes: int
    admitted: float


class Server:
    """Continuous-batching loop over a list of timed requests (see the module docstring)."""

    def __init__(self, memory_budget: int | None = None, max_batch: int = 64, backend=None, clock=time.perf_counter):
        self.backend = backend if backend is not None else _HubbardBackend()
        self.budget = memory_budget if memory_budget is not None else int(0.8 * device_free_bytes())
        self.max_batch = max_batch
        self.clock = clock
        self.peak_batches = 0
        self.wall = 0.0

    def _work(self, b: _Batch, t0: float, done):
        """One batch from admission to results: models, context, init, sweeps, download."""
        try:
            eng = self.backend.start(b.reqs)
            try:
                self.backend.run(eng, b.reqs[0].nsweeps)
                out = []
                for z, r in enumerate(b.reqs):
                    f, g, ph, ctl = self.backend.collect(eng, z)
                    out.append(Result(r.rid, r.arrival, b.admitted, self.clock() - t0, r.nsweeps, len(b.reqs),
                                      f, g, ph, ctl))
            finally:
                eng.close()
            done.put((b, out, None))
        except Exception as exc:          # reported by the loop
            done.put((b, None, exc))

    def run(self, requests) -> list:
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        queue, results, workers = [], [], {}
        done = _queue.Queue()
        used = 0
        per_chain = {}
        t0 = self.clock()
        while pending or queue or workers:
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.pop(0))
            # admission in arrival order; the head request takes the queued requests of its shape
            while queue:
                head = queue[0]
                same = [r for r in queue if r.shape() == head.shape() and r.nsweeps == head.nsweeps]
                same = same[:self.max_batch]
                key = head.shape()
                if key not in per_chain:
                    per_chain[key] = self.backend.chain_bytes(head)
                need = per_chain[key] * len(same)
                if used + need > self.budget:
                    if workers:
                        break               # wait for memory to free up
                    if per_chain[key] > self.budget:
                        raise RuntimeError(f"request {head.rid} needs {per_chain[key]} B, above the budget")
                    same = same[:max(1, self.budget // per_chain[key])]
                    need = per_chain[key] * len(same)
                b = _Batch(same, need, self.clock() - t0)
                used += need
                th = threading.Thread(target=self._work, args=(b, t0, done), daemon=True)
                workers[id(b)] = th
                th.start()
                ids = {r.rid for r in same}
                queue = [r for r in queue if r.rid not in ids]
            self.peak_batches = max(self.peak_batches, len(workers))
            # wait for a batch to finish or for the next arrival
            timeout = None
            if pending:
                timeout = max(0.0, pending[0].arrival - (self.clock() - t0))
            elif not workers:
                continue
            try:
                item = done.get(timeout=timeout)
            except _queue.Empty:
                continue
            while True:
                b, out, exc = item
                workers.pop(id(b)).join()
                used -= b.nbytes
                if exc is not None:
                    for th in workers.values():
                        th.join()
                    raise exc
                results.extend(out)
                try:
                    item = done.get_nowait()
                except _queue.Empty:
                    break
        self.wall = self.clock() - t0
        return sorted(results, key=lambda x: x.rid)


def poisson_requests(n: int, rate: float, seed: int = 0, nsweeps: int = 500, sizes=(8, 12, 16, 24, 32),
                     weights=(4, 3, 2, 1, 1), betas=(4.0, 8.0, 16.0), us=(2.0, 4.0, 6.0)):
    """A stream of Hubbard requests: Poisson arrivals at `rate` per second, square lattices of
    linear size drawn from `sizes` with `weights`, beta and U uniform over their sets."""
    rng = np.random.default_rng(seed)
    p = np.asarray(weights, float) / np.sum(weights)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        out.append(Request(rid=i, L=int(rng.choice(sizes, p=p)), beta=float(rng.choice(betas)),
                           U=float(rng.choice(us)), seed=1000 + i, nsweeps=nsweeps, arrival=t))
    return out


def latency_summary(results, wall: float) -> dict:
    lat = np.array([r.latency for r in results])
    sweeps = sum(r.nsweeps for r in results)
    return {"requests": len(results), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "max_s": float(lat.max()), "aggregate_sweeps_per_s": sweeps / wall, "wall_s": wall}
