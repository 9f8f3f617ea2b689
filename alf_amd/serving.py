"""Serving loop: DQMC simulations requested over time, run as batched chains on one GPU.

ALF itself runs one parameter set per job and one Markov chain per MPI rank (no
counterpart in the reference).  Here a request asks for ``nsweeps`` sweeps of a Hubbard
model (lattice, linear size, beta, U, seed) and the loop serves a stream of them with
continuous batching:

* Requests of one shape (lattice, size, Ltrot, Nwrap, real or complex hopping) run in one
  lane: a batched context (``BatchEngine``: every kernel launch covers all its slots) driven
  by a host thread of its own on its own HIP stream, sweeping one sweep at a time; couplings
  and seeds differ per slot.  Lanes of different shapes overlap on the device.
* A request joins a running lane of its shape that has a slot free at its next sweep
  boundary: it is initialised in a small context of its own (model set-up, fields, Philox
  stream, the UDV stack and G at tau = 0) and moved into the free slot (``alfq_chain_load``);
  the other slots keep sweeping.  A slot whose request has its sweeps returns the chain's
  fields, G, phase and precision statistics to the host and becomes free.  Free slots keep a
  copy of a chain and sweep it unread until a request takes them.
* A request without such a slot opens a new lane as soon as the device memory of its slots
  (``alfq_chain_bytes`` per slot) fits in the budget beside the lanes already running; the
  slot count is the number of requests of its shape waiting then (at least ``min_slots``, at
  most ``max_batch``).  Without the memory it waits behind a running lane of its shape (and
  joins it when a slot frees) or, with no such lane, for memory.  A lane with no request in
  its slots and none waiting closes and frees its memory.

Time-to-result of a request = its arrival -> its result on the host.  A chain's trajectory
does not depend on the lane, the slot or the other chains (chains are independent, seeded
per request, and move between contexts with their whole state), so a request's result is the
same alone and in any mix.
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

    def prepare(self, r: Request):
        """A single-chain context of the request, initialised (G and the UDV stack at tau = 0)."""
        h = self.model(r)
        eng = self.rt.Engine(h, delay=self.delay)
        with self._lock:
            h.nsigma.set_uniform_four_state(r.seed)
            f = h.nsigma.f.copy()
        eng.set_fields(f)
        eng.set_philox(seed=r.seed, stream=0)
        eng.init()
        return eng

    def lane(self, reqs, nslot: int):
        """A batched context of nslot slots; the requests take the first slots, the rest hold
        copies of the first request's chain."""
        fill = [reqs[min(z, len(reqs) - 1)] for z in range(nslot)]
        hams = [self.model(r) for r in fill]
        eng = self.rt.BatchEngine(hams, delay=self.delay)
        # two launches per delayed-update round in the lanes: a stream of 500-sweep requests
        # over nine concurrent lanes stalled twice with the one-launch rounds (cause not found)
        eng.round_fusion(False)
        for z, r in enumerate(fill):
            with self._lock:
                hams[z].nsigma.set_uniform_four_state(r.seed)
                f = hams[z].nsigma.f.copy()
            eng.chain(z).set_fields(f)
            eng.chain(z).set_philox(seed=r.seed, stream=0)
        eng.init()
        return eng

    def sweep(self, eng):
        eng.sweep(1)
        eng.sync()

    def collect(self, eng, z: int):
        eng.chain(z)
        return eng.get_fields(), eng.get_green(), eng.get_phase(), eng.control()


class _Lane:
    """One shape's batched context and the host thread that drives it."""

    def __init__(self, server, key, reqs, nslot, nbytes):
        self.server, self.key, self.nslot, self.nbytes = server, key, nslot, nbytes
        self.waiting = list(reqs)           # requests of this shape not yet in a slot
        self.busy = min(len(reqs), nslot)   # occupied slots
        self.cv = threading.Condition()
        self.closed = False
        self.thread = threading.Thread(target=self._run, daemon=True)

    def offer(self, r, wait: bool = False) -> bool:
        """Queue a request for this lane when a slot is free for it at the next sweep boundary
        (or, with wait, behind the lane's other waiting requests); False when the lane is full
        or has closed."""
        with self.cv:
            if self.closed or (not wait and self.busy + len(self.waiting) >= self.nslot):
                return False
            self.waiting.append(r)
            return True

    def _run(self):
        srv, be = self.server, self.server.backend
        try:
            with self.cv:
                first = self.waiting[:self.nslot]
                del self.waiting[:len(first)]
            eng = be.lane(first, self.nslot)
            t_adm = srv.clock() - srv.t0
            slots = {z: [r, r.nsweeps, t_adm] for z, r in enumerate(first)}
            swept = 0
            try:
                while True:
                    if slots:
                        be.sweep(eng)
                        swept += 1
                        if srv.lane_note and srv.note_every and swept % srv.note_every == 0:
                            srv.lane_note(self, swept, sorted(s[0].rid for s in slots.values()))
                    done = []
                    for z, s in list(slots.items()):
                        s[1] -= 1
                        if s[1] <= 0:
                            f, g, ph, ctl = be.collect(eng, z)
                            r = s[0]
                            done.append(Result(r.rid, r.arrival, s[2], srv.clock() - srv.t0, r.nsweeps,
                                               self.nslot, f, g, ph, ctl))
                            del slots[z]
                    if done:
                        with self.cv:
                            self.busy = len(slots)
                        srv.done.put(("results", self, done))
                    # refill free slots at this sweep boundary
                    while True:
                        with self.cv:
                            free = [z for z in range(self.nslot) if z not in slots]
                            if not free or not self.waiting:
                                if not slots and not self.waiting:
                                    self.closed = True
                                break
                            r = self.waiting.pop(0)
                        src = be.prepare(r)
                        try:
                            eng.load_chain(free[0], src)
                        finally:
                            src.close()
                        slots[free[0]] = [r, r.nsweeps, srv.clock() - srv.t0]
                        with self.cv:
                            self.busy = len(slots)
                    if self.closed:
                        break
                    srv.lane_peak = max(srv.lane_peak, len(slots))
            finally:
                eng.close()
            srv.done.put(("closed", self, None))
        except Exception as exc:          # reported by the loop
            with self.cv:
                self.closed = True
            srv.done.put(("error", self, exc))


def n_lanes(lanes) -> int:
    return sum(len(group) for group in lanes.values())


class Server:
    """Continuous-batching loop over a list of timed requests (see the module docstring)."""

    def __init__(self, memory_budget: int | None = None, max_batch: int = 64, backend=None, clock=time.perf_counter,
                 min_slots: int = 2, progress=None, lane_note=None, note_every: int = 0):
        self.backend = backend if backend is not None else _HubbardBackend()
        self.budget = memory_budget if memory_budget is not None else int(0.8 * device_free_bytes())
        self.max_batch = max_batch
        self.min_slots = min_slots
        self.progress = progress           # called with each list of results as it arrives
        self.lane_note = lane_note         # called by a lane's thread every note_every sweeps
        self.note_every = note_every
        self.clock = clock
        self.peak_batches = 0              # lanes running at once
        self.lane_peak = 0                 # occupied slots of one lane at a sweep boundary
        self.joined = 0                    # requests that joined a lane already running
        self.wall = 0.0

    def run(self, requests) -> list:
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        queue, results, lanes = [], [], {}
        self.done = _queue.Queue()
        used = 0
        per_chain = {}
        self.t0 = self.clock()
        while pending or queue or n_lanes(lanes):
            now = self.clock() - self.t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.pop(0))
            # a request joins a running lane of its shape that has a slot free at its next sweep
            # boundary; the others open new lanes as memory allows, and otherwise wait behind
            # a running lane of their shape
            rest = []
            for r in queue:
                if any(ln.offer(r) for ln in lanes.get(r.shape(), [])):
                    self.joined += 1
                else:
                    rest.append(r)
            queue = rest
            while queue:
                head = queue[0]
                key = head.shape()
                same = [r for r in queue if r.shape() == key]
                if key not in per_chain:
                    per_chain[key] = self.backend.chain_bytes(head)
                nslot = min(self.max_batch, max(len(same), self.min_slots))
                if used + per_chain[key] * nslot > self.budget:
                    nslot = min(self.max_batch, max(len(same), 1))
                need = per_chain[key] * nslot
                if used + need > self.budget:
                    running = lanes.get(key, [])
                    if running and running[0].offer(head, wait=True):
                        self.joined += 1
                        queue.pop(0)
                        continue
                    if n_lanes(lanes):
                        break               # wait for memory to free up
                    if per_chain[key] > self.budget:
                        raise RuntimeError(f"request {head.rid} needs {per_chain[key]} B, above the budget")
                    nslot = max(1, self.budget // per_chain[key])
                    need = per_chain[key] * nslot
                # the lane takes every waiting request of its shape: the first nslot start at once,
                # the others join as slots free up
                ln = _Lane(self, key, same, nslot, need)
                lanes.setdefault(key, []).append(ln)
                used += need
                self.joined += max(0, len(same) - nslot)
                ids = {r.rid for r in same}
                queue = [r for r in queue if r.rid not in ids]
                ln.thread.start()
            self.peak_batches = max(self.peak_batches, n_lanes(lanes))
            timeout = None
            if pending:
                timeout = max(0.0, pending[0].arrival - (self.clock() - self.t0))
            elif not n_lanes(lanes):
                continue
            try:
                item = self.done.get(timeout=timeout)
            except _queue.Empty:
                continue
            while True:
                kind, ln, payload = item
                if kind == "results":
                    results.extend(payload)
                    if self.progress is not None:
                        self.progress(payload)
                elif kind == "closed":
                    ln.thread.join()
                    group = lanes.get(ln.key, [])
                    if ln in group:
                        group.remove(ln)
                    if not group:
                        lanes.pop(ln.key, None)
                    used -= ln.nbytes
                    with ln.cv:                 # requests offered after the lane's last check
                        queue.extend(ln.waiting)
                        ln.waiting = []
                else:
                    every = [o for group in lanes.values() for o in group]
                    for other in every:
                        with other.cv:
                            other.waiting = []
                    for other in every:
                        other.thread.join()
                    raise payload
                try:
                    item = self.done.get_nowait()
                except _queue.Empty:
                    break
        self.wall = self.clock() - self.t0
        return sorted(results, key=lambda x: x.rid)


def uneven_batches(requests, max_ratio: float | None = None, max_batch: int = 256):
    """Groups for batches of uneven Ltrot: requests of one lattice, size, Nwrap, hopping type and
    sweep count, sorted by Ltrot; with max_ratio a group is cut where its largest Ltrot would
    exceed max_ratio times its smallest (less padding, more batches)."""
    groups = {}
    for r in requests:
        groups.setdefault((r.lattice, r.L, r.nwrap, r.n_phi != 0, r.nsweeps), []).append(r)
    out = []
    for key in sorted(groups):
        rs = sorted(groups[key], key=lambda r: (r.ltrot, r.rid))
        cur = []
        for r in rs:
            if cur and (len(cur) >= max_batch or (max_ratio is not None and r.ltrot > max_ratio * cur[0].ltrot)):
                out.append(cur)
                cur = []
            cur.append(r)
        if cur:
            out.append(cur)
    return out


def run_uneven(requests, max_ratio: float | None = None, max_batch: int = 256, delay: int = 32,
               clock=time.perf_counter, backend=None):
    """Every request's chain in a batch of uneven Ltrot (uneven_batches), the batches advancing
    concurrently from one host thread each on its own stream; returns (results by rid, wall s,
    number of batches).  A chain whose Ltrot is below its batch's largest sweeps the largest
    one with identity slices beyond its own (alfq_ctx_create_batch).  A backend passed in keeps
    its host models between calls."""
    be = backend if backend is not None else _HubbardBackend(delay=delay)
    batches = uneven_batches(requests, max_ratio, max_batch)
    results, errors = [], []
    lock = threading.Lock()
    t0 = clock()

    def work(reqs):
        try:
            hams = [be.model(r) for r in reqs]
            eng = be.rt.BatchEngine(hams, delay=delay)
            try:
                for z, (r, h) in enumerate(zip(reqs, hams)):
                    with be._lock:
                        h.nsigma.set_uniform_four_state(r.seed)
                        f = h.nsigma.f.copy()
                    eng.chain(z).set_fields(f)
                    eng.chain(z).set_philox(seed=r.seed, stream=0)
                eng.init()
                t_adm = clock() - t0
                eng.sweep(reqs[0].nsweeps)
                eng.sync()
                out = []
                for z, r in enumerate(reqs):
                    f, g, ph, ctl = be.collect(eng, z)
                    out.append(Result(r.rid, r.arrival, t_adm, clock() - t0, r.nsweeps, len(reqs), f, g, ph, ctl))
            finally:
                eng.close()
            with lock:
                results.extend(out)
        except Exception as exc:          # reported after the join
            with lock:
                errors.append(exc)

    threads = [threading.Thread(target=work, args=(b,), daemon=True) for b in batches]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    if errors:
        raise errors[0]
    return sorted(results, key=lambda x: x.rid), clock() - t0, len(batches)


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


# Hardware queues of a process that runs several lanes or batches at once.  HIP maps its
# streams onto GPU_MAX_HW_QUEUES queues (4 unless set), and streams that share a queue run
# their kernels one after another; the uneven-size mix runs 8 batches on 8 streams (measured
# 224 sweeps/s with 4 queues, 265 with 8).  Read once, when the HIP runtime starts, so it has
# to be in the environment before the first GPU call of the process.
HW_QUEUES = "8"


def use_hw_queues() -> None:
    """Ask for HW_QUEUES hardware queues unless the environment already names a count;
    call before anything in the process touches the GPU."""
    import os
    os.environ.setdefault("GPU_MAX_HW_QUEUES", HW_QUEUES)


def main(argv=None) -> None:
    """A Poisson stream of serving-configuration requests through the loop, summary as JSON:

        python3 -m alf_amd.serving --requests 16 --sweeps 500 --rate 50 --seed 0
    """
    import argparse
    import json
    use_hw_queues()
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=16)
    ap.add_argument("--sweeps", type=int, default=500)
    ap.add_argument("--rate", type=float, default=50.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--min-slots", type=int, default=2)
    ap.add_argument("--note-every", type=int, default=50, help="progress line per lane every this many sweeps")
    args = ap.parse_args(argv)
    reqs = poisson_requests(args.requests, args.rate, seed=args.seed, nsweeps=args.sweeps)
    import sys

    def progress(done):
        for r in done:
            print(f"[serving] request {r.rid} done after {r.latency:.1f} s", file=sys.stderr, flush=True)
    def lane_note(lane, swept, rids):
        print(f"[serving] {srv.clock() - srv.t0:7.1f} s  lane L={lane.key[1]} Ltrot={lane.key[2]}: sweep {swept}, "
              f"requests {rids}", file=sys.stderr, flush=True)
    srv = Server(min_slots=args.min_slots, progress=progress, lane_note=lane_note, note_every=args.note_every)
    res = srv.run(reqs)
    out = latency_summary(res, srv.wall)
    out.update({"sweeps_per_request": args.sweeps, "rate_per_s": args.rate, "seed": args.seed,
                "shapes": sorted({f"L={r.L} Ltrot={r.ltrot}" for r in reqs}),
                "peak_lanes": srv.peak_batches, "joined_running_lanes": srv.joined,
                "precision_green_max": max(r.control["XMAXG"] for r in res),
                "precision_green_mean_worst": max(r.control["precision_green_mean"] for r in res),
                "latency_s_by_request": {r.rid: round(r.latency, 3) for r in res}})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

