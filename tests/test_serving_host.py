"""Host logic of the continuous-batching serving loop (lanes per shape, slots refilled at sweep
boundaries, admission by memory, overlap of lanes, latency accounting) with a stand-in backend
that sleeps per sweep: no device is touched."""
import threading
import time

import numpy as np
import pytest

from alf_amd.serving import Request, Server, latency_summary, poisson_requests

SWEEP_S = 0.02      # one sweep of the stand-in backend


class FakeLane:
    def __init__(self, reqs, nslot):
        self.slots = [reqs[min(z, len(reqs) - 1)] for z in range(nslot)]
        self.sweeps = 0
        self.loads = []
        self.closed = False

    def load_chain(self, z, src):
        self.slots[z] = src.r
        self.loads.append((self.sweeps, z, src.r.rid))

    def close(self):
        self.closed = True


class FakeSrc:
    def __init__(self, r):
        self.r = r

    def close(self):
        pass


class FakeBackend:
    def __init__(self, bytes_per_chain, fail_rid=None):
        self.bpc = bytes_per_chain
        self.fail_rid = fail_rid
        self.lanes = []
        self.lock = threading.Lock()

    def chain_bytes(self, r):
        return self.bpc[r.L]

    def lane(self, reqs, nslot):
        e = FakeLane(list(reqs), nslot)
        with self.lock:
            self.lanes.append(e)
        if self.fail_rid is not None and any(r.rid == self.fail_rid for r in reqs):
            raise RuntimeError("engine failure")
        return e

    def prepare(self, r):
        return FakeSrc(r)

    def sweep(self, eng):
        time.sleep(SWEEP_S)
        eng.sweeps += 1

    def collect(self, eng, z):
        r = eng.slots[z]
        return np.full((2, 2), r.rid), np.eye(2), 1 + 0j, {"XMAXG": 0.0}


def run(reqs, budget, bpc=None, max_batch=64, fail_rid=None, min_slots=1):
    be = FakeBackend(bpc or {4: 10, 6: 20}, fail_rid)
    srv = Server(memory_budget=budget, max_batch=max_batch, backend=be, min_slots=min_slots)
    return srv, be, srv.run(reqs)


def test_same_shape_requests_share_a_lane():
    reqs = [Request(i, 4, 1.0, 4.0 + i, seed=i, nsweeps=3, arrival=0.0) for i in range(3)]
    srv, be, res = run(reqs, budget=1000)
    assert len(be.lanes) == 1 and all(r.batch_size == 3 for r in res)
    assert all(e.closed for e in be.lanes)
    assert [r.rid for r in res] == [0, 1, 2]
    assert all(int(r.fields[0, 0]) == r.rid for r in res)
    assert all(3 * SWEEP_S <= r.latency < 3 * SWEEP_S + 0.5 for r in res)


def test_late_request_joins_the_running_lane_at_a_free_slot():
    # a lane of two slots and no memory for a second lane: request 1 finishes after 2 sweeps,
    # request 2 (arriving meanwhile) takes its slot at the next sweep boundary while request 0
    # keeps sweeping
    a = Request(0, 4, 1.0, 4.0, seed=0, nsweeps=12, arrival=0.0)
    b = Request(1, 4, 1.0, 3.0, seed=1, nsweeps=2, arrival=0.0)
    c = Request(2, 4, 1.0, 2.0, seed=2, nsweeps=3, arrival=0.01)
    srv, be, res = run([a, b, c], budget=20)
    assert len(be.lanes) == 1 and srv.joined == 1
    lane = be.lanes[0]
    assert lane.loads and lane.loads[0][1:] == (1, 2)       # slot 1, request 2
    assert lane.loads[0][0] == 2                            # after request 1's two sweeps
    ra, rb, rc = res
    assert int(rc.fields[0, 0]) == 2
    assert rc.admitted >= rb.done - 1e-3 and rc.done < ra.done


def test_late_request_opens_a_second_lane_when_its_lane_is_full():
    a = Request(0, 4, 1.0, 4.0, seed=0, nsweeps=12, arrival=0.0)
    b = Request(1, 4, 1.0, 3.0, seed=1, nsweeps=12, arrival=0.0)
    c = Request(2, 4, 1.0, 2.0, seed=2, nsweeps=3, arrival=0.01)
    srv, be, res = run([a, b, c], budget=1000)
    assert len(be.lanes) == 2 and srv.joined == 0 and srv.peak_batches == 2
    assert res[2].done < res[0].done                 # c did not wait for a slot of the first lane


def test_memory_budget_serialises_lanes_in_arrival_order():
    a = Request(0, 4, 1.0, 4.0, seed=0, nsweeps=2, arrival=0.0)
    b = Request(1, 6, 1.0, 4.0, seed=1, nsweeps=2, arrival=0.0)
    srv, be, res = run([a, b], budget=25)           # 10 + 20 B do not fit together
    ra, rb = res
    assert srv.peak_batches == 1
    assert rb.admitted >= ra.done - 1e-3
    assert rb.latency > ra.latency


def test_lanes_run_concurrently_and_short_ones_finish_first():
    a = Request(0, 4, 1.0, 4.0, seed=0, nsweeps=10, arrival=0.0)
    b = Request(1, 6, 1.0, 4.0, seed=1, nsweeps=2, arrival=0.05)
    srv, be, res = run([a, b], budget=1000)
    ra, rb = res
    assert srv.peak_batches == 2
    assert 0.05 <= rb.admitted < ra.done              # admitted while a runs
    assert rb.done < ra.done                          # and finished before it
    assert ra.done < 10 * SWEEP_S + 10 * SWEEP_S      # the two overlapped


def test_slot_cap_min_slots_oversize_request_and_failures():
    reqs = [Request(i, 4, 1.0, 4.0, seed=i, nsweeps=1, arrival=0.0) for i in range(5)]
    srv, be, res = run(reqs, budget=1000, max_batch=2)
    assert len(be.lanes) == 1 and be.lanes[0].slots and len(be.lanes[0].slots) == 2
    assert [r.rid for r in res] == [0, 1, 2, 3, 4] and srv.joined == 3
    # a lone request still gets min_slots slots (room for arrivals of its shape)
    srv, be, res = run(reqs[:1], budget=1000, min_slots=3)
    assert len(be.lanes[0].slots) == 3 and res[0].batch_size == 3
    with pytest.raises(RuntimeError, match="above the budget"):
        run([Request(9, 6, 1.0, 4.0, seed=0, nsweeps=1)], budget=10)
    with pytest.raises(RuntimeError, match="engine failure"):
        run(reqs[:2], budget=1000, fail_rid=1)


def test_poisson_stream_and_summary():
    reqs = poisson_requests(40, rate=50.0, seed=3, nsweeps=2)
    gaps = np.diff([r.arrival for r in reqs])
    assert np.all(gaps > 0) and 0.005 < gaps.mean() < 0.05
    assert {r.L for r in reqs} <= {8, 12, 16, 24, 32} and {r.ltrot for r in reqs} <= {40, 80, 160}
    srv, be, res = run(reqs[:6], budget=10 ** 9, bpc={L: 1 for L in (8, 12, 16, 24, 32)})
    s = latency_summary(res, srv.wall)
    assert s["requests"] == 6 and s["p50_s"] <= s["p99_s"] <= s["max_s"]
    assert abs(s["aggregate_sweeps_per_s"] - 12 / srv.wall) < 1e-9
    assert all(r.latency >= 2 * SWEEP_S for r in res)


def test_lane_progress_notes():
    # a lane reports every note_every sweeps from its own thread, with the requests it holds
    notes = []
    be = FakeBackend({4: 10, 6: 20})
    srv = Server(memory_budget=1000, backend=be, lane_note=lambda ln, n, rids: notes.append((ln.key[1], n, rids)),
                 note_every=2)
    srv.run([Request(0, 4, 1.0, 4.0, seed=0, nsweeps=5, arrival=0.0),
             Request(1, 4, 1.0, 3.0, seed=1, nsweeps=3, arrival=0.0)])
    assert [n for _, n, _ in notes] == [2, 4]
    assert notes[0] == (4, 2, [0, 1]) and notes[1] == (4, 4, [0])
