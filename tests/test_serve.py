"""Serving layer: request schema <-> argfile keys, HBM admission with a
stub runner (CPU), and mixed vs alone results on the GPU."""
import threading
import types
import time

import numpy as np
import pytest

import ddx
from ddx import serve


def small_request(kind, n, rid, seed=0, T=0.03):
    nx, ny, nz = n
    d = 10.0
    shots = np.array([[nx // 2 * d, ny // 2 * d, 3 * d]], np.float32)
    g = np.arange(2, nx - 2, 4) * d
    rec = np.stack([g, np.full(g.size, (ny // 2) * d), np.full(g.size, 2 * d)], 1).astype(np.float32)
    return serve.Request(kind=kind, n_xyz=n, pml_num=6, fm=25.0, dt=0.001, T=T, rec_width=2, cfl_policy=1,
                         model_seed=seed, shots=shots, receivers=rec, rid=rid)


def test_request_argfile_round_trip():
    r = small_request("rtm", (30, 28, 24), "a")
    a = r.to_argfile()
    for k in ("pml_num", "fm", "dt", "T", "x_divide_num", "y_divide_num", "z_divide_num", "rec_width"):
        assert k in a  # the reference's ArgList keys plus the engine keys
    r2 = serve.Request.from_argfile(a, "rtm", r.shots, r.receivers, rid="a")
    assert r2.to_argfile() == a and r2.ntime == r.ntime == ddx.ntime(0.03, 0.001)
    with pytest.raises(ValueError):
        serve.Request(kind="fwd", n_xyz=(16, 16, 16))


def test_footprint_grows_with_size_and_rtm():
    small = small_request("model", (24, 24, 24), "s").footprint()
    big = small_request("model", (48, 48, 48), "b").footprint()
    rtm = small_request("rtm", (24, 24, 24), "r").footprint()
    assert 0 < small < big and rtm > small
    half = small_request("rtm", (24, 24, 24), "h")
    half.storage = "bf16"
    assert half.footprint() < rtm and half.to_argfile()["storage"] == "bf16"
    ring = small_request("rtm", (24, 24, 24), "g")
    ring.ring_storage = "fp16"
    assert half.footprint() < ring.footprint() < rtm and ring.to_argfile()["ring_storage"] == "fp16"
    bfp = small_request("rtm", (24, 24, 24), "p")
    bfp.ring_storage = "bfp16"  # 16-bit codes plus an exponent byte per 16 entries
    assert ring.footprint() < bfp.footprint() < rtm and bfp.to_argfile()["ring_storage"] == "bfp16"
    with pytest.raises(ValueError):
        serve.Request(kind="rtm", n_xyz=(16, 16, 16), storage="bf16", ring_storage="fp16")
    with pytest.raises(ValueError):
        serve.Request(kind="rtm", n_xyz=(16, 16, 16), ring_storage="fp8")


def test_admission_respects_budget_and_backfills():
    """Stub runner: the sum of admitted footprints never exceeds the budget,
    a request that does not fit waits while smaller later ones are
    admitted, and everything completes."""
    reqs = [small_request("model", n, f"r{i}") for i, n in
            enumerate([(48, 48, 48), (48, 48, 48), (24, 24, 24), (24, 24, 24), (40, 40, 40)])]
    fps = [r.footprint() for r in reqs]
    budget = fps[0] + fps[2] + fps[3] + 1.0  # one large and the two small at once, not two large
    live = {"bytes": 0.0, "max": 0.0}
    order = []
    lk = threading.Lock()

    def runner(req, model):
        fp = req.footprint()
        with lk:
            live["bytes"] += fp
            live["max"] = max(live["max"], live["bytes"])
            order.append(req.rid)
        time.sleep(0.05)
        with lk:
            live["bytes"] -= fp
        return {"cells": 1.0}

    srv = serve.Server(budget, runner=runner)
    futs = [srv.submit(r) for r in reqs]
    for f in futs:
        f.result(timeout=30)
    srv.close()
    assert live["max"] <= budget
    # r1 (large) waits for r0 while the two small ones backfill: r0, r2 and
    # r3 start together (in any order), r1 after them
    assert sorted(order[:3]) == ["r0", "r2", "r3"] and "r1" in order[3:]
    st = srv.stats()
    assert st["done"] == 5 and st["peak_bytes"] <= budget and st["peak_inflight"] >= 2


def test_close_with_requests_still_queued():
    """close() while requests wait behind the budget: the queued ones are
    still admitted and run as memory frees up, every future resolves and
    close() returns; submit() after close() is refused."""
    reqs = [small_request("model", (32, 32, 32), f"q{i}") for i in range(4)]
    fp = reqs[0].footprint()
    ran = []

    def runner(req, model):
        time.sleep(0.05)
        ran.append(req.rid)
        return {"cells": 1.0}

    srv = serve.Server(fp * 1.5, runner=runner)  # one at a time
    futs = [srv.submit(r) for r in reqs]
    closer = threading.Thread(target=srv.close)
    closer.start()
    closer.join(timeout=30)
    assert not closer.is_alive()
    assert all(f.done() for f in futs) and sorted(ran) == [r.rid for r in reqs]
    with pytest.raises(RuntimeError):
        srv.submit(reqs[0])


def test_oversized_request_is_refused():
    r = small_request("rtm", (40, 40, 40), "x")
    srv = serve.Server(r.footprint() / 2, runner=lambda q, m: {})
    with pytest.raises(ValueError):
        srv.submit(r)
    srv.close()


@pytest.mark.gpu
@pytest.mark.parametrize("batching", [False, True])
def test_mixed_requests_match_alone(batching):
    """Six modelling and RTM requests of different sizes through one server
    whose budget admits a few at a time (with continuous batching: the FP32
    and BF16 requests, modelling and RTM, share fused batches, one RTM
    request with the block-scaled store among them): each result equals the
    same request run alone."""
    shapes = [(32, 30, 26), (24, 24, 24), (40, 28, 24), (28, 36, 30), (24, 26, 28), (36, 32, 24)]
    reqs = [small_request("rtm" if i % 2 else "model", n, f"q{i}", seed=i) for i, n in enumerate(shapes)]
    reqs[3].ring_storage = "bfp16"
    reqs[4].storage = "bf16"  # the storage format is chosen per request
    reqs[5].storage = "fp16"
    budget = 2.5 * max(r.footprint() for r in reqs)
    srv = serve.Server(budget, batching=batching, chunk=17)
    futs = [srv.submit(r) for r in reqs]
    mixed = [f.result(timeout=300) for f in futs]
    if batching:
        assert sorted(srv.lanes) == sorted({serve.lane_key(r) for r in reqs} - {None})
    st = srv.stats()
    srv.close()
    print("serving stats:", st, flush=True)
    assert st["peak_inflight"] >= 2
    for r, m in zip(reqs, mixed):
        alone = serve.run_request(r)
        key = "seismograms" if r.kind == "model" else "images"
        for a, b in zip(np.asarray(alone[key], dtype=object).ravel(), np.asarray(m[key], dtype=object).ravel()):
            assert np.array_equal(a, b), r.rid
        assert abs(m["device_bytes"] - r.footprint()) <= 0.05 * r.footprint() + (1 << 20)


@pytest.mark.gpu
@pytest.mark.parametrize("batching", [False, True])
def test_diverged_and_cancelled_requests_leave_the_rest_unchanged(batching):
    """Among five requests, one diverges (a time step above the stability
    limit, accepted with cfl_policy = 1) and one long one is cancelled while
    it runs: they fail with RequestDiverged and RequestCancelled, their
    budget returns, and the three others equal their runs alone."""
    reqs = [small_request("model", (32, 30, 26), "ok0", seed=1), small_request("rtm", (24, 24, 24), "ok1", seed=2),
            small_request("model", (28, 26, 24), "ok2", seed=3)]
    bad = small_request("model", (30, 28, 26), "diverges", seed=4, T=0.6)
    bad.dt = 0.0025  # vp dt / h = 0.75 on the model below (the 8th-order limit is 0.449)
    bad.T = 0.6
    # an explicit homogeneous model: the layered default clips vp to the
    # stable limit of the request's dt (the FP32 oracle overflows at step 46)
    vp = np.full((26, 28, 30), 3000.0, np.float32)
    bad_model = (vp, (vp / np.sqrt(3)).astype(np.float32), np.full_like(vp, 2000.0))
    # about 6 s of time steps alone on an MI355X: still running at the cancel
    long = small_request("model", (40, 36, 32), "cancelled", seed=5, T=300.0)
    srv = serve.Server(1e12, batching=batching, chunk=20)
    f_bad = srv.submit(bad, bad_model)
    f_long = srv.submit(long)
    futs = [srv.submit(r) for r in reqs]
    time.sleep(1.0)
    assert srv.cancel(f_long)
    mixed = [f.result(timeout=300) for f in futs]
    with pytest.raises(serve.RequestDiverged):
        f_bad.result(timeout=300)
    with pytest.raises(serve.RequestCancelled):
        f_long.result(timeout=300)
    srv.close()
    st = srv.stats()
    print("serving stats:", st, flush=True)
    assert st["failed"] == 2 and st["done"] == 3 and srv.used == 0.0
    for r, m in zip(reqs, mixed):
        alone = serve.run_request(r)
        key = "seismograms" if r.kind == "model" else "images"
        for a, b in zip(np.asarray(alone[key], dtype=object).ravel(), np.asarray(m[key], dtype=object).ravel()):
            assert np.array_equal(a, b), r.rid


def test_lane_keys():
    """Which requests share a continuous batch (no GPU needed)."""
    a = small_request("model", (24, 24, 24), "a")
    b = small_request("model", (40, 32, 28), "b")
    assert serve.lane_key(a) == serve.lane_key(b) is not None
    b.storage = "bf16"
    assert serve.lane_key(b) != serve.lane_key(a)
    b.storage = "fp16"  # per-request scales: runs alone
    assert serve.lane_key(b) is None
    c = small_request("rtm", (24, 24, 24), "c")
    assert serve.lane_key(c) == serve.lane_key(a)  # RTM and modelling requests share a lane
    c.image_fp64 = 1
    assert serve.lane_key(c) not in (None, serve.lane_key(a))
    c.ring_storage = "fp16"  # an FP16 VOWR store over FP32 wavefields: runs alone
    assert serve.lane_key(c) is None
    c.ring_storage = "bfp16"  # the block-scaled store joins its wavefields' lane
    c.image_fp64 = 0
    assert serve.lane_key(c) == serve.lane_key(a)


class _StubBatch:
    """Stands in for ddx.Batch: a member leaves after its own number of steps."""

    def __init__(self):
        self.members = {}

    def add(self, engine, ishot=0, rtm=False):
        if getattr(engine, "bad", False):
            raise RuntimeError("cannot join")
        self.members[engine.h.value] = [engine, engine.steps]

    def run(self, max_steps):
        if not self.members:
            return []
        k = min(max_steps, min(v[1] for v in self.members.values()))
        done = []
        for key in list(self.members):
            self.members[key][1] -= k
            self.members[key][0].ran += k
            if self.members[key][1] <= 0:
                done.append(self.members.pop(key)[0])
        return done

    def remove(self, engine):
        del self.members[engine.h.value]

    def __len__(self):
        return len(self.members)

    def close(self):
        pass


class _StubEngine:
    def __init__(self, steps, bad=False, diverge_after=None):
        self.steps = steps
        self.bad = bad
        self.diverge_after = diverge_after
        self.ran = 0
        self.h = types.SimpleNamespace(value=id(self))

    def finite(self):
        return self.diverge_after is None or self.ran < self.diverge_after


def test_batch_lane_joins_and_leaves(monkeypatch):
    """Shots handed to a lane from several threads, of different lengths and
    arriving at different times, all complete; a shot that cannot join fails
    on its own caller only."""
    monkeypatch.setattr(serve, "Batch", _StubBatch)
    lane = serve.BatchLane(chunk=7)
    engines = [_StubEngine(s) for s in (30, 5, 60, 12)] + [_StubEngine(10, bad=True)]
    errors, done = [], []

    def submit(e, delay):
        time.sleep(delay)
        try:
            lane.run_shot(e, 0)
            done.append(e)
        except RuntimeError as ex:
            errors.append((e, str(ex)))

    threads = [threading.Thread(target=submit, args=(e, 0.01 * q)) for q, e in enumerate(engines)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=30)
    lane.close()
    assert sorted(map(id, done)) == sorted(map(id, engines[:4]))
    assert len(errors) == 1 and errors[0][0] is engines[4] and "cannot join" in errors[0][1]


def test_batch_lane_drops_cancelled_and_diverged_members(monkeypatch):
    """A member whose wavefield diverges is taken out of the batch with
    RequestDiverged, a cancelled one with RequestCancelled; the others
    finish."""
    monkeypatch.setattr(serve, "Batch", _StubBatch)
    lane = serve.BatchLane(chunk=5)
    ok = [_StubEngine(40), _StubEngine(23)]
    bad = _StubEngine(60, diverge_after=12)
    slow = _StubEngine(10**6)
    cancel = threading.Event()
    res = {}

    def submit(e, ev=None):
        try:
            lane.run_shot(e, 0, ev, rid=str(id(e)))
            res[id(e)] = "done"
        except Exception as ex:  # noqa: BLE001 - the kind is what is checked
            res[id(e)] = type(ex).__name__

    threads = [threading.Thread(target=submit, args=(e,)) for e in ok + [bad]]
    threads.append(threading.Thread(target=submit, args=(slow, cancel)))
    for t in threads:
        t.start()
    time.sleep(0.2)
    cancel.set()
    for t in threads:
        t.join(timeout=30)
    lane.close()
    assert [res[id(e)] for e in ok] == ["done", "done"]
    assert res[id(bad)] == "RequestDiverged" and res[id(slow)] == "RequestCancelled"


def test_cancel_queued_and_running_requests():
    """Server.cancel: a queued request fails at once, a running one at its
    next check; both release their budget and the rest complete."""
    reqs = [small_request("model", (32, 32, 32), f"c{i}") for i in range(4)]
    fp = reqs[0].footprint()
    started = threading.Event()

    def runner(req, model, cancel=None):
        if req.rid == "c0":
            started.set()
            while not cancel.is_set():
                time.sleep(0.01)
            raise serve.RequestCancelled(req.rid)
        time.sleep(0.02)
        return {"cells": 1.0}

    srv = serve.Server(fp * 1.5, runner=runner)  # one at a time
    futs = [srv.submit(r) for r in reqs]
    assert started.wait(10)
    assert srv.cancel(futs[2])  # still queued
    with pytest.raises(serve.RequestCancelled):
        futs[2].result(timeout=10)
    assert srv.cancel(futs[0])  # running
    with pytest.raises(serve.RequestCancelled):
        futs[0].result(timeout=10)
    assert futs[1].result(timeout=10)["cells"] == 1.0 and futs[3].result(timeout=10)["cells"] == 1.0
    assert not srv.cancel(futs[1])  # finished already
    srv.close()
    st = srv.stats()
    assert st["done"] == 2 and st["failed"] == 2 and srv.used == 0.0
