"""Serving layer: a request loop that admits shot-modelling and RTM requests
of different model sizes by HBM budget and runs the admitted ones
concurrently on one GPU.

A request carries the keys of the reference's argfile
(BASE/fdtd3d_base_1_interface.hpp:27-49: pml_num, fm, dt, T, the divide
numbers) plus the engine's own argfile keys (rec_width, image_fp64,
source_mode, cfl_policy, smooth_xyz, storage, ring_storage), the model (grid and a seed for the
layered generator, or arrays) and the survey in metres.  Request.to_argfile
/ Request.from_argfile map it onto an argfile.json dictionary.

Admission: each request's device footprint comes from ddx_estimate_bytes
(wavefields, CPML state, images, seismograms and, for RTM, the VOWR boundary
store).  The scheduler admits queued requests in arrival order, and backfills
later ones that fit while the head waits, as long as the admitted footprints
stay within the budget; every admitted request runs on its own engine
context (its own HIP stream), so their kernels overlap on the device, and its
memory is released when it completes.  Results are identical to a request
run alone (contexts share nothing).

Failure handling: admitted requests run their shots in pieces of `chunk`
time steps (ddx_model_steps / ddx_rtm_steps).  After every piece the
wavefields are checked for infinities and NaNs (ddx_check_finite); a
request that diverged fails with RequestDiverged, and one cancelled with
Server.cancel fails with RequestCancelled (at once while queued, at the
next piece boundary while running).  Either way its slot and HBM budget
return to the server and the other requests run on unchanged.

Continuous batching (Server(batching=True)): modelling requests in FP32 or
BF16 storage join a running fused batch (ddx_batch) of requests with the
same dt, spacing and storage at the next chunk of time steps, and leave it
when their shot is done, so one kernel launch per step covers all of them.
Their results are still identical to running alone.
"""
import collections
import concurrent.futures
import dataclasses
import inspect
import threading
import time
from typing import Optional, Sequence

import numpy as np

from . import Batch, Engine, estimate_bytes, make_desc, ricker, smooth_model
from . import synth

CHUNK = 100  # time steps between a running request's cancellation and divergence checks


class RequestCancelled(Exception):
    """The request was cancelled (Server.cancel)."""


class RequestDiverged(RuntimeError):
    """The request's wavefields became infinite or NaN."""


def _run_shot_in_pieces(e, kind, ishot, cancel, chunk, rid):
    """One shot of engine e in pieces of `chunk` stages, checking for
    cancellation before and for divergence after every piece."""
    nt = e.desc.ntime
    stages = nt if kind == "model" else 2 * nt
    step = e.model_steps if kind == "model" else e.rtm_steps
    for s0 in range(0, stages, chunk):
        if cancel is not None and cancel.is_set():
            raise RequestCancelled(f"request {rid} cancelled")
        n = min(chunk, stages - s0)
        step(ishot, s0, n)
        if not e.finite():
            raise RequestDiverged(f"request {rid}: the wavefield diverged (non-finite values) in "
                                  f"{'modelling' if kind == 'model' else 'RTM'} shot {ishot} before stage {s0 + n}")


ARGFILE_KEYS = ("pml_num", "fm", "dt", "T", "x_divide_num", "y_divide_num", "z_divide_num", "rec_width",
                "image_fp64", "source_mode", "cfl_policy", "smooth_xyz", "storage", "ring_storage", "n_xyz",
                "d_xyz")


@dataclasses.dataclass
class Request:
    kind: str                                  # "model" (seismograms) or "rtm" (images)
    n_xyz: Sequence[int]                       # physical grid
    d_xyz: Sequence[float] = (10.0, 10.0, 10.0)
    pml_num: int = 20
    fm: float = 15.0
    dt: float = 0.001
    T: float = 1.0
    x_divide_num: int = 1
    y_divide_num: int = 1
    z_divide_num: int = 1
    rec_width: int = 1
    image_fp64: int = 0
    source_mode: int = 0
    cfl_policy: int = 0
    smooth_xyz: Sequence[int] = (0, 0, 0)
    storage: str = "fp32"                      # wavefield storage: "fp32", "fp16" or "bf16"
    ring_storage: str = "fp32"                 # VOWR store over FP32 wavefields: "fp16", "bf16" or "bfp16"
    model_seed: int = 0                        # layered model generator seed (when no arrays are given)
    shots: Optional[np.ndarray] = None         # [nshot][3] metres
    receivers: Optional[np.ndarray] = None     # [nrec][3] metres
    rid: str = ""

    def __post_init__(self):
        if self.kind not in ("model", "rtm"):
            raise ValueError("kind must be 'model' or 'rtm'")
        if len(self.n_xyz) != 3 or min(self.n_xyz) < 8:
            raise ValueError("n_xyz must be three sizes of at least 8")
        if self.d_xyz[0] != self.d_xyz[1] or self.d_xyz[0] != self.d_xyz[2]:
            raise ValueError("the engine takes one grid spacing (d_xyz all equal)")
        if self.storage not in ("fp32", "fp16", "bf16"):
            raise ValueError("storage must be fp32, fp16 or bf16")
        if self.ring_storage not in ("fp32", "fp16", "bf16", "bfp16") or (
                self.ring_storage != "fp32" and self.storage != "fp32"):
            raise ValueError("ring_storage must be fp16, bf16 or bfp16, over fp32 storage")

    @property
    def ntime(self):
        from . import ntime
        return ntime(self.T, self.dt)

    def to_argfile(self):
        """argfile.json dictionary of the request (ArgList keys plus the
        engine keys; model and survey paths are the caller's)."""
        a = {k: getattr(self, k) for k in ARGFILE_KEYS}
        for k in ("smooth_xyz", "n_xyz", "d_xyz"):
            a[k] = list(a[k])
        return a

    @classmethod
    def from_argfile(cls, argfile, kind, shots, receivers, model_seed=0, rid=""):
        """Request from an argfile.json dictionary (absent engine keys take
        their defaults, as the argfile reader does)."""
        kw = {k: argfile[k] for k in ARGFILE_KEYS if k in argfile}
        return cls(kind=kind, shots=np.asarray(shots, np.float32), receivers=np.asarray(receivers, np.float32),
                   model_seed=model_seed, rid=rid, **kw)

    def desc(self):
        nx, ny, nz = (int(v) for v in self.n_xyz)
        return make_desc(nx, ny, nz, float(self.d_xyz[0]), int(self.pml_num), float(self.dt), float(self.fm),
                         self.ntime, divide=(self.x_divide_num, self.y_divide_num, self.z_divide_num),
                         rec_width=self.rec_width, image_fp64=self.image_fp64, cfl_policy=self.cfl_policy,
                         storage=self.storage, ring_storage=self.ring_storage)

    def footprint(self):
        """Device bytes of the request's context (ddx_estimate_bytes)."""
        nshot = 0 if self.shots is None else len(self.shots)
        nrec = 0 if self.receivers is None else len(self.receivers)
        return estimate_bytes(self.desc(), nrec, nshot, rtm=self.kind == "rtm",
                              migration=any(int(v) > 0 for v in self.smooth_xyz))


def default_model(req):
    nx, ny, nz = (int(v) for v in req.n_xyz)
    vp, vs, rho = synth.layered(nx, ny, nz, req.model_seed, nlayer=(3, 6), vp_range=(2000.0, 3500.0),
                                anomaly_radius=max(2, min(nx, ny, nz) // 6))
    return synth.clip_for_stability(vp, vs, rho, req.dt, float(req.d_xyz[0]))[:3]


def load_engine(req, model=None):
    """Engine context of a request with its model, wavelet and survey set."""
    vp, vs, rho = model if model is not None else default_model(req)
    e = Engine(req.desc())
    try:
        e.set_model(vp, vs, rho)
        if any(int(v) > 0 for v in req.smooth_xyz):
            e.set_migration_model(*[smooth_model(a, req.smooth_xyz) for a in (vp, vs, rho)])
        e.set_wavelet(ricker(req.fm, req.dt, req.ntime))
        e.set_geometry(req.shots, req.receivers)
    except Exception:
        e.close()
        raise
    return e


class BatchLane:
    """One thread advancing a continuous batch (ddx.Batch) in chunks of
    `chunk` time steps; shots handed to run_shot() join at the next chunk
    and run_shot() returns when theirs is done."""

    def __init__(self, chunk=50):
        self.chunk = chunk
        self.batch = Batch()
        self.cv = threading.Condition()
        self.pending = collections.deque()
        self.members = {}
        self.closed = False
        self.thread = threading.Thread(target=self._loop, daemon=True)
        self.thread.start()

    def run_shot(self, engine, ishot, cancel=None, rid="", kind="model"):
        """Run shot ishot of engine in the batch, a modelling shot or (kind
        "rtm") an RTM shot whose observed data are set; returns when it is
        done.  Raises RequestCancelled when `cancel` is set before the shot
        ends, RequestDiverged when its wavefields become non-finite."""
        done = threading.Event()
        box = {"cancel": cancel, "rid": rid, "kind": kind}
        with self.cv:
            if self.closed:
                raise RuntimeError("batch lane closed")
            self.pending.append((engine, ishot, done, box))
            self.cv.notify_all()
        done.wait()
        if "error" in box:
            raise box["error"]

    def _fail_all(self, ex, pending=False):
        """Fail every member (and, with pending, every shot waiting to join)
        with ex.  The batch is closed (its stream synchronized) before any
        waiter wakes, so a woken caller may destroy its context."""
        waiters = list(self.members.values())
        self.members.clear()
        if pending:
            with self.cv:
                waiters += [(done, box) for _, _, done, box in self.pending]
                self.pending.clear()
        try:
            self.batch.close()
        finally:
            self.batch = Batch()
            for done, box in waiters:
                box["error"] = ex
                done.set()

    def _loop(self):
        try:
            self._loop_body()
        except BaseException as ex:  # never leave a caller waiting on a dead lane
            with self.cv:
                self.closed = True
            self._fail_all(RuntimeError(f"batch lane stopped: {ex!r}"), pending=True)
            raise

    def _loop_body(self):
        while True:
            with self.cv:
                while not self.pending and not self.members and not self.closed:
                    self.cv.wait()
                if self.closed and not self.pending and not self.members:
                    return
                adds = list(self.pending)
                self.pending.clear()
            for e, ishot, done, box in adds:
                try:
                    self.batch.add(e, ishot, rtm=box.get("kind") == "rtm")
                    box["engine"] = e
                    self.members[e.h.value] = (done, box)
                except Exception as ex:
                    box["error"] = ex
                    done.set()
            if not self.members:
                continue
            try:
                finished = self.batch.run(self.chunk)
            except Exception as ex:  # every member fails with the batch
                self._fail_all(ex)
                continue
            for e in finished:
                done, _ = self.members.pop(e.h.value)
                done.set()
            self._drop_failed()

    def _drop_failed(self):
        """Take cancelled and diverged members out of the batch; the others
        run on."""
        for key, (done, box) in list(self.members.items()):
            e = box.get("engine")
            cancel = box.get("cancel")
            err = None
            if cancel is not None and cancel.is_set():
                err = RequestCancelled(f"request {box.get('rid', '')} cancelled")
            elif e is not None and not e.finite():
                err = RequestDiverged(f"request {box.get('rid', '')}: the wavefield diverged (non-finite values)")
            if err is None:
                continue
            self.batch.remove(e)
            del self.members[key]
            box["error"] = err
            done.set()

    def close(self):
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        self.thread.join()
        self.batch.close()


def lane_key(req):
    """Batch lane of a request (modelling and RTM requests share lanes), or
    None when it runs on its own context: FP16 storage (per-request scales),
    FP16 or BF16 VOWR stores over FP32 wavefields, and smoothed migration
    models.  The block-scaled store ("bfp16") joins the lane of its
    wavefield storage: a batch saves and injects each member's store on its
    own."""
    if (req.storage not in ("fp32", "bf16") or req.ring_storage not in ("fp32", "bfp16") or
            any(int(v) > 0 for v in req.smooth_xyz)):
        return None
    return (float(req.dt), tuple(float(v) for v in req.d_xyz), req.storage, int(req.image_fp64))


def run_request_in_lane(req, model, lane, cancel=None):
    """A request whose shots run in a continuous batch: a modelling request
    its modelling shots; an RTM request, per shot, the modelling of its
    observed data, then the RTM shot."""
    e = load_engine(req, model)
    try:
        c0 = e.cell_updates()
        out = {}
        if req.kind == "model":
            seis = []
            for ishot in range(len(req.shots)):
                lane.run_shot(e, ishot, cancel, req.rid)
                seis.append(e.seismograms(ishot))
            out["seismograms"] = seis
        else:
            for ishot in range(len(req.shots)):
                lane.run_shot(e, ishot, cancel, req.rid)
                lane.run_shot(e, ishot, cancel, req.rid, kind="rtm")
            e.finalize_images()
            out["images"] = [[e.image(q, r) for q in range(9)] for r in range(e.nsub)]
        e.synchronize()
        out.update(cells=e.cell_updates() - c0, device_bytes=e.device_bytes())
        return out
    finally:
        e.close()


def run_request(req, model=None, graphs=False, autotune=True, cancel=None, chunk=CHUNK):
    """Run one request on its own engine context; returns a dict with the
    seismograms (kind "model") or the per-subdomain images (kind "rtm").
    Shots run in pieces of `chunk` steps with a cancellation check (the
    threading.Event `cancel`) and a divergence check between pieces.
    graphs: replay the modelling steps from a captured graph instead, whole
    shots without checks (measured no faster for these request sizes, so
    off by default)."""
    e = load_engine(req, model)
    try:
        if autotune:
            e.autotune()  # brick shapes, cached per process by size bucket
        e.set_graphs(graphs)
        out = {"cells": 0.0}
        c0 = e.cell_updates()

        def shot(kind, ishot):
            if graphs:
                (e.model_shot if kind == "model" else e.rtm_shot)(ishot)
            else:
                _run_shot_in_pieces(e, kind, ishot, cancel, chunk, req.rid)

        if req.kind == "model":
            seis = []
            for ishot in range(len(req.shots)):
                shot("model", ishot)
                seis.append(e.seismograms(ishot))
            out["seismograms"] = seis
        else:
            for ishot in range(len(req.shots)):
                shot("model", ishot)
                shot("rtm", ishot)
            e.finalize_images()
            out["images"] = [[e.image(q, r) for q in range(9)] for r in range(e.nsub)]
        e.synchronize()
        out["cells"] = e.cell_updates() - c0
        out["device_bytes"] = e.device_bytes()
        return out
    finally:
        e.close()


class Server:
    """Admits requests by HBM budget and runs the admitted ones concurrently.

    submit() returns a Future whose result is run_request's dict plus the
    request's latency (seconds from submission to completion) and queueing
    delay.  `runner` replaces run_request (tests of the admission logic).
    batching: modelling and RTM requests in FP32/BF16 join continuous
    batches (one per dt, spacing, storage and image precision), advanced
    `chunk` steps at a time."""

    def __init__(self, hbm_budget, max_inflight=16, runner=None, batching=False, chunk=50):
        self.budget = float(hbm_budget)
        self.lanes = {}
        self.chunk = chunk
        self.runner = runner or (self._run_batched if batching else run_request)
        try:  # runners that take a `cancel` event get the request's
            self._pass_cancel = "cancel" in inspect.signature(self.runner).parameters
        except (TypeError, ValueError):
            self._pass_cancel = False
        self.cancels = {}     # future -> its request's cancel event
        self.failed = []      # (rid, error name)
        self.pool = concurrent.futures.ThreadPoolExecutor(max_workers=max_inflight)
        self.max_inflight = max_inflight
        self.lock = threading.Condition()
        self.queue = collections.deque()
        self.used = 0.0
        self.inflight = 0
        self.peak_inflight = 0
        self.peak_used = 0.0
        self.done = []        # (rid, latency, queue_delay, cells)
        self.closed = False
        self.t0 = time.perf_counter()
        self.thread = threading.Thread(target=self._schedule, daemon=True)
        self.thread.start()

    def submit(self, req, model=None):
        if self.closed:
            raise RuntimeError("server closed")
        fp = req.footprint()
        if fp > self.budget:
            raise ValueError(f"request {req.rid} needs {fp / 1e9:.2f} GB, above the budget {self.budget / 1e9:.2f} GB")
        fut = concurrent.futures.Future()
        with self.lock:
            self.cancels[fut] = threading.Event()
            self.queue.append((req, model, fp, time.perf_counter(), fut))
            self.lock.notify_all()
        return fut

    def cancel(self, fut):
        """Cancel a submitted request: a queued one fails at once with
        RequestCancelled; a running one stops at its next piece boundary
        (its slot and budget are then released).  Returns False when the
        request has already finished."""
        with self.lock:
            for k, item in enumerate(self.queue):
                if item[4] is fut:
                    del self.queue[k]
                    self.cancels.pop(fut, None)
                    self.failed.append((item[0].rid, "RequestCancelled"))
                    fut.set_exception(RequestCancelled(f"request {item[0].rid} cancelled while queued"))
                    self.lock.notify_all()
                    return True
            ev = self.cancels.get(fut)
        if ev is None or fut.done():
            return False
        ev.set()
        return True

    def _schedule(self):
        # close() stops new submissions; requests already queued are still
        # admitted as memory frees up, and the thread ends once none is left
        while True:
            with self.lock:
                while not self._admit_one():
                    if self.closed and not self.queue:
                        return
                    self.lock.wait()

    def _admit_one(self):
        """Admit the first queued request that fits (arrival order, backfill)."""
        if self.inflight >= self.max_inflight:
            return False
        for k, (req, model, fp, t_sub, fut) in enumerate(self.queue):
            if self.used + fp <= self.budget:
                del self.queue[k]
                self.used += fp
                self.inflight += 1
                self.peak_inflight = max(self.peak_inflight, self.inflight)
                self.peak_used = max(self.peak_used, self.used)
                self.pool.submit(self._run, req, model, fp, t_sub, time.perf_counter(), fut)
                return True
        return False

    def _run(self, req, model, fp, t_sub, t_adm, fut):
        try:
            with self.lock:
                cancel = self.cancels.get(fut)
            if self._pass_cancel:
                out = self.runner(req, model, cancel=cancel)
            else:
                out = self.runner(req, model)
            t_end = time.perf_counter()
            out = dict(out)
            out["latency_s"] = t_end - t_sub
            out["queue_s"] = t_adm - t_sub
            with self.lock:
                self.done.append((req.rid, t_end - t_sub, t_adm - t_sub, float(out.get("cells", 0.0))))
            fut.set_result(out)
        except Exception as ex:  # the caller sees the failure on its future
            with self.lock:
                self.failed.append((req.rid, type(ex).__name__))
            fut.set_exception(ex)
        finally:
            with self.lock:
                self.cancels.pop(fut, None)
                self.used -= fp
                self.inflight -= 1
                self.lock.notify_all()

    def _run_batched(self, req, model, cancel=None):
        key = lane_key(req)
        if key is None:
            return run_request(req, model, cancel=cancel)
        with self.lock:
            lane = self.lanes.get(key)
            if lane is None:
                lane = self.lanes[key] = BatchLane(self.chunk)
        return run_request_in_lane(req, model, lane, cancel)

    def close(self):
        with self.lock:
            self.closed = True
            self.lock.notify_all()
        self.thread.join()
        self.pool.shutdown(wait=True)
        for lane in self.lanes.values():
            lane.close()

    def stats(self):
        """Requests done, requests/s, Gcell-updates/s, latency p50/p99, peaks."""
        with self.lock:
            lat = [d[1] for d in self.done]
            cells = sum(d[3] for d in self.done)
        wall = time.perf_counter() - self.t0
        return {"done": len(lat), "failed": len(self.failed), "requests_per_s": len(lat) / wall if wall > 0 else 0.0,
                "gcell_updates_per_s": cells / wall / 1e9 if wall > 0 else 0.0,
                "latency_p50_s": float(np.percentile(lat, 50)) if lat else None,
                "latency_p99_s": float(np.percentile(lat, 99)) if lat else None,
                "peak_inflight": self.peak_inflight, "peak_bytes": self.peak_used, "budget_bytes": self.budget}
