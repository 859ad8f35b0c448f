"""Load test of the serving layer: a Poisson stream of modelling and RTM
requests of random sizes through ddx.serve.Server, at one or more arrival
rates; prints one JSON line per rate with requests/s, Gcell-updates/s and
p50/p99 latency.

python -m ddx.serve_bench [--requests N] [--rates 5,10,20,50] [--sizes 48,128]
                          [--steps 200,500] [--bf16 0.5] [--kinds model,rtm]
                          [--graphs 0] [--batching 0] [--budget-gb G] [--seed S]

Each request: kind model or RTM with equal odds (--kinds), an isotropic physical grid
with each side uniform in --sizes, dx = 10 m, 20 CPML layers, steps uniform
in --steps, a layered model from its own seed, one surface shot and a
receiver line, FP32 or BF16 storage (BF16 with probability --bf16).  Models
are built before the clock starts; a request's latency runs from its arrival
to its result.
"""
import argparse
import itertools
import json
import math
import sys
import time

import numpy as np

from . import serve


def make_requests(n, sizes, steps, p_bf16, seed, kinds=("model", "rtm")):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        s = int(rng.integers(sizes[0], sizes[1] + 1))
        nt = int(rng.integers(steps[0], steps[1] + 1))
        kind = kinds[int(rng.integers(0, len(kinds)))]
        st = "bf16" if rng.random() < p_bf16 else "fp32"
        dx = 10.0
        shots = np.array([[(s // 2) * dx, (s // 2) * dx, 3 * dx]], np.float32)
        xs = np.arange(2, s - 2, 4) * dx
        rec = np.stack([xs, np.full(xs.size, (s // 2) * dx), np.full(xs.size, 3 * dx)], 1).astype(np.float32)
        reqs.append(serve.Request(kind=kind, n_xyz=(s, s, s), T=nt * 0.001, dt=0.001, fm=15.0, pml_num=20,
                                  shots=shots, receivers=rec, model_seed=1000 + i, storage=st, rid=f"r{i}"))
    return reqs


def device_budget(frac=0.8):
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return frac * free.value


def run_rate(reqs, models, rate, budget, max_inflight, seed, graphs=False, batching=False):
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(reqs))
    if batching:
        srv = serve.Server(budget, max_inflight=max_inflight, batching=True)
    else:
        srv = serve.Server(budget, max_inflight=max_inflight,
                           runner=lambda r, m: serve.run_request(r, m, graphs=graphs))
    t0 = time.perf_counter()
    futs = []
    arrival = 0.0
    finished = itertools.count(1)  # next() is one call, safe from the server's threads

    def progress(f, rid):  # one line on stderr per finished request
        state = "failed" if f.exception() is not None else "done"
        print(f"serve_bench: {rid} {state} at {time.perf_counter() - t0:.1f} s ({next(finished)}/{len(reqs)})",
              file=sys.stderr, flush=True)

    for r, m, g in zip(reqs, models, gaps):
        arrival += g
        wait = t0 + arrival - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        f = srv.submit(r, m)
        f.add_done_callback(lambda f, rid=r.rid: progress(f, rid))
        futs.append(f)
    outs = [f.result() for f in futs]
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = np.array([o["latency_s"] for o in outs])
    cells = sum(o["cells"] for o in outs)
    return {"rate_per_s": rate, "requests": len(reqs), "wall_s": round(wall, 3),
            "requests_per_s": round(len(reqs) / wall, 3), "gcell_updates_per_s": round(cells / wall / 1e9, 3),
            "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
            "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
            "queue_p50_s": round(float(np.percentile([o["queue_s"] for o in outs], 50)), 4),
            "peak_inflight": st["peak_inflight"], "peak_bytes_gb": round(st["peak_bytes"] / 1e9, 2),
            "budget_gb": round(budget / 1e9, 1),
            "bf16_requests": sum(r.storage == "bf16" for r in reqs)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--rates", default="5,10,20,50")
    ap.add_argument("--sizes", default="48,128")
    ap.add_argument("--steps", default="200,500")
    ap.add_argument("--bf16", type=float, default=0.5)
    ap.add_argument("--kinds", default="model,rtm")
    ap.add_argument("--graphs", type=int, default=0, help="replay modelling steps from captured graphs")
    ap.add_argument("--batching", type=int, default=0,
                    help="continuous batching: FP32/BF16 modelling requests share fused batches")
    ap.add_argument("--budget-gb", type=float, default=0.0, help="HBM budget (default: 80%% of free memory)")
    ap.add_argument("--max-inflight", type=int, default=16)
    ap.add_argument("--seed", type=int, default=5)
    a = ap.parse_args()
    sizes = [int(v) for v in a.sizes.split(",")]
    steps = [int(v) for v in a.steps.split(",")]
    reqs = make_requests(a.requests, sizes, steps, a.bf16, a.seed, tuple(a.kinds.split(",")))
    models = [serve.default_model(r) for r in reqs]
    budget = a.budget_gb * 1e9 if a.budget_gb > 0 else device_budget()
    # one untimed request warms the HIP runtime and the brick cache
    serve.run_request(reqs[0], models[0])
    for k, rate in enumerate(float(v) for v in a.rates.split(",")):
        out = run_rate(reqs, models, rate, budget, a.max_inflight, a.seed + 17 * k, graphs=bool(a.graphs),
                       batching=bool(a.batching))
        out.update({"kinds": a.kinds, "graphs": bool(a.graphs), "batching": bool(a.batching), "sizes": sizes,
                    "steps": steps})
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
