"""Benchmark of the MI355X elastic FDTD / RTM engine.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rtm|model]

A step is one RTM shot of the 256x256x200 layered model (20 CPML layers,
decomposed 2x2x1 with every subdomain on the one GPU, 2000 time steps,
dt = 1 ms, fm = 15 Hz, one surface shot, 128x128 receivers, VOWR ring
width 1): forward modelling of the observed data, forward propagation with
the VOWR boundary save, then the reverse loop (backward reconstruction,
receiver back-propagation, imaging).  Inputs (model, wavelet, geometry) are
uploaded before the timed region; observed data stay in HBM.

value = cell-updates per second over the whole job (one cell-update is one
velocity+stress step of one gridext cell, gridphy for the backward pass).
The JSON line also carries shots/s, per-shot latency percentiles, the
HBM roofline of the dominant kernel (HIP events on the engine stream) and a
CPU baseline (the FP32 CPU oracle on this host, bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "domain-decomposition-release_amd"))

import numpy as np  # noqa: E402

import ddx  # noqa: E402
from ddx import synth  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "Gcell-updates/s + shots/s, 1 MI355X, vs plain FP32; p50/p99 shot latency"
KERNEL_NAMES = {0: "velocity update, physical region (k_vel_phy)", 1: "stress update, physical region (k_stress_phy)",
                2: "CPML velocity (k_vel_pml)", 3: "CPML stress (k_stress_pml)"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_engine(cfg):
    d = ddx.make_desc(cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["pml"], cfg["dt"], cfg["fm"], cfg["ntime"],
                      divide=cfg["divide"], rec_width=cfg.get("rec_width", 1), image_fp64=0, cfl_policy=0)
    e = ddx.Engine(d)
    e.set_model(cfg["vp"], cfg["vs"], cfg["rho"])
    e.set_wavelet(ddx.ricker(cfg["fm"], cfg["dt"], cfg["ntime"]))
    e.set_geometry(cfg["shots"], cfg["recv"])
    return e


def cpu_baseline(cfg, budget_s=12.0):
    """FP32 CPU oracle (the restated reference method, OpenMP over grid
    planes) on this host: pure forward modelling steps on the same global
    grid until about budget_s seconds of stepping have run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    threads = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    o = po.Oracle(cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["pml"], cfg["dt"], cfg["fm"], cfg["ntime"], 1,
                  cfg["vp"], cfg["vs"], cfg["rho"], prec=4)
    w = ddx.ricker(cfg["fm"], cfg["dt"], cfg["ntime"])
    node = o.node(*cfg["shots"][0])
    o.reset(0)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.vel(0, po.PURE_FORWARD)
        o.source(0, node, float(w[n]), 1, 0)
        o.stress(0, po.PURE_FORWARD)
        n += 1
        el = time.perf_counter() - t0
        if (el > budget_s and n >= 2) or n >= cfg["ntime"]:
            break
    cells = float(o.NX * o.NY * o.NZ) * n
    o.close()
    return {"value": cells / el / 1e9, "unit": "Gcell-updates/s", "cores": int(os.environ.get("OMP_NUM_THREADS", threads)),
            "kind": "port",
            "sample": f"{n} pure-forward time steps of the same {cfg['nx']}x{cfg['ny']}x{cfg['nz']} grid "
                      f"(+{cfg['pml']} CPML) in {el:.1f} s, FP32 oracle"}


def copy_calibration(n, nbytes=512 << 20):
    """hipMemcpy device-to-device of a known size (one copy kernel per call):
    the reference point that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(n):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"bench: copy calibration {n} x {nbytes} bytes")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["rtm", "model"], default="rtm")
    ap.add_argument("--roofline-kernel", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ntime", type=int, default=0, help="override the time steps (profiling runs only)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="first run N device-to-device copies of 512 MiB (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {a.gpus}); "
              "every subdomain of the decomposition is resident on that GPU.", file=sys.stderr)
        sys.exit(2)

    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
    cfg = synth.config2(rec_width=1)
    if a.ntime:
        cfg["ntime"] = a.ntime
    log(f"bench: workload {a.workload}, grid {cfg['nx']}x{cfg['ny']}x{cfg['nz']} pml {cfg['pml']} "
        f"divide {cfg['divide']} ntime {cfg['ntime']} receivers {len(cfg['recv'])} (vp clipped in "
        f"{cfg['clipped']} cells for stability)")
    e = make_engine(cfg)

    def one_step():
        e.model_shot(0)
        if a.workload == "rtm":
            e.rtm_shot(0)
        e.synchronize()

    for i in range(a.warmup):
        t = time.perf_counter()
        one_step()
        log(f"bench: warmup step {i} {time.perf_counter() - t:.3f} s")
    c0 = e.cell_updates()
    e.profile(a.roofline_kernel)
    e.synchronize()
    lat = []
    t0 = time.perf_counter()
    for i in range(a.steps):
        t = time.perf_counter()
        one_step()
        lat.append(time.perf_counter() - t)
        log(f"bench: step {i} {lat[-1]:.3f} s")
    e.synchronize()
    total = time.perf_counter() - t0
    cells = e.cell_updates() - c0
    kms, kn, kbytes = e.kernel_time()
    achieved = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    value = cells / total / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(total / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("rtm_shot" if a.workload == "rtm" else "model_shot") +
                        f"_{cfg['nx']}x{cfg['ny']}x{cfg['nz']}_pml{cfg['pml']}_divide2x2x1_{cfg['ntime']}steps_R1",
            "grid": [cfg["nx"], cfg["ny"], cfg["nz"]], "pml": cfg["pml"], "divide": list(cfg["divide"]),
            "ntime": cfg["ntime"], "receivers": int(len(cfg["recv"])), "rec_width": 1,
        },
        "shots_per_s": round(a.steps / total, 4),
        "latency_ms_p50": round(float(np.percentile(lat, 50)) * 1e3, 2),
        "latency_ms_p99": round(float(np.percentile(lat, 99)) * 1e3, 2),
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": KERNEL_NAMES.get(a.roofline_kernel, str(a.roofline_kernel)),
            "launches": kn, "avg_launch_us": round(kms * 1e3 / max(kn, 1), 2),
            "bytes_per_launch": round(kbytes / max(kn, 1)),
        },
    }
    e.close()
    if not a.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(cfg)
        except Exception as ex:  # the baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": str(ex)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
