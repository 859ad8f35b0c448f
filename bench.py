"""Benchmark of the MI355X elastic FDTD / RTM engine.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rtm|model|precision|batch]
                [--storage fp32|fp16|bf16] [--ring-storage fp16|bf16|bfp16]

A step is one RTM shot of the 256x256x200 layered model (20 CPML layers,
decomposed 2x2x1 with every subdomain on the one GPU, 2000 time steps,
dt = 1 ms, fm = 15 Hz, one surface shot, 128x128 receivers, VOWR ring
width 1): forward modelling of the observed data, forward propagation with
the VOWR boundary save, then the reverse loop (backward reconstruction,
receiver back-propagation, imaging).  Inputs (model, wavelet, geometry) are
uploaded before the timed region; observed data stay in HBM.
--workload precision: one RTM shot of the 676x676x210 layered model with 2%
correlated heterogeneity (dx = 12.5 m, 20 CPML layers, one subdomain, 3000
steps, fm = 12 Hz, receivers every 16 cells).
--workload batch: --shots independent modelling shots of uneven grids (each
side 96-320 cells, 10-20 CPML layers, 1500 steps, fm 10-20 Hz), first one
at a time with the default bricks (the plain per-shot path), then all
submitted at once to the serving layer (autotuned bricks, up to --inflight
concurrent contexts), then as one fused batch (ddx_model_batch: each time
step launches the bricks and CPML kernels once over all shots; the
headline value, context setup included); a step is one pass over the
batch.  --storage
selects the wavefield storage format (FP32 by default; the 16-bit formats
keep FP32 arithmetic); --ring-storage keeps FP32 wavefields and stores only
the VOWR boundary in 16 bits.

value = cell-updates per second over the whole job (one cell-update is one
velocity+stress step of one gridext cell, gridphy for the backward pass).
The JSON line also carries shots/s, per-shot latency percentiles, the
HBM roofline of the dominant kernel (HIP events on the engine stream over
the timed region), per-class rows for every kernel class (one more,
untimed shot with every class timed), the plain per-shot FP32 kernels'
rate on the same shot and GPU (plain_fp32, vs_plain = value / its value)
and a CPU baseline (the CPU oracle on this host's cores: the 100^3
modelling configuration in full, FP32 and FP64).
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
                2: "CPML velocity (k_vel_pml)", 3: "CPML stress (k_stress_pml)", 4: "CMM halo exchange (k_halo)",
                5: "imaging (k_image)"}
# kernel-name prefixes of each class, as rocprofv3 reports them
KERNEL_PREFIX = {0: ("ddx::k_vel_phy", "ddx::k_vel_lds"), 1: ("ddx::k_stress_phy", "ddx::k_stress_lds"),
                 2: ("ddx::k_vel_pml",), 3: ("ddx::k_stress_pml",), 4: ("ddx::k_halo",), 5: ("ddx::k_image",)}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
KERNEL_SRC = os.path.join(ROOT, "domain-decomposition-release_amd", "csrc", "kernels.hip")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_engine(cfg, storage="fp32", ring_storage="fp32"):
    d = ddx.make_desc(cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["pml"], cfg["dt"], cfg["fm"], cfg["ntime"],
                      divide=cfg["divide"], rec_width=cfg.get("rec_width", 1),
                      image_fp64=int(cfg.get("image_fp64", 0)), cfl_policy=0,
                      storage=storage, ring_storage=ring_storage)
    e = ddx.Engine(d)
    e.set_model(cfg["vp"], cfg["vs"], cfg["rho"])
    e.set_wavelet(ddx.ricker(cfg["fm"], cfg["dt"], cfg["ntime"]))
    e.set_geometry(cfg["shots"], cfg["recv"])
    return e


def host_info():
    """nproc, the CPU model name and socket count (lscpu), the threads this
    process may use, and whether the CPU has AVX-512 (x86-64-v4)."""
    import subprocess
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Model name":
                info["model_name"] = v.strip()
            elif k.strip() == "Socket(s)":
                info["sockets"] = int(v.strip()) if v.strip().isdigit() else v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as fh:
            flags = next((ln.split(":", 1)[1].split() for ln in fh if ln.startswith("flags")), [])
        info["avx512"] = all(f in flags for f in ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"))
    except OSError:
        info["avx512"] = False
    return info


def cpu_baseline():
    """The CPU oracle (the restated reference method: same stencils, CPML,
    source and receiver rules and operation order; OpenMP over grid planes)
    on this host's cores, on the 100^3 modelling configuration in full: 2
    shots x 1000 steps of the 120^3 extended grid with 8 receivers, in FP32,
    then the same in FP64 (time and error against FP32 reported apart).  The
    reference has no CPU path of its own."""
    info = host_info()
    threads = int(os.environ.get("OMP_NUM_THREADS") or info["affinity_cpus"])
    os.environ["OMP_NUM_THREADS"] = str(threads)
    if info.get("avx512"):
        os.environ["DDX_ORACLE_ISA"] = "v4"
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    c = synth.config1()
    nt = c["ntime"]
    w = po.ricker(c["fm"], c["dt"], nt)
    res = {}
    for prec in (4, 8):
        o = po.Oracle(c["nx"], c["ny"], c["nz"], c["dx"], c["pml"], c["dt"], c["fm"], nt, 1, c["vp"], c["vs"],
                      c["rho"], prec=prec)
        nodes = [o.node(*r) for r in c["recv"]]
        t0 = time.perf_counter()
        seis = [o.model_shot(o.node(*sh), w, nodes) for sh in c["shots"]]
        el = time.perf_counter() - t0
        cells = float(o.NX * o.NY * o.NZ) * nt * len(c["shots"])
        o.close()
        res[prec] = (el, cells, np.array(seis))
        log(f"bench: CPU oracle FP{prec * 8}: {el:.1f} s for {len(c['shots'])} shots")
    el, cells, s32 = res[4]
    el8, _, s64 = res[8]
    err = float(np.linalg.norm(s32 - s64) / np.linalg.norm(s64))
    return {"value": round(cells / el / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"100^3 + 10 CPML (120^3 extended), {nt} steps, {len(c['shots'])} shots, 8 receivers, in "
                      f"full; FP32 CPU oracle, OpenMP {threads} threads, "
                      f"{'x86-64-v4 (AVX-512)' if info.get('avx512') else 'x86-64-v3 (AVX2)'} build",
            "shots_per_s": round(len(c["shots"]) / el, 4), "wall_s": round(el, 2),
            "fp64": {"value": round(cells / el8 / 1e9, 4), "wall_s": round(el8, 2),
                     "seismogram_rel_l2_fp32_vs_fp64": err},
            "host": info}


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


def workload_key(kind, rec_width=1, storage="fp32"):
    """Workload name without the step count (per-launch figures do not
    depend on it): the model grid, CPML layers, split, and the storage
    format unless FP32."""
    grid = "676x676x210_pml20_divide1x1x1" if kind == "precision" else "256x256x200_pml20_divide2x2x1"
    return (("model_shot" if kind == "model" else "rtm_shot") + f"_{grid}_R{rec_width}" +
            ("" if storage == "fp32" else f"_{storage}"))


def source_digest():
    import hashlib
    with open(KERNEL_SRC, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def kernel_class(name):
    short = name.split("(")[0].replace("void ", "").strip()
    for k, prefixes in KERNEL_PREFIX.items():
        if short.startswith(prefixes):
            return k
    return None


def pmc_summary(fetch_db, write_db, workload):
    """HBM traffic per launch of each kernel class from two rocprofv3 --pmc
    runs (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what
    the counter reads for the 512 MiB device-to-device copies that
    --copy-calibration puts in the same run.  Prints a markdown table and
    writes profiles/traffic.json, tagged with the kernel source digest."""
    import sqlite3
    from collections import defaultdict

    def load(db, counter):
        out = defaultdict(list)
        con = sqlite3.connect(db)
        for name, v in con.execute("select name, counter_value from pmc_events where counter_name = ?", (counter,)):
            out[name].append(v)
        return out
    f, w = load(fetch_db, "FETCH_SIZE"), load(write_db, "WRITE_SIZE")

    def calib(d):
        big = sorted(v for n, vs in d.items() if "copyBuffer" in n for v in vs)[-2:]
        return (512 * 1024) / (sum(big) / len(big))  # KiB of a 512 MiB copy / counter reading
    cf, cw = calib(f), calib(w)
    per_class = defaultdict(lambda: [0, 0.0, 0.0])
    per_kernel = {}
    rows = []
    for name in sorted(f, key=lambda n: -sum(f[n])):
        k = kernel_class(name)
        if name not in w or k is None:
            continue
        fa = sum(f[name]) / len(f[name]) * cf * 1024.0
        wa = sum(w[name]) / len(w[name]) * cw * 1024.0
        short = name.split("(")[0].replace("void ", "")
        rows.append((short, len(f[name]), fa / 1e6, wa / 1e6))
        per_kernel[short] = {"dispatches": len(f[name]), "fetch_bytes": fa, "write_bytes": wa,
                             "bytes_per_launch": fa + wa}
        pc = per_class[k]
        pc[0] += len(f[name])
        pc[1] += fa * len(f[name])
        pc[2] += wa * len(f[name])
    print(f"calibration: FETCH x{cf:.3f}, WRITE x{cw:.3f} (512 MiB device-to-device copy)")
    print("| kernel | dispatches | FETCH MB/dispatch | WRITE MB/dispatch |")
    print("|---|---:|---:|---:|")
    for r in rows:
        print(f"| `{r[0]}` | {r[1]} | {r[2]:.1f} | {r[3]:.1f} |")
    doc = {"workload": workload, "kernel_source_sha256_16": source_digest(), "fetch_factor": round(cf, 4),
           "write_factor": round(cw, 4),
           "classes": {str(k): {"name": KERNEL_NAMES[k], "dispatches": v[0], "fetch_bytes": v[1] / v[0],
                                "write_bytes": v[2] / v[0], "bytes_per_launch": (v[1] + v[2]) / v[0]}
                       for k, v in sorted(per_class.items())},
           "kernels": per_kernel}
    with open(TRAFFIC_FILE, "w") as fh:
        json.dump(doc, fh, indent=1)
    print(f"wrote {TRAFFIC_FILE}")


def brick_kernel_names(e, kernel, modes, storage):
    """Kernel names (as rocprofv3 reports them) of the brick shapes engine e
    uses for the physical-region kernel class `kernel` in `modes`."""
    st = {"fp32": "ddx::F32", "fp16": "ddx::F16", "bf16": "ddx::BF16"}[storage]
    names = []
    for m in modes:
        tx, ty, cy, kz, form = e.brick(kernel, m)
        tm = 0 if (kernel == 0 and m == ddx.RTM_FORWARD) else m  # forward velocity shares the pure-forward kernel
        base = "k_vel_phy_zb" if kernel == 0 else "k_stress_phy_zb"
        if form == 2:  # paired bricks
            names.append(f"ddx::{base}2<{tm}, {tx}, {ty}, {cy}, {kz}, {st}>")
        elif form == 4:  # the y arm from LDS
            names.append(f"ddx::{'k_vel_phy_yl' if kernel == 0 else 'k_stress_phy_yl'}<{tm}, {ty}, {kz}, {st}>")
        else:  # one cell per thread; form 3 takes the x arm from neighbouring lanes
            names.append(f"ddx::{base}<{tm}, {tx}, {ty}, {cy}, {kz}, {st}, {'true' if form == 3 else 'false'}>")
    return names


def measured_traffic(kernel, workload, names=None):
    """HBM bytes per launch of `kernel` from profiles/traffic.json when it was
    measured on this workload with the current kernel source, else None.
    For the physical-region classes, `names` (the brick kernels this run
    launches, equally often) select the per-kernel figures, so the shapes the
    autotuner timed in the profiled run do not enter the average."""
    try:
        with open(TRAFFIC_FILE) as fh:
            doc = json.load(fh)
        if doc.get("kernel_source_sha256_16") != source_digest() or doc.get("workload") != workload:
            return None
        if names:
            return sum(doc["kernels"][n]["bytes_per_launch"] for n in names) / len(names)
        return doc["classes"][str(kernel)]["bytes_per_launch"]
    except (OSError, KeyError, ValueError, ZeroDivisionError):
        return None


def class_rows(e, one_step, workload, storage):
    """Every kernel class of one more shot (untimed, after the timed region),
    timed at once with HIP events on the engine stream: launches, mean launch
    time, algorithmic bytes per launch, achieved GB/s and its fraction of the
    8 TB/s HBM peak, share of the summed kernel time, and the measured HBM
    traffic per launch where profiles/traffic.json holds it for this
    workload and kernel source."""
    e.profile(-1)
    one_step()
    rows = {}
    tot = 0.0
    for k in range(6):
        ms, n, by = e.class_time(k)
        tot += ms
        rows[k] = (ms, n, by)
    e.profile(-2)
    out = {}
    key = workload_key(workload, storage=storage)
    for k, (ms, n, by) in rows.items():
        if n == 0:
            continue
        gbs = by / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        t = measured_traffic(k, key) if k >= 2 else None
        out[str(k)] = {"kernel": KERNEL_NAMES[k], "launches": n, "avg_launch_us": round(ms * 1e3 / n, 2),
                       "algorithmic_bytes_per_launch": round(by / n), "achieved_gbs": round(gbs, 1),
                       "frac": round(gbs / HBM_PEAK_GBS, 4), "share_of_kernel_time": round(ms / tot, 4),
                       "traffic_bytes_per_launch": None if t is None else round(t)}
    log("bench: kernel classes " + ", ".join(f"{v['kernel'].split(' (')[0]} {v['avg_launch_us']} us "
                                              f"{v['frac']:.3f}" for v in out.values()))
    return out


def plain_shot(cfg, workload):
    """One shot of the same workload on the plain per-shot FP32 kernels:
    one thread per cell and one plane per block (64 x 4 threads) in every
    physical-region mode, CPML blocks in launch order, no autotuning.  That
    is the reference's kernel structure, on this GPU.  One untimed modelling
    shot first builds the tables."""
    env = {"DDX_TILING": "64x4x1x0", "DDX_PML_GROUP": "0"}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        e = make_engine(cfg)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    try:
        e.model_shot(0)
        e.synchronize()
        c0 = e.cell_updates()
        t0 = time.perf_counter()
        e.model_shot(0)
        if workload != "model":
            e.rtm_shot(0)
        e.synchronize()
        dt = time.perf_counter() - t0
        cells = e.cell_updates() - c0
        bricks = {f"{k}{m}": list(e.brick(ki, m)) for ki, k in enumerate(("vel", "stress")) for m in range(4)}
    finally:
        e.close()
    log(f"bench: plain per-shot FP32 kernels {dt:.3f} s per shot")
    return {"value": round(cells / dt / 1e9, 3), "unit": "Gcell-updates/s", "shots_per_s": round(1.0 / dt, 4),
            "steps": 1, "kernels": "one thread per cell, one plane per block (64x4), CPML in launch order, "
            "no autotune", "bricks": bricks}


def fused_batch(reqs, models, steps):
    """The same shots as one fused batch (ddx_model_batch): every context
    created and loaded by 8 host threads (the setup, timed apart), the first
    one autotuned, then `steps` passes in which each time step launches the
    bricks and CPML kernels once over all the shots."""
    from concurrent.futures import ThreadPoolExecutor

    def load(arg):
        r, (vp, vs, rho) = arg
        e = ddx.Engine(r.desc())
        try:
            e.set_model(vp, vs, rho)
            e.set_wavelet(ddx.ricker(r.fm, r.dt, r.ntime))
            e.set_geometry(r.shots, r.receivers)
        except Exception:
            e.close()
            raise
        return e

    t0 = time.perf_counter()
    engines = []
    try:
        with ThreadPoolExecutor(max_workers=8) as ex:
            futs = [ex.submit(load, x) for x in zip(reqs, models)]
            errors = []
            for f in futs:
                try:
                    engines.append(f.result())
                except Exception as err:  # keep the others to close them
                    errors.append(err)
            if errors:
                raise errors[0]
        engines[0].autotune()
        for e in engines:
            e.synchronize()
        t_setup = time.perf_counter() - t0
        c0 = sum(e.cell_updates() for e in engines)
        t0 = time.perf_counter()
        for k in range(steps):
            ddx.model_batch(engines)
            log(f"bench: fused pass {k} {time.perf_counter() - t0:.2f} s")
        t_run = time.perf_counter() - t0
        cells = sum(e.cell_updates() for e in engines) - c0
    finally:
        for e in engines:
            e.close()
    n = steps * len(reqs)
    return {"shots_per_s": round(n / t_run, 4), "gcell_updates_per_s": round(cells / t_run / 1e9, 3),
            "setup_s": round(t_setup, 2), "shots_per_s_with_setup": round(n / (t_run + t_setup), 4),
            "gcell_updates_per_s_with_setup": round(cells / (t_run + t_setup) / 1e9, 3),
            "seconds": t_run + t_setup, "bricks": "autotuned on the first shot"}


def run_batch(a):
    """Batched and served modelling shots against the plain per-shot path."""
    from ddx import serve
    cfgs = synth.config3(a.shots)
    reqs, models = [], []
    for i, c in enumerate(cfgs):
        reqs.append(serve.Request(kind="model", n_xyz=(c["nx"], c["ny"], c["nz"]), pml_num=c["pml"], fm=c["fm"],
                                  dt=c["dt"], T=c["ntime"] * c["dt"], shots=c["shots"], receivers=c["recv"],
                                  storage=a.storage, rid=f"b{i}"))
        models.append((c["vp"], c["vs"], c["rho"]))
    log(f"bench: batch of {len(reqs)} shots, cells per step "
        f"{sum((c['nx'] + 2 * c['pml']) * (c['ny'] + 2 * c['pml']) * (c['nz'] + 2 * c['pml']) for c in cfgs) / 1e6:.1f} M")
    for i in range(a.warmup):
        serve.run_request(reqs[0], models[0])
    # plain per-shot path: one context at a time, default bricks
    t0 = time.perf_counter()
    cells_plain = 0.0
    for r, m in zip(reqs, models):
        cells_plain += serve.run_request(r, m, autotune=False)["cells"]
    t_plain = time.perf_counter() - t0
    log(f"bench: plain per-shot {t_plain:.2f} s")
    lat, cells, t_all = [], 0.0, 0.0
    for k in range(a.steps):
        srv = serve.Server(1e15, max_inflight=a.inflight)
        t0 = time.perf_counter()
        futs = [srv.submit(r, m) for r, m in zip(reqs, models)]
        outs = [f.result() for f in futs]
        t_all += time.perf_counter() - t0
        srv.close()
        lat += [o["latency_s"] for o in outs]
        cells += sum(o["cells"] for o in outs)
        log(f"bench: batched pass {k} {t_all:.2f} s")
    fused = fused_batch(reqs, models, a.steps)
    shots_b = a.steps * len(reqs) / t_all
    shots_p = len(reqs) / t_plain
    # the headline is the fused batch, setup included (contexts created and loaded)
    out = {
        "metric": METRIC, "value": fused["gcell_updates_per_s_with_setup"], "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(fused["seconds"] / a.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batch_{len(reqs)}_modelling_shots_96to320_pml10to20_1500steps", "shots": len(reqs),
                   "form": "fused batch (ddx_model_batch), setup included", "storage": a.storage},
        "shots_per_s": fused["shots_per_s_with_setup"],
        "fused": fused,
        "served": {"shots_per_s": round(shots_b, 4), "gcell_updates_per_s": round(cells / t_all / 1e9, 3),
                   "inflight": a.inflight, "latency_ms_p50": round(float(np.percentile(lat, 50)) * 1e3, 2),
                   "latency_ms_p99": round(float(np.percentile(lat, 99)) * 1e3, 2)},
        "plain_per_shot": {"shots_per_s": round(shots_p, 4), "gcell_updates_per_s": round(cells_plain / t_plain / 1e9, 3)},
        "vs_plain": round(fused["shots_per_s_with_setup"] / shots_p, 3),
        "served_vs_plain": round(shots_b / shots_p, 3),
        "roofline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["rtm", "model", "precision", "batch"], default="rtm")
    ap.add_argument("--shots", type=int, default=16, help="batch workload: shots per batch")
    ap.add_argument("--inflight", type=int, default=16, help="batch workload: concurrent contexts")
    ap.add_argument("--roofline-kernel", type=int, default=1,
                    help="kernel class of the roofline entry (0 vel phy, 1 stress phy, 2/3 CPML, 4 halo, 5 image)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-classes", action="store_true",
                    help="skip the untimed shot that times every kernel class (the per-class rows)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain per-shot FP32 kernels' shot (the vs_plain ratio)")
    ap.add_argument("--storage", choices=["fp32", "fp16", "bf16"], default="fp32",
                    help="wavefield storage format (arithmetic stays FP32)")
    ap.add_argument("--ring-storage", choices=["fp32", "fp16", "bf16", "bfp16"], default="fp32",
                    help="VOWR boundary store format over FP32 wavefields (fp32: as --storage)")
    ap.add_argument("--exchange", type=int, default=-1,
                    help="CMM exchange mode (ddx_set_exchange): 0 direct copy (default), 1 contiguous message "
                         "blocks, 2 pack on store (needs DDX_MIRROR=1 in the environment)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default brick shapes instead of ddx_autotune's choice")
    ap.add_argument("--ntime", type=int, default=0, help="override the time steps (profiling runs only)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="first run N device-to-device copies of 512 MiB (counter calibration)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DB", "WRITE_DB"),
                    help="summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) into profiles/traffic.json")
    a = ap.parse_args()
    if a.pmc_summary:
        pmc_summary(a.pmc_summary[0], a.pmc_summary[1], workload_key(a.workload, storage=a.storage))
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); "
              "every subdomain of the decomposition is resident on that GPU.", file=sys.stderr)
        sys.exit(2)

    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
    if a.workload == "batch":
        run_batch(a)
        return
    cfg = synth.config4() if a.workload == "precision" else synth.config2(rec_width=1)
    if a.ntime:
        cfg["ntime"] = a.ntime
    log(f"bench: workload {a.workload}, grid {cfg['nx']}x{cfg['ny']}x{cfg['nz']} pml {cfg['pml']} "
        f"divide {cfg['divide']} ntime {cfg['ntime']} receivers {len(cfg['recv'])} (vp clipped in "
        f"{cfg['clipped']} cells for stability)")
    if a.ring_storage != "fp32" and a.storage != "fp32":
        ap.error("--ring-storage needs --storage fp32")
    e = make_engine(cfg, a.storage, a.ring_storage)
    if a.exchange >= 0:
        e.set_exchange(a.exchange)
    tuned = None
    if not a.no_autotune:
        t = time.perf_counter()
        n = e.autotune()
        tuned = {"seconds": round(time.perf_counter() - t, 2), "shapes_timed": n,
                 "bricks": {f"{k}{m}": list(e.brick(ki, m)) for ki, k in enumerate(("vel", "stress")) for m in range(4)}}
        log(f"bench: autotune {tuned['seconds']} s, {n} shapes timed: {tuned['bricks']}")

    def one_step():
        e.model_shot(0)
        if a.workload != "model":
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
    dev_gib = e.device_bytes() / 2**30  # after the first shot: the VOWR store is allocated then
    achieved = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    value = cells / total / 1e9
    names = None
    if a.roofline_kernel in (0, 1):
        modes = [ddx.PURE_FORWARD] if a.workload == "model" else [ddx.PURE_FORWARD, ddx.RTM_FORWARD,
                                                                   ddx.RTM_BACKWARD, ddx.RTM_REVERSE]
        names = brick_kernel_names(e, a.roofline_kernel, modes, a.storage)
    traffic = measured_traffic(a.roofline_kernel, workload_key(a.workload, storage=a.storage), names)
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
            "workload": ("model_shot" if a.workload == "model" else "rtm_shot") +
                        f"_{cfg['nx']}x{cfg['ny']}x{cfg['nz']}_pml{cfg['pml']}_divide"
                        f"{'x'.join(str(v) for v in cfg['divide'])}_{cfg['ntime']}steps_R1" +
                        ("" if a.storage == "fp32" else f"_{a.storage}_storage") +
                        ("" if a.ring_storage == "fp32" else f"_{a.ring_storage}_ring"),
            "grid": [cfg["nx"], cfg["ny"], cfg["nz"]], "pml": cfg["pml"], "divide": list(cfg["divide"]),
            "ntime": cfg["ntime"], "receivers": int(len(cfg["recv"])), "rec_width": 1, "storage": a.storage,
            "ring_storage": a.ring_storage, "image_fp64": int(cfg.get("image_fp64", 0)),
            "bricks": "default" if tuned is None else "autotuned (ddx_autotune, outside the timed region)",
            "exchange": e.exchange(),
        },
        "shots_per_s": round(a.steps / total, 4),
        "device_gib": round(dev_gib, 2),
        "latency_ms_p50": round(float(np.percentile(lat, 50)) * 1e3, 2),
        "latency_ms_p99": round(float(np.percentile(lat, 99)) * 1e3, 2),
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic if traffic is None else round(traffic),
            "traffic_source": "profiles/traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
                              if traffic is not None else "not measured for this kernel source",
            "kernel": KERNEL_NAMES.get(a.roofline_kernel, str(a.roofline_kernel)),
            "launches": kn, "avg_launch_us": round(kms * 1e3 / max(kn, 1), 2),
            "bytes_per_launch": round(kbytes / max(kn, 1)),
        },
    }
    if tuned is not None:
        out["autotune"] = tuned
    if not a.no_classes:
        out["classes"] = class_rows(e, one_step, a.workload, a.storage)
        # the whole shot against the roofline: every class's algorithmic bytes
        # over the summed kernel time, and over the shot's wall time
        by = sum(v["algorithmic_bytes_per_launch"] * v["launches"] for v in out["classes"].values())
        ks = sum(v["avg_launch_us"] * 1e-6 * v["launches"] for v in out["classes"].values())
        wall = total / a.steps
        out["shot_roofline"] = {
            "algorithmic_bytes_per_shot": round(by), "kernel_s_per_shot": round(ks, 4), "wall_s_per_shot": round(wall, 4),
            "frac_of_kernel_time": round(by / ks / 1e9 / HBM_PEAK_GBS, 4) if ks > 0 else None,
            "frac_of_wall_time": round(by / wall / 1e9 / HBM_PEAK_GBS, 4)}
    e.close()
    if not a.no_plain and a.workload in ("rtm", "model", "precision") and a.storage == "fp32":
        out["plain_fp32"] = plain_shot(cfg, a.workload)
        out["vs_plain"] = round(value / out["plain_fp32"]["value"], 3)
    if not a.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline()
        except Exception as ex:  # the baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": str(ex)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
