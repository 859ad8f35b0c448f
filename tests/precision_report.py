"""Accuracy of the wavefield storage formats against the FP64 oracle on the
200^3 crop of the 676x676x210 precision workload (3000 steps, one RTM shot).

Not collected by pytest (it takes tens of minutes).  Three steps:
  python tests/precision_report.py gpu OUT_DIR [RUNS [STRIDE]]  # on the GPU: fp32, fp16, bf16,
                                                     # ring_fp16, ring_fp16_slice, ring_bf16, ring_bfp16
  python tests/precision_report.py oracle OUT_DIR    # FP64 CPU oracle (about 40 min on 8 cores)
  python tests/precision_report.py compare OUT_DIR   # relative L2 errors, JSON
Saved per run: the seismograms (3 x 144 receivers x 3000 steps), the forward
vx snapshot after step 1499 and the raw RTM images Ip, Ipp, all on every
second cell of the crop's gridphy (the comparison grid; the GPU runs may
save every STRIDE-th cell, a multiple of 2, and compare() then takes the
same cells of the oracle's).  ring_fp16 and ring_bf16 keep FP32 wavefields
and store only the VOWR boundary in 16 bits; ring_fp16_slice does so with
one power-of-two scale per ring field and time slice (ddx_set_ring_scaling),
ring_bfp16 as 16-bit integers with one exponent per 64 store entries.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "domain-decomposition-release_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

from ddx import synth  # noqa: E402

CROP = 200
EVERY = 1500
RUNS = {"fp32": ("fp32", "fp32"), "fp16": ("fp16", "fp32"), "bf16": ("bf16", "fp32"),
        "ring_fp16": ("fp32", "fp16"), "ring_fp16_slice": ("fp32", "fp16"), "ring_bf16": ("fp32", "bf16"),
        "ring_bfp16": ("fp32", "bfp16")}


def run_gpu(out, runs="fp32,fp16,bf16,ring_fp16,ring_fp16_slice,ring_bf16,ring_bfp16", stride="2"):
    import ddx
    k = int(stride)
    c = synth.config4(crop=CROP)
    for st in runs.split(","):
        t = time.perf_counter()
        storage, ring = RUNS[st]
        d = ddx.make_desc(c["nx"], c["ny"], c["nz"], c["dx"], c["pml"], c["dt"], c["fm"], c["ntime"],
                          rec_width=1, image_fp64=c["image_fp64"], storage=storage, ring_storage=ring)
        e = ddx.Engine(d)
        e.set_model(c["vp"], c["vs"], c["rho"])
        e.set_wavelet(ddx.ricker(c["fm"], c["dt"], c["ntime"]))
        e.set_geometry(c["shots"], c["recv"])
        if st.endswith("_slice"):
            e.set_ring_scaling(1)
        e.autotune()
        e.model_shot(0)
        seis = e.seismograms(0)
        e.set_snapshots(ddx.VX, EVERY)
        e.rtm_shot(0)
        snap = e.snapshot(0, 1)[::k, ::k, ::k]
        ip = e.image(ddx.IP, 0)[::k, ::k, ::k]
        ipp = e.image(ddx.IPP, 0)[::k, ::k, ::k]
        scales = (e.storage_scale(ddx.CLASS_VELOCITY), e.storage_scale(ddx.CLASS_STRESS))
        e.close()
        np.savez_compressed(os.path.join(out, f"gpu_{st}.npz"), seis=seis, snap=snap, ip=ip, ipp=ipp,
                            scales=np.array(scales))
        print(f"precision crop: {st} done in {time.perf_counter() - t:.1f} s, scales {scales}", flush=True)


def run_oracle(out):
    import pyoracle as po
    c = synth.config4(crop=CROP)
    nx, ny, nz, dx, pml, nt = c["nx"], c["ny"], c["nz"], c["dx"], c["pml"], c["ntime"]
    subs = po.decompose((nx, ny, nz), (dx,) * 3, pml, (1, 1, 1))
    rows = po.kept_receivers(subs, c["recv"], (dx,) * 3, pml)
    nodes = [po.ext_index(n, nx + 2 * pml, ny + 2 * pml) for (_, _, n) in rows]
    w = po.ricker(c["fm"], c["dt"], nt)
    o = po.Oracle(nx, ny, nz, dx, pml, c["dt"], c["fm"], nt, 1, c["vp"], c["vs"], c["rho"], prec=8)
    shot = o.node(*c["shots"][0])
    seis = o.model_shot(shot, w, nodes)
    fs, _ = o.rtm_shot(shot, w, nodes, seis.astype(np.float32), snap_every=EVERY)
    m = pml
    snap = fs[1][m:m + nz, m:m + ny, m:m + nx][::2, ::2, ::2]
    np.savez_compressed(os.path.join(out, "oracle_fp64.npz"), seis=seis, snap=snap,
                        ip=o.image(0)[::2, ::2, ::2], ipp=o.image(1)[::2, ::2, ::2])


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def compare(out):
    ref = np.load(os.path.join(out, "oracle_fp64.npz"))
    rep = {}
    for st in RUNS:
        if not os.path.exists(os.path.join(out, f"gpu_{st}.npz")):
            continue
        g = np.load(os.path.join(out, f"gpu_{st}.npz"))
        q = ref["ip"].shape[0] // g["ip"].shape[0] if g["ip"].shape[0] else 1
        ref_st = {n: ref[n][::q, ::q, ::q] for n in ("snap", "ip", "ipp")}
        comp_g = np.where(g["ip"] > 3.4e-38, g["ipp"] / np.maximum(g["ip"], 1e-30), g["ipp"])
        comp_r = np.where(ref_st["ip"] > 3.4e-38, ref_st["ipp"] / np.maximum(ref_st["ip"], 1e-30), ref_st["ipp"])
        rep[st] = {"seismograms": rel(g["seis"], ref["seis"]), "vx_snapshot_step1499": rel(g["snap"], ref_st["snap"]),
                   "image_ip": rel(g["ip"], ref_st["ip"]), "image_ipp_raw": rel(g["ipp"], ref_st["ipp"]),
                   "image_ipp_compensated": rel(comp_g, comp_r), "scales_log2_v_s": g["scales"].tolist()}
    print(json.dumps(rep, indent=1))
    return rep


if __name__ == "__main__":
    what, out = sys.argv[1], sys.argv[2]
    os.makedirs(out, exist_ok=True)
    {"gpu": run_gpu, "oracle": run_oracle, "compare": compare}[what](out, *sys.argv[3:])
