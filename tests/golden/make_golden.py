"""Write the golden vectors of tests/golden/small_case.npz from the FP32 CPU
oracle (run from the repository root: python tests/golden/make_golden.py).

Case: 20x18x16 layered model (seed 11), 6 CPML layers, 60 steps, dt 1 ms,
fm 25 Hz, one shot, receivers on a plane at z = 2 cells; decomposition
2x2x1; VOWR ring widths 1, 2 and 4.  Stored: the kept receiver rows, the
three seismograms, vz after the last step, and per ring width the
per-subdomain Ipp image and its Laplacian.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"),
                os.path.join(ROOT, "domain-decomposition-release_amd")]

import pyoracle as po  # noqa: E402
from cases import case, oracle_for, oracle_rows  # noqa: E402

DIVIDE = (2, 2, 1)


def main():
    out = {}
    c = case(nx=20, ny=18, nz=16, pml=6, ntime=60)
    rows, subs = oracle_rows(c, DIVIDE)
    nodes = [n for _, _, n in rows]
    o = oracle_for(c)
    w = po.ricker(c["fm"], c["dt"], c["ntime"])
    shot = o.node(*c["shots"][0])
    seis = o.model_shot(shot, w, nodes)
    out["rows"] = np.array([(r, i) for r, i, _ in rows], np.int32)
    out["seis"] = seis.astype(np.float32)
    out["vz_last"] = o.get(0, po.VZ).astype(np.float32)
    obs = seis.astype(np.float32)
    for R in (1, 2, 4):
        cR = dict(c, rec_width=R)
        oR = oracle_for(cR)
        oR.rtm_shot(shot, w, nodes, obs)
        for s in subs:
            i0, j0, k0 = s["phy_off"]
            ni, nj, nk = s["phy_n"]
            img, lap = oR.finalize_box(i0, j0, k0, ni, nj, nk)
            out[f"ipp_R{R}_rank{s['rank']}"] = img[0].astype(np.float32)
            out[f"ipp_lap_R{R}_rank{s['rank']}"] = lap[0].astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "small_case.npz"), **out)
    print("wrote", os.path.join(HERE, "small_case.npz"), sum(v.nbytes for v in out.values()), "bytes")


if __name__ == "__main__":
    main()
