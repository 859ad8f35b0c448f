"""Golden vectors (tests/golden/small_case.npz, written by
tests/golden/make_golden.py from the FP32 oracle): the oracle must keep
reproducing them exactly, and the GPU engine must match them."""
import os

import numpy as np
import pytest

import ddx
import pyoracle as po
from cases import case, oracle_for, oracle_rows
from conftest import rel_l2

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "small_case.npz"))
DIVIDE = (2, 2, 1)
C = case(nx=20, ny=18, nz=16, pml=6, ntime=60)


def test_oracle_reproduces_golden():
    rows, subs = oracle_rows(C, DIVIDE)
    assert np.array_equal(np.array([(r, i) for r, i, _ in rows], np.int32), G["rows"])
    o = oracle_for(C)
    w = po.ricker(C["fm"], C["dt"], C["ntime"])
    seis = o.model_shot(o.node(*C["shots"][0]), w, [n for _, _, n in rows])
    assert np.array_equal(seis.astype(np.float32), G["seis"])
    assert np.array_equal(o.get(0, po.VZ).astype(np.float32), G["vz_last"])


@pytest.mark.gpu
@pytest.mark.parametrize("R", [1, 2, 4])
def test_engine_matches_golden(R):
    d = ddx.make_desc(C["nx"], C["ny"], C["nz"], C["dx"], C["pml"], C["dt"], C["fm"], C["ntime"], divide=DIVIDE,
                      rec_width=R, cfl_policy=1)
    with ddx.Engine(d) as e:
        e.set_model(C["vp"], C["vs"], C["rho"])
        e.set_wavelet(ddx.ricker(C["fm"], C["dt"], C["ntime"]))
        e.set_geometry(C["shots"], C["recv"])
        rank, inp = e.receivers()
        assert np.array_equal(np.stack([rank, inp], 1), G["rows"])
        e.model_shot(0)
        assert rel_l2(e.seismograms(0), G["seis"]) <= 1e-5
        assert rel_l2(e.get(0, ddx.VZ), G["vz_last"]) <= 1e-5
        e.rtm_shot(0)
        e.finalize_images()
        for r in range(4):
            assert rel_l2(e.image(ddx.IPP, r), G[f"ipp_R{R}_rank{r}"]) <= 1e-5
            assert rel_l2(e.image(ddx.IPP_LAP, r), G[f"ipp_lap_R{R}_rank{r}"]) <= 1e-5
