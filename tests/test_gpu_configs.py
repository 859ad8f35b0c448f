"""The five workload configurations at their own grid sizes, on the GPU,
against the CPU oracle (through the C ABI).

Bar: relative L2 error <= 1e-5 against the FP32 oracle (16-bit storage:
against the oracle's mirror with the same quantized stores); the error
against the FP64 oracle is printed and held to the 1e-3 target.  Runs are
shortened in time where the oracle would take minutes; the grids,
decompositions, receiver layouts and ring widths are the configurations'.

- modelling, 100^3 + 10 CPML layers, 1000 steps, 2 shots, 8 receivers:
  explosive and vertical-stress sources, in full;
- RTM, 256x256x200 + 20 layers, 2x2x1 subdomains, 128x128 receivers,
  R = 1, 2, 4, 150 steps: seismograms and all nine images per subdomain;
- batching: the four smallest of the 64 uneven shots, fused in one batch,
  100 steps;
- precision: the 200^3 crop of the 676x676x210 model, FP32, FP16 and BF16
  storage (seismograms) and the FP32 RTM with FP64 images, 140 steps;
- serving: requests of the serving distribution (grids 64-400, both kinds,
  FP32 and BF16) are covered in tests/test_serve.py at reduced sizes.
"""
import numpy as np
import pytest

import ddx
import pyoracle as po
from conftest import rel_l2
from ddx import synth

pytestmark = pytest.mark.gpu
TOL = 1e-5


def say(msg):
    print(msg, flush=True)


def engine(c, ntime, divide=(1, 1, 1), rec_width=1, image_fp64=0, source_mode=0, storage="fp32"):
    d = ddx.make_desc(c["nx"], c["ny"], c["nz"], c["dx"], c["pml"], c["dt"], c["fm"], ntime, divide=divide,
                      rec_width=rec_width, image_fp64=image_fp64, cfl_policy=1, source_mode=source_mode,
                      storage=storage)
    e = ddx.Engine(d)
    e.set_model(c["vp"], c["vs"], c["rho"])
    e.set_wavelet(ddx.ricker(c["fm"], c["dt"], ntime))
    e.set_geometry(c["shots"], c["recv"])
    return e


def oracle(c, ntime, rec_width=1, prec=4):
    return po.Oracle(c["nx"], c["ny"], c["nz"], c["dx"], c["pml"], c["dt"], c["fm"], ntime, rec_width, c["vp"],
                     c["vs"], c["rho"], prec=prec)


def rows(c, divide):
    d = (c["dx"],) * 3
    subs = po.decompose((c["nx"], c["ny"], c["nz"]), d, c["pml"], divide)
    kept = po.kept_receivers(subs, c["recv"], d, c["pml"])
    NX, NY = c["nx"] + 2 * c["pml"], c["ny"] + 2 * c["pml"]
    return [po.ext_index(n, NX, NY) for _, _, n in kept], subs


# ---------------------------------------------------------------------------
# modelling, 100^3, in full
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("source_mode", [0, 1])
def test_config1_modelling_full(source_mode):
    c = synth.config1()
    nt = c["ntime"]
    say(f"config 1: 100^3 + 10, {nt} steps, 2 shots, 8 receivers, source_mode {source_mode}")
    e = engine(c, nt, source_mode=source_mode)
    e.autotune()
    nodes, _ = rows(c, (1, 1, 1))
    assert len(nodes) == 8
    o = oracle(c, nt)
    w = po.ricker(c["fm"], c["dt"], nt)
    shots = range(2) if source_mode == 0 else range(1)
    for s in shots:
        e.model_shot(s)
        got = e.seismograms(s)
        ref = o.model_shot(o.node(*c["shots"][s]), w, nodes, src_mode=source_mode)
        err = rel_l2(got, ref)
        say(f"  shot {s}: rel L2 vs FP32 oracle {err:.2e}")
        assert err <= TOL, (s, err)
        if s == 0:
            o8 = oracle(c, nt, prec=8)
            ref8 = o8.model_shot(o8.node(*c["shots"][s]), w, nodes, src_mode=source_mode)
            e8 = rel_l2(got, ref8)
            say(f"  shot {s}: rel L2 vs FP64 oracle {e8:.2e}")
            assert e8 <= 1e-3
            o8.close()
    e.close()


# ---------------------------------------------------------------------------
# RTM, 256x256x200, 2x2x1, 128x128 receivers
# ---------------------------------------------------------------------------
C2_STEPS = 150


@pytest.fixture(scope="module")
def c2_observed():
    """The configuration and its FP32-oracle seismograms over C2_STEPS (the
    observed data of every RTM run below)."""
    c = synth.config2(1)
    nodes, subs = rows(c, (2, 2, 1))
    o = oracle(c, C2_STEPS)
    w = po.ricker(c["fm"], c["dt"], C2_STEPS)
    obs = o.model_shot(o.node(*c["shots"][0]), w, nodes).astype(np.float32)
    o.close()
    return c, nodes, subs, obs


def test_config2_seismograms(c2_observed):
    c, nodes, _, obs = c2_observed
    say(f"config 2: 256x256x200 + 20, 2x2x1, {len(nodes)} receivers, {C2_STEPS} steps: modelling")
    e = engine(c, C2_STEPS, divide=(2, 2, 1))
    e.autotune()
    assert e.nrec == len(nodes) == 16384
    e.model_shot(0)
    err = rel_l2(e.seismograms(0), obs)
    say(f"  rel L2 vs FP32 oracle {err:.2e}")
    assert err <= TOL
    e.close()


# brick forms: R = 1 autotuned; R = 2 and R = 4 the LDS-tile bricks (form 4)
# of two shapes for every kernel and mode
C2_BRICKS = {1: None, 2: (64, 4, 1, 4), 4: (64, 8, 1, 2)}


@pytest.mark.parametrize("rec_width", [1, 2, 4])
def test_config2_rtm_images(c2_observed, rec_width):
    c, nodes, subs, obs = c2_observed
    say(f"config 2: RTM R = {rec_width}, {C2_STEPS} steps")
    e = engine(c, C2_STEPS, divide=(2, 2, 1), rec_width=rec_width)
    shape = C2_BRICKS[rec_width]
    if shape is None:
        e.autotune()
    else:
        for kernel in (0, 1):
            for mode in range(4):
                e.set_brick(kernel, mode, *shape, px=4)
    e.set_observed(0, obs)
    e.rtm_shot(0)
    e.finalize_images()
    o = oracle(c, C2_STEPS, rec_width=rec_width)
    w = po.ricker(c["fm"], c["dt"], C2_STEPS)
    o.rtm_shot(o.node(*c["shots"][0]), w, nodes, obs)
    worst = 0.0
    for s in subs:
        i0, j0, k0 = s["phy_off"]
        ni, nj, nk = s["phy_n"]
        img, lap = o.finalize_box(i0, j0, k0, ni, nj, nk)
        ip_ref = o.image(0)[k0:k0 + nk, j0:j0 + nj, i0:i0 + ni]
        errs = [rel_l2(e.image(ddx.IP, s["rank"]), ip_ref)]
        errs += [rel_l2(e.image(ddx.IPP + q, s["rank"]), img[q]) for q in range(4)]
        errs += [rel_l2(e.image(ddx.IPP_LAP + q, s["rank"]), lap[q]) for q in range(4)]
        worst = max(worst, max(errs))
        assert max(errs) <= TOL, (s["rank"], errs)
    say(f"  9 images x 4 subdomains: worst rel L2 vs FP32 oracle {worst:.2e}")
    o.close()
    e.close()


# ---------------------------------------------------------------------------
# batching: uneven grids in one fused batch
# ---------------------------------------------------------------------------
def test_config3_fused_batch_smallest_four():
    shots = synth.config3()
    shots.sort(key=lambda c: (c["nx"] + 2 * c["pml"]) * (c["ny"] + 2 * c["pml"]) * (c["nz"] + 2 * c["pml"]))
    four = shots[:4]
    nt = 100
    say("config 3: fused batch of " + ", ".join(f"{c['nx']}x{c['ny']}x{c['nz']}+{c['pml']}" for c in four) +
        f", {nt} steps")
    es = [engine(c, nt) for c in four]
    es[0].autotune()
    ddx.model_batch(es)
    for c, e in zip(four, es):
        nodes, _ = rows(c, (1, 1, 1))
        o = oracle(c, nt)
        ref = o.model_shot(o.node(*c["shots"][0]), po.ricker(c["fm"], c["dt"], nt), nodes)
        err = rel_l2(e.seismograms(0), ref)
        say(f"  {c['nx']}x{c['ny']}x{c['nz']}: rel L2 vs FP32 oracle {err:.2e}")
        assert err <= TOL
        o.close()
        e.close()


# ---------------------------------------------------------------------------
# precision: the 200^3 crop
# ---------------------------------------------------------------------------
C4_STEPS = 140


@pytest.fixture(scope="module")
def c4_crop():
    return synth.config4(crop=200)


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
def test_config4_crop_storage_formats(c4_crop, storage):
    c = c4_crop
    say(f"config 4: 200^3 crop + 20, {storage} storage, {C4_STEPS} steps: modelling")
    e = engine(c, C4_STEPS, storage=storage)
    nodes, _ = rows(c, (1, 1, 1))
    o = oracle(c, C4_STEPS)
    if storage != "fp32":
        o.set_storage(ddx.storage_id(storage), e.storage_scale(0), e.storage_scale(1))
    e.model_shot(0)
    got = e.seismograms(0)
    ref = o.model_shot(o.node(*c["shots"][0]), po.ricker(c["fm"], c["dt"], C4_STEPS), nodes)
    err = rel_l2(got, ref)
    say(f"  rel L2 vs the oracle with the same stores {err:.2e}")
    assert err <= TOL
    o.close()
    e.close()


def test_config4_crop_rtm_fp64_images(c4_crop):
    c = c4_crop
    say(f"config 4: 200^3 crop + 20, FP32 RTM with FP64 images, {C4_STEPS} steps")
    e = engine(c, C4_STEPS, image_fp64=1)
    nodes, subs = rows(c, (1, 1, 1))
    w = po.ricker(c["fm"], c["dt"], C4_STEPS)
    o = oracle(c, C4_STEPS)
    obs = o.model_shot(o.node(*c["shots"][0]), w, nodes).astype(np.float32)
    e.set_observed(0, obs)
    e.rtm_shot(0)
    e.finalize_images()
    o.rtm_shot(o.node(*c["shots"][0]), w, nodes, obs)
    img, lap = o.finalize_box(0, 0, 0, c["nx"], c["ny"], c["nz"])
    errs = [rel_l2(e.image(ddx.IP, 0), o.image(0))]
    errs += [rel_l2(e.image(ddx.IPP + q, 0), img[q]) for q in range(4)]
    errs += [rel_l2(e.image(ddx.IPP_LAP + q, 0), lap[q]) for q in range(4)]
    say(f"  9 images: worst rel L2 vs FP32 oracle {max(errs):.2e}")
    assert max(errs) <= TOL, errs
    o.close()
    e.close()
