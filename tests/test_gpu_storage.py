"""GPU parity of the 16-bit wavefield storage formats (FP16 with
power-of-two class scales, BF16) against the oracle's FP32 mirror with the
same quantized stores, through the C ABI.

Bar: relative L2 <= 1e-5, as for FP32 storage.  The engine computes every
value in FP32 with the oracle's operation order and rounds each store to
nearest even, so equal FP32 results give equal 16-bit codes; the scales are
powers of two and fold into the constants exactly.  The error of the
16-bit results against the FP64 oracle is reported (and bounded) separately.
"""
import numpy as np
import pytest

import ddx
import pyoracle as po
from cases import case, oracle_for, oracle_rows, random_fields
from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5
FORMATS = [("fp16", po.STORE_FP16), ("bf16", po.STORE_BF16)]


def engine_for(c, storage, divide=(1, 1, 1), image_fp64=0, ring_storage="fp32"):
    d = ddx.make_desc(c["nx"], c["ny"], c["nz"], c["dx"], c["pml"], c["dt"], c["fm"], c["ntime"], divide=divide,
                      rec_width=c["rec_width"], image_fp64=image_fp64, cfl_policy=1, storage=storage,
                      ring_storage=ring_storage)
    e = ddx.Engine(d)
    e.set_model(c["vp"], c["vs"], c["rho"])
    e.set_wavelet(ddx.ricker(c["fm"], c["dt"], c["ntime"]))
    e.set_geometry(c["shots"], c["recv"])
    return e


def oracle_like(c, e, fmt):
    """FP32 oracle storing as the engine does (its current class scales)."""
    o = oracle_for(c)
    o.set_storage(fmt, e.storage_scale(ddx.CLASS_VELOCITY), e.storage_scale(ddx.CLASS_STRESS))
    return o


@pytest.mark.parametrize("name,fmt", FORMATS)
@pytest.mark.parametrize("divide", [(1, 1, 1), (2, 2, 1)])
@pytest.mark.parametrize("mode", [po.PURE_FORWARD, po.RTM_FORWARD, po.RTM_BACKWARD, po.RTM_REVERSE])
def test_half_steps_match_oracle(name, fmt, divide, mode):
    """Random initial fields (rounded to the format by set_field), three
    velocity+stress steps with CPML memory variables; every field compared."""
    c = case(nx=24, ny=22, nz=20, rec_width=2)
    e = engine_for(c, name, divide)
    if fmt == po.STORE_FP16:
        # random velocities and stresses of the same size drive the stresses
        # to about 2e4 within three steps (velocities stay under 6e-3): keep
        # both inside the FP16 range
        e.set_storage_scale(ddx.CLASS_VELOCITY, 17)
        e.set_storage_scale(ddx.CLASS_STRESS, -3)
    o = oracle_like(c, e, fmt)
    for f, arr in enumerate(random_fields(e.ext_shape, seed=200 + mode, scale=1e-3)):
        e.set(0, f, arr)
        o.set(0, f, arr.astype(np.float64))
    for _ in range(3):
        e.step_vel(0, mode)
        o.vel(0, mode)
        e.step_stress(0, mode)
        o.stress(0, mode)
    for f in range(13):
        if mode == po.PURE_FORWARD and f == ddx.SAU:
            continue  # the engine skips the unread CPML sau write of pure forward modelling
        err = rel_l2(e.get(0, f), o.get(0, f))
        assert err <= TOL, (name, ddx.FIELD_NAMES[f], err)


@pytest.mark.parametrize("name,fmt", FORMATS)
def test_field_round_trip(name, fmt):
    c = case(ntime=10)
    e = engine_for(c, name, (2, 2, 1))
    rng = np.random.default_rng(12)
    arr = (rng.standard_normal(e.ext_shape) * 1e-4).astype(np.float32)
    for f in (ddx.VX, ddx.SYZ, ddx.VZP):
        cls = ddx.CLASS_STRESS if ddx.SXX <= f <= ddx.SAU else ddx.CLASS_VELOCITY
        e.set(0, f, arr)
        q, _ = po.quantize(fmt, arr, e.storage_scale(cls))
        assert np.array_equal(e.get(0, f), q)


@pytest.mark.parametrize("name,fmt", FORMATS)
@pytest.mark.parametrize("divide", [(1, 1, 1), (2, 2, 1)])
def test_model_shot_seismograms(name, fmt, divide):
    """Seismograms with the default scales (FP16: from the wavelet peak and
    the model) against the storage oracle."""
    c = case(ntime=120)
    e = engine_for(c, name, divide)
    if fmt == po.STORE_BF16:
        assert e.storage_scale(ddx.CLASS_VELOCITY) == 0 and e.storage_scale(ddx.CLASS_STRESS) == 0
    rows, _ = oracle_rows(c, divide)
    o = oracle_like(c, e, fmt)
    nodes = [n for _, _, n in rows]
    w = po.ricker(c["fm"], c["dt"], c["ntime"])
    e.model_shot(0)
    g = e.seismograms(0)
    ref = o.model_shot(o.node(*c["shots"][0]), w, nodes)
    assert np.isfinite(g).all()
    err = rel_l2(g, ref)
    assert err <= TOL, (name, divide, err)
    assert rel_l2(e.get(0, ddx.SZZ), o.get(0, po.SZZ)) <= TOL


@pytest.mark.parametrize("name", ["fp16", "bf16"])
def test_decomposition_is_exact(name):
    c = case(nx=30, ny=26, nz=24, ntime=80)
    outs = []
    for divide in [(1, 1, 1), (2, 2, 1), (3, 2, 2)]:
        e = engine_for(c, name, divide)
        e.model_shot(0)
        rank, inp = e.receivers()
        order = np.argsort(inp, kind="stable")
        outs.append((e.seismograms(0)[:, order], e.get(0, ddx.VZ), e.get(0, ddx.SXY)))
        e.close()
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("name,fmt", FORMATS)
@pytest.mark.parametrize("rec_width", [1, 4])
def test_rtm_images_and_snapshots(name, fmt, rec_width):
    """RTM shot (VOWR store in the 16-bit format): images, and forward and
    reconstructed vx snapshots, against the storage oracle."""
    c = case(ntime=70, rec_width=rec_width)
    divide = (2, 2, 1)
    e = engine_for(c, name, divide)
    rows, subs = oracle_rows(c, divide)
    o = oracle_like(c, e, fmt)
    w = po.ricker(c["fm"], c["dt"], c["ntime"])
    nodes = [n for _, _, n in rows]
    shot = o.node(*c["shots"][0])
    obs = o.model_shot(shot, w, nodes).astype(np.float32)
    every = 23
    e.set_observed(0, obs)
    e.set_snapshots(ddx.VX, every)
    e.rtm_shot(0)
    fs, bs = o.rtm_shot(shot, w, nodes, obs, snap_every=every)
    e.finalize_images()
    m = c["pml"]
    crop = (slice(m, m + c["nz"]), slice(m, m + c["ny"]), slice(m, m + c["nx"]))
    for q in range(fs.shape[0]):
        assert rel_l2(e.snapshot(0, q), fs[q][crop]) <= TOL, ("forward snapshot", q)
        if q > 0:
            assert rel_l2(e.snapshot(1, q), bs[q][crop]) <= TOL, ("backward snapshot", q)
    for s in subs:
        i0, j0, k0 = s["phy_off"]
        ni, nj, nk = s["phy_n"]
        img, lap = o.finalize_box(i0, j0, k0, ni, nj, nk)
        ip_ref = o.image(0)[k0:k0 + nk, j0:j0 + nj, i0:i0 + ni]
        assert rel_l2(e.image(ddx.IP, s["rank"]), ip_ref) <= TOL
        for q in range(4):
            assert rel_l2(e.image(ddx.IPP + q, s["rank"]), img[q]) <= TOL, ("img", q)
            assert rel_l2(e.image(ddx.IPP_LAP + q, s["rank"]), lap[q]) <= TOL, ("lap", q)


@pytest.mark.parametrize("name,fmt", FORMATS)
@pytest.mark.parametrize("rec_width", [1, 4])
def test_ring_storage_rtm_matches_oracle(name, fmt, rec_width):
    """FP32 wavefields with the VOWR store alone in a 16-bit format: every
    saved boundary value is scaled, rounded and widened back on injection;
    images and reconstructed snapshots against the oracle with the same
    rounding of its store."""
    c = case(ntime=70, rec_width=rec_width)
    divide = (2, 2, 1)
    e = engine_for(c, "fp32", divide, ring_storage=name)
    lv, ls = e.storage_scale(ddx.CLASS_VELOCITY), e.storage_scale(ddx.CLASS_STRESS)
    if fmt == po.STORE_FP16:
        assert lv > ls  # auto scales: velocities are about rho vp below the stresses
    else:
        assert lv == 0 and ls == 0
    rows, subs = oracle_rows(c, divide)
    o = oracle_for(c)
    o.set_ring_storage(fmt, lv, ls)
    w = po.ricker(c["fm"], c["dt"], c["ntime"])
    nodes = [n for _, _, n in rows]
    shot = o.node(*c["shots"][0])
    obs = o.model_shot(shot, w, nodes).astype(np.float32)
    every = 23
    e.set_observed(0, obs)
    e.set_snapshots(ddx.VX, every)
    e.rtm_shot(0)
    fs, bs = o.rtm_shot(shot, w, nodes, obs, snap_every=every)
    e.finalize_images()
    m = c["pml"]
    crop = (slice(m, m + c["nz"]), slice(m, m + c["ny"]), slice(m, m + c["nx"]))
    for q in range(1, fs.shape[0]):
        assert rel_l2(e.snapshot(1, q), bs[q][crop]) <= TOL, ("backward snapshot", q)
    for s in subs:
        i0, j0, k0 = s["phy_off"]
        ni, nj, nk = s["phy_n"]
        img, lap = o.finalize_box(i0, j0, k0, ni, nj, nk)
        assert rel_l2(e.image(ddx.IP, s["rank"]), o.image(0)[k0:k0 + nk, j0:j0 + nj, i0:i0 + ni]) <= TOL
        for q in range(4):
            assert rel_l2(e.image(ddx.IPP + q, s["rank"]), img[q]) <= TOL, ("img", q)
    e.close()


def test_ring_storage_rounding_reaches_the_images():
    """The 16-bit store changes the images (the rounding is not lost on the
    way), by far less than 16-bit wavefields do."""
    c = case(ntime=70)
    rows, _ = oracle_rows(c, (1, 1, 1))
    o = oracle_for(c)
    w = po.ricker(c["fm"], c["dt"], c["ntime"])
    obs = o.model_shot(o.node(*c["shots"][0]), w, [n for _, _, n in rows]).astype(np.float32)
    ims = {}
    for key, st, ring in (("fp32", "fp32", "fp32"), ("ring", "fp32", "bf16"), ("bf16", "bf16", "fp32")):
        e = engine_for(c, st, ring_storage=ring)
        e.set_observed(0, obs)
        e.rtm_shot(0)
        e.finalize_images()
        ims[key] = e.image(ddx.IPP, 0)
        e.close()
    d_ring = rel_l2(ims["ring"], ims["fp32"])
    d_wave = rel_l2(ims["bf16"], ims["fp32"])
    print("image change: bf16 store", d_ring, "bf16 wavefields", d_wave)
    assert 0 < d_ring < d_wave


def test_error_against_fp64_oracle():
    """What the 16-bit formats cost in accuracy on this case (120 steps,
    default scales): seismogram relative L2 against the FP64 oracle."""
    c = case(ntime=120)
    rows, _ = oracle_rows(c, (1, 1, 1))
    nodes = [n for _, _, n in rows]
    o64 = oracle_for(c, prec=8)
    ref = o64.model_shot(o64.node(*c["shots"][0]), po.ricker(c["fm"], c["dt"], c["ntime"]), nodes)
    errs = {}
    for name in ("fp32", "fp16", "bf16"):
        e = engine_for(c, name)
        e.model_shot(0)
        errs[name] = rel_l2(e.seismograms(0), ref)
        e.close()
    print("seismogram rel L2 vs FP64:", errs)
    assert errs["fp32"] < 1e-5
    assert errs["fp32"] < errs["fp16"] < errs["bf16"]
    assert errs["fp16"] < 1e-2 and errs["bf16"] < 5e-2


def test_unsupported_combinations_raise():
    c = case(ntime=10)
    e = engine_for(c, "fp16")
    with pytest.raises(ddx.DdxError, match="FWI"):
        e.step_vel(0, po.FWI)
    with pytest.raises(ddx.DdxError):
        e.set_tiling(64, 8, 0, 4)  # LDS-tiled kernels are FP32 only
    with pytest.raises(ddx.DdxError):
        e.set_tiling(64, 4, 1, 0)  # one plane per block: FP32 only
    e.set_tiling(64, 2, 1, 8)
    f = engine_for(c, "fp32")
    with pytest.raises(ddx.DdxError):
        f.set_storage_scale(ddx.CLASS_STRESS, 3)
    g = engine_for(c, "fp32", ring_storage="fp16")
    g.set_storage_scale(ddx.CLASS_STRESS, 3)  # the store's scale
    assert g.storage_scale(ddx.CLASS_STRESS) == 3
    with pytest.raises(ddx.DdxError, match="ring_storage"):
        engine_for(c, "fp16", ring_storage="bf16")
    with pytest.raises(ddx.DdxError, match="ring_storage"):
        engine_for(c, "fp16", ring_storage="bfp16")
    with pytest.raises(ddx.DdxError, match="storage"):
        engine_for(c, "bfp16")  # a store format only
    h = engine_for(c, "fp32", ring_storage="bfp16")
    with pytest.raises(ddx.DdxError):
        h.set_storage_scale(ddx.CLASS_STRESS, 3)  # runs carry their own exponents
    with pytest.raises(ddx.DdxError):
        h.set_ring_scaling(1)


@pytest.mark.parametrize("divide", [(1, 1, 1), (2, 2, 1)])
def test_bfp16_store_round_trip(divide):
    """The block-scaled store through the step hooks: a slice saved, the
    fields cleared, the slice injected.  Every injected value v' of an
    original v keeps |v' - v| <= 2^-13 |v| + 2^(s-20), where 2^s <= twice
    the largest |value| of v's run, so <= 2^-13 |v| + 2^-19 times the
    field's largest ring |value| (a run's exponent comes from its own 16
    entries, and no run mixes boxes: the velocities are 3e9 times smaller
    than the stresses here, so a run shared with a stress box would lose
    them); and the rounding is there."""
    c = case(ntime=10, rec_width=2)
    e = engine_for(c, "fp32", divide, ring_storage="bfp16")
    ring_fields = [ddx.VX, ddx.VY, ddx.VZ, ddx.SXX, ddx.SYY, ddx.SZZ, ddx.SAU, ddx.SXY, ddx.SXZ, ddx.SYZ]
    for f, arr in enumerate(random_fields(e.ext_shape, seed=5, scale=1e-3)):
        e.set(0, f, arr if f <= ddx.VZ else arr * np.float32(3e6))
    before = {f: e.get(0, f) for f in ring_fields}
    e.vowr_save(0)
    zero = np.zeros(e.ext_shape, np.float32)
    for f in ring_fields:
        e.set(0, f, zero)
    e.vowr_inject(1, 0)
    e.vowr_inject(2, 0)
    for f in ring_fields:
        got = e.get(0, f)
        cells = got != 0
        assert cells.sum() > 0, ddx.FIELD_NAMES[f]
        ref = before[f][cells]
        err = np.abs(got[cells].astype(np.float64) - ref)
        bound = np.abs(ref) * 2.0 ** -13 + np.abs(ref).max() * 2.0 ** -19
        assert np.all(err <= bound), (ddx.FIELD_NAMES[f], (err / bound).max())
        assert err.max() > 0, ddx.FIELD_NAMES[f]
    e.close()


@pytest.mark.parametrize("divide", [(1, 1, 1), (2, 2, 1), (3, 2, 2)])
@pytest.mark.parametrize("rec_width", [1, 4])
def test_bfp16_store_rtm_matches_oracle(divide, rec_width):
    """FP32 wavefields with the block-scaled store: images and reconstructed
    snapshots against the oracle whose store rounds the same runs (each
    subdomain's part of a ring box, runs of 16 from its origin), at the 1e-5
    bar; the decompositions split the ring boxes along one, two and three
    axes (a 36x30x28 grid, so that 3x2x2 subdomains hold 8 cells per axis)."""
    c = case(nx=36, ny=30, nz=28, ntime=70, rec_width=rec_width)
    e = engine_for(c, "fp32", divide, ring_storage="bfp16")
    rows, subs = oracle_rows(c, divide)
    o = oracle_for(c)
    w = po.ricker(c["fm"], c["dt"], c["ntime"])
    nodes = [n for _, _, n in rows]
    shot = o.node(*c["shots"][0])
    obs = o.model_shot(shot, w, nodes).astype(np.float32)
    o.set_ring_bfp(subs)
    every = 23
    e.set_observed(0, obs)
    e.set_snapshots(ddx.VX, every)
    e.rtm_shot(0)
    fs, bs = o.rtm_shot(shot, w, nodes, obs, snap_every=every)
    e.finalize_images()
    m = c["pml"]
    crop = (slice(m, m + c["nz"]), slice(m, m + c["ny"]), slice(m, m + c["nx"]))
    for q in range(1, fs.shape[0]):
        assert rel_l2(e.snapshot(1, q), bs[q][crop]) <= TOL, ("backward snapshot", q)
    for s in subs:
        i0, j0, k0 = s["phy_off"]
        ni, nj, nk = s["phy_n"]
        img, lap = o.finalize_box(i0, j0, k0, ni, nj, nk)
        assert rel_l2(e.image(ddx.IP, s["rank"]), o.image(0)[k0:k0 + nk, j0:j0 + nj, i0:i0 + ni]) <= TOL
        for q in range(4):
            assert rel_l2(e.image(ddx.IPP + q, s["rank"]), img[q]) <= TOL, ("img", q)
    e.close()


@pytest.mark.parametrize("rec_width", [1, 4])
def test_bfp16_store_rtm_images(rec_width):
    """RTM with FP32 wavefields and the block-scaled store, and with the FP16
    store, both against the FP32 oracle without store rounding: the
    reconstructed snapshots and the final images of the BFP16 store are
    closer to it than the FP16 store's (its rounding is 2^-13 of each value
    within 2^-7 of its run's largest, FP16's 2^-11), by at least 2x at R = 1
    (measured about 4x and 3.3x); and its raw images closer to the
    FP32-store run.  At R = 4 the FP32 reconstruction is exact and the
    reconstructed source field decays by more than two orders of magnitude
    toward t = 0, so the stores' absolute rounding dominates the last
    snapshot and the compensated images in relative terms (about 12 and 50
    for the snapshot); the CPU oracle's mirrors of both stores show the same
    smooth growth.  All distances are printed."""
    c = case(ntime=70, rec_width=rec_width)
    divide = (2, 2, 1)
    rows, subs = oracle_rows(c, divide)
    o = oracle_for(c)
    w = po.ricker(c["fm"], c["dt"], c["ntime"])
    nodes = [n for _, _, n in rows]
    shot = o.node(*c["shots"][0])
    obs = o.model_shot(shot, w, nodes).astype(np.float32)
    every = 23
    fs, bs = o.rtm_shot(shot, w, nodes, obs, snap_every=every)
    m = c["pml"]
    crop = (slice(m, m + c["nz"]), slice(m, m + c["ny"]), slice(m, m + c["nx"]))
    want = []
    for s in subs:
        i0, j0, k0 = s["phy_off"]
        ni, nj, nk = s["phy_n"]
        want.append(o.finalize_box(i0, j0, k0, ni, nj, nk)[0])
    raw, snap, fin = {}, {}, {}
    for name in ("fp32", "fp16", "bfp16"):
        e = engine_for(c, "fp32", divide, ring_storage=name)
        e.set_observed(0, obs)
        e.set_snapshots(ddx.VX, every)
        e.rtm_shot(0)
        raw[name] = [e.image(ddx.IP, s["rank"]) for s in subs] + [e.image(ddx.IPP, s["rank"]) for s in subs]
        snap[name] = [rel_l2(e.snapshot(1, q), bs[q][crop]) for q in range(1, fs.shape[0])]
        e.finalize_images()
        fin[name] = max(rel_l2(e.image(ddx.IPP + q, s["rank"]), want[n][q])
                        for n, s in enumerate(subs) for q in range(4))
        e.close()
    d16 = max(rel_l2(a, b) for a, b in zip(raw["fp16"], raw["fp32"]))
    dp = max(rel_l2(a, b) for a, b in zip(raw["bfp16"], raw["fp32"]))
    print(f"raw Ip/Ipp change against the FP32 store: FP16 store {d16:.2e}, BFP16 store {dp:.2e}; "
          f"snapshots against the oracle: FP32 {snap['fp32']}, FP16 {snap['fp16']}, BFP16 {snap['bfp16']}; "
          f"final images: FP32 {fin['fp32']:.2e}, FP16 {fin['fp16']:.2e}, BFP16 {fin['bfp16']:.2e}", flush=True)
    assert fin["fp32"] <= TOL and max(snap["fp32"]) <= TOL
    k = 2 if rec_width == 1 else 1
    assert 0 < dp < d16
    assert fin["bfp16"] * k < fin["fp16"]
    for a, b in zip(snap["bfp16"], snap["fp16"]):
        assert a * k < b


@pytest.mark.parametrize("name,fmt", FORMATS)
@pytest.mark.parametrize("rec_width", [1, 4])
def test_per_slice_ring_scaling_matches_oracle(name, fmt, rec_width):
    """ddx_set_ring_scaling(1): one power-of-two scale per ring field and
    time slice; images and reconstructed snapshots against the oracle's
    mirror of the same scaling.  FP16: the per-slice scales do not depend
    on a class scale, so they stay close to the FP32 store where a poor
    class scale (2^-20, most velocities below FP16's range) does not; the
    distance with the default class scales is printed beside them."""
    c = case(ntime=70, rec_width=rec_width)
    divide = (2, 2, 1)
    rows, subs = oracle_rows(c, divide)
    o = oracle_for(c)
    w = po.ricker(c["fm"], c["dt"], c["ntime"])
    nodes = [n for _, _, n in rows]
    shot = o.node(*c["shots"][0])
    obs = o.model_shot(shot, w, nodes).astype(np.float32)
    o.set_ring_storage(fmt, 0, 0)
    o.set_ring_dynamic(True)
    every = 23
    e = engine_for(c, "fp32", divide, ring_storage=name)
    e.set_ring_scaling(1)
    e.set_observed(0, obs)
    e.set_snapshots(ddx.VX, every)
    e.rtm_shot(0)
    raw = e.image(ddx.IPP, 0)  # before the illumination compensation
    e.finalize_images()
    fs, bs = o.rtm_shot(shot, w, nodes, obs, snap_every=every)
    m = c["pml"]
    crop = (slice(m, m + c["nz"]), slice(m, m + c["ny"]), slice(m, m + c["nx"]))
    for q in range(1, fs.shape[0]):
        assert rel_l2(e.snapshot(1, q), bs[q][crop]) <= TOL, ("backward snapshot", q)
    for s in subs:
        i0, j0, k0 = s["phy_off"]
        ni, nj, nk = s["phy_n"]
        img, lap = o.finalize_box(i0, j0, k0, ni, nj, nk)
        for q in range(4):
            assert rel_l2(e.image(ddx.IPP + q, s["rank"]), img[q]) <= TOL, ("img", q)
    if fmt == po.STORE_FP16:
        # against the FP32 store: per-slice scales, default class scales and
        # class scales of 2^-20
        ref = engine_for(c, "fp32", divide)
        cls = engine_for(c, "fp32", divide, ring_storage=name)
        poor = engine_for(c, "fp32", divide, ring_storage=name)
        poor.set_storage_scale(ddx.CLASS_VELOCITY, -20)
        poor.set_storage_scale(ddx.CLASS_STRESS, -20)
        out = []
        for x in (ref, cls, poor):
            x.set_observed(0, obs)
            x.rtm_shot(0)
            out.append(x.image(ddx.IPP, 0))  # raw: the compensation divides by Ip, tiny far from the source
            x.close()
        d_dyn = rel_l2(raw, out[0])
        d_cls = rel_l2(out[1], out[0])
        d_poor = rel_l2(out[2], out[0])
        print(f"raw Ipp change against the FP32 store: per-slice {d_dyn:.2e}, default class scales {d_cls:.2e}, "
              f"class scales 2^-20 {d_poor:.2e}", flush=True)
        assert d_dyn < d_poor
    e.close()
