"""16-bit wavefield storage on the CPU: the oracle's conversions against
numpy's float16 and torch's bfloat16 casts (both round to nearest, ties to
even), the scaling identity, the storage oracle's behaviour against the FP64
oracle, and the C ABI's footprint for the 16-bit formats.

The 16-bit formats have no counterpart in the reference (it stores FP32);
the oracle's FP32 mirror with quantized stores is their parity checker.
"""
import numpy as np
import pytest
import torch

import ddx
import pyoracle as po
from cases import case, oracle_for, oracle_rows
from conftest import rel_l2


def _samples(n=400_000, seed=3):
    """Random finite float32 bit patterns (all exponents), decimal-scaled
    normals, and the edge cases of both formats."""
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[np.isfinite(x)]
    edges = np.array([0.0, -0.0, 1.0, -1.0, 65504.0, 65519.996, 65520.0, 65536.0, 2.0 ** -14, 2.0 ** -24,
                      2.0 ** -25, np.nextafter(np.float32(2.0 ** -25), np.float32(1)), 2.0 ** -26, 1e-30,
                      1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8,
                      3.3895314e38, -3.3895314e38], np.float32)
    ties16 = (np.arange(1, 4096, dtype=np.float32) + np.float32(0.5)) * np.float32(2.0 ** -24)  # subnormal ties
    scaled = (rng.standard_normal(100_000) * 10.0 ** rng.integers(-12, 7, 100_000)).astype(np.float32)
    return np.concatenate([x, edges, ties16, scaled])


def test_f16_codes_match_numpy():
    x = _samples()
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16)
    val, code = po.quantize(po.STORE_FP16, x)
    assert np.array_equal(code, ref.view(np.uint16))
    fin = np.isfinite(ref)
    assert np.array_equal(val[fin], ref[fin].astype(np.float32))


def test_bf16_codes_match_torch():
    x = _samples(seed=4)
    t = torch.from_numpy(x.copy()).to(torch.bfloat16)
    val, code = po.quantize(po.STORE_BF16, x)
    assert np.array_equal(code, t.view(torch.int16).numpy().view(np.uint16))
    assert np.array_equal(val, t.float().numpy())


@pytest.mark.parametrize("fmt", [po.STORE_FP16, po.STORE_BF16])
def test_power_of_two_scale_is_exact(fmt):
    """Storing x at scale 2^k holds exactly the value x * 2^k holds at scale
    1, divided by 2^k (the identity the engine's folded scale ratios rely on)."""
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(50_000) * 1e-3).astype(np.float32)
    for k in (-7, 0, 9, 17):
        a, _ = po.quantize(fmt, x, k)
        b, _ = po.quantize(fmt, (x * np.float32(2.0 ** k)).astype(np.float32), 0)
        assert np.array_equal(a, (b / np.float32(2.0 ** k)).astype(np.float32))
        # and the stored value is within half a unit in the last place
        ulp = 2.0 ** -11 if fmt == po.STORE_FP16 else 2.0 ** -8
        big = np.abs(x * 2.0 ** k) > 2.0 ** -14
        assert np.all(np.abs(a - x)[big] <= ulp * np.abs(x)[big])


def test_fp32_storage_mode_is_the_fp32_mirror():
    c = case(ntime=50)
    rows, _ = oracle_rows(c, (1, 1, 1))
    nodes = [n for _, _, n in rows]
    w = po.ricker(c["fm"], c["dt"], c["ntime"])
    a = oracle_for(c)
    b = oracle_for(c)
    b.set_storage(po.STORE_FP32, 5, -3)  # scales mean nothing for FP32
    shot = a.node(*c["shots"][0])
    assert np.array_equal(a.model_shot(shot, w, nodes), b.model_shot(shot, w, nodes))


def test_storage_error_against_fp64_oracle():
    """Quantized stores perturb the result by about the format's unit
    roundoff, growing with the steps: FP32 < FP16 < BF16 against FP64, all
    small on this case (120 steps)."""
    c = case(ntime=120)
    rows, _ = oracle_rows(c, (1, 1, 1))
    nodes = [n for _, _, n in rows]
    w = po.ricker(c["fm"], c["dt"], c["ntime"])
    o64 = oracle_for(c, prec=8)
    shot = o64.node(*c["shots"][0])
    ref = o64.model_shot(shot, w, nodes)
    errs = {}
    for fmt, lv, ls in ((po.STORE_FP32, 0, 0), (po.STORE_FP16, 20, -6), (po.STORE_BF16, 0, 0)):
        o = oracle_for(c)
        o.set_storage(fmt, lv, ls)
        errs[fmt] = rel_l2(o.model_shot(shot, w, nodes), ref)
    assert errs[po.STORE_FP32] < 1e-5
    assert errs[po.STORE_FP32] < errs[po.STORE_FP16] < errs[po.STORE_BF16]
    assert errs[po.STORE_FP16] < 5e-3 and errs[po.STORE_BF16] < 5e-2, errs


def test_set_field_rounds_by_class():
    c = case(ntime=10)
    o = oracle_for(c)
    o.set_storage(po.STORE_FP16, 12, -4)
    rng = np.random.default_rng(8)
    arr = rng.standard_normal(o.ext_shape) * 1e-2
    for f, lg in ((po.VX, 12), (po.VZP, 12), (po.SXY, -4), (po.SAU, -4)):
        o.set(0, f, arr)
        q, _ = po.quantize(po.STORE_FP16, arr.astype(np.float32), lg)
        assert np.array_equal(o.get(0, f), q.astype(np.float64))


def test_estimate_bytes_by_storage():
    """16-bit storage halves the wavefields, memory variables and the VOWR
    store; the model, CPML coefficients, images and seismograms stay."""
    kw = dict(divide=(2, 2, 1), rec_width=2)
    d32 = ddx.make_desc(64, 60, 48, 10.0, 10, 0.001, 15.0, 300, **kw)
    d16 = ddx.make_desc(64, 60, 48, 10.0, 10, 0.001, 15.0, 300, storage="fp16", **kw)
    db16 = ddx.make_desc(64, 60, 48, 10.0, 10, 0.001, 15.0, 300, storage="bf16", **kw)
    ring32 = ddx.estimate_bytes(d32, 100, rtm=True) - ddx.estimate_bytes(d32, 100, rtm=False)
    ring16 = ddx.estimate_bytes(d16, 100, rtm=True) - ddx.estimate_bytes(d16, 100, rtm=False)
    assert ring16 * 2 == ring32
    # a 16-bit store over FP32 wavefields: only the store shrinks
    dr = ddx.make_desc(64, 60, 48, 10.0, 10, 0.001, 15.0, 300, ring_storage="bf16", **kw)
    assert ddx.estimate_bytes(dr, 100, rtm=False) == ddx.estimate_bytes(d32, 100, rtm=False)
    assert ddx.estimate_bytes(dr, 100, rtm=True) - ddx.estimate_bytes(dr, 100, rtm=False) == ring16
    assert ddx.estimate_bytes(d16, 100) == ddx.estimate_bytes(db16, 100)
    assert 0.5 < ddx.estimate_bytes(d16, 100) / ddx.estimate_bytes(d32, 100) < 0.75
    assert ddx.storage_id("bf16") == ddx.STORE_BF16 and ddx.storage_id(1) == ddx.STORE_FP16
    with pytest.raises(ValueError):
        ddx.storage_id("fp8")
    # block-scaled store: 16-bit entries, each box padded to whole runs of
    # 16, plus one exponent byte per run
    dp = ddx.make_desc(64, 60, 48, 10.0, 10, 0.001, 15.0, 300, ring_storage="bfp16", **kw)
    ringp = ddx.estimate_bytes(dp, 100, rtm=True) - ddx.estimate_bytes(dp, 100, rtm=False)
    assert ddx.estimate_bytes(dp, 100, rtm=False) == ddx.estimate_bytes(d32, 100, rtm=False)
    assert ring16 < ringp < ring16 * 1.05


def test_create_rejects_unknown_storage():
    """Checked before any device call (runs without a GPU)."""
    with pytest.raises(ddx.DdxError, match="storage"):
        ddx.Engine(ddx.make_desc(16, 16, 16, 10.0, 6, 0.001, 15.0, 10, storage=3))  # bfp16: a store format only
    with pytest.raises(ddx.DdxError, match="ring_storage"):
        ddx.Engine(ddx.make_desc(16, 16, 16, 10.0, 6, 0.001, 15.0, 10, ring_storage=4))
    with pytest.raises(ddx.DdxError, match="ring_storage"):
        ddx.Engine(ddx.make_desc(16, 16, 16, 10.0, 6, 0.001, 15.0, 10, storage="bf16", ring_storage="bfp16"))
    with pytest.raises(ddx.DdxError, match="ring_storage"):
        ddx.Engine(ddx.make_desc(16, 16, 16, 10.0, 6, 0.001, 15.0, 10, storage="fp16", ring_storage="bf16"))


def _oracle_rtm(c, storage=None, ring=None):
    rows, _ = oracle_rows(c, (1, 1, 1))
    nodes = [n for _, _, n in rows]
    w = po.ricker(c["fm"], c["dt"], c["ntime"])
    o = oracle_for(c)
    shot = o.node(*c["shots"][0])
    obs = o.model_shot(shot, w, nodes).astype(np.float32)
    o = oracle_for(c)
    if storage:
        o.set_storage(*storage)
    if ring:
        o.set_ring_storage(*ring)
    o.rtm_shot(shot, w, nodes, obs)
    return o.image(1)


def test_oracle_ring_storage():
    """The oracle's 16-bit VOWR store: over 16-bit wavefields of the same
    format and scales it changes nothing (their values are already
    representable: rounding is idempotent); over FP32 wavefields it changes
    the image, by less than 16-bit wavefields do."""
    c = case(ntime=50)
    ref = _oracle_rtm(c)
    wave = _oracle_rtm(c, storage=(po.STORE_BF16, 0, 0))
    assert np.array_equal(_oracle_rtm(c, storage=(po.STORE_BF16, 0, 0), ring=(po.STORE_BF16, 0, 0)), wave)
    wave16 = _oracle_rtm(c, storage=(po.STORE_FP16, 14, -6))
    assert np.array_equal(_oracle_rtm(c, storage=(po.STORE_FP16, 14, -6), ring=(po.STORE_FP16, 14, -6)), wave16)
    ring = _oracle_rtm(c, ring=(po.STORE_BF16, 0, 0))
    assert 0 < rel_l2(ring, ref) < rel_l2(wave, ref)


def test_bfp16_store_rounding_pins():
    """The block-scaled store's rounding (the oracle's bfp_round), pinned by
    what the format fixes: runs of 16 share the exponent s of their largest
    |value| (2^(s-1) <= largest < 2^s); u = |v| 2^-s is kept as
    (4096 + M) 2^(E-20), E = 1..7, for u >= 2^-7 and as a multiple of 2^-19
    below, rounded to nearest even; the largest code is 8191 2^-13."""
    rng = np.random.default_rng(3)
    # values on the format's grid are kept bit for bit (s = 5: largest 24)
    E = rng.integers(1, 8, 10)
    M = rng.integers(0, 4096, 10)
    c = rng.integers(0, 4096, 5)
    grid = np.concatenate([[0.75], np.ldexp((4096 + M).astype(np.float64), E - 20),
                           np.ldexp(c.astype(np.float64), -19)])
    run = (np.ldexp(grid, 5) * rng.choice([-1.0, 1.0], 16)).astype(np.float32)
    assert np.array_equal(po.bfp_round(run), run)
    # ties to even, a mantissa rounding up into the next binade, the 2^-7
    # threshold, the largest code (s = 0: largest 0.75)
    q = 2.0 ** -14  # the quantum of u in [0.25, 0.5)
    x = np.array([0.75, 5000.5 * q, 5001.5 * q, 8191.5 * q, 2.0 ** -7 - 2.0 ** -20, 3.25 * 2.0 ** -19,
                  3.5 * 2.0 ** -19, 2.0 ** -30], np.float32)
    want = np.array([0.75, 5000 * q, 5002 * q, 0.5, 2.0 ** -7, 3 * 2.0 ** -19, 4 * 2.0 ** -19, 0.0], np.float32)
    assert np.array_equal(po.bfp_round(x), want)
    top = np.array([np.nextafter(np.float32(1), np.float32(0))], np.float32)
    assert po.bfp_round(top)[0] == np.float32(8191 * 2.0 ** -13)
    # runs are independent: 1e-6 next to 1.0 (s = 1) falls below 2^(s-20)
    # and is lost; in a run of its own it keeps 2^-13 of itself
    x = np.zeros(32, np.float32)
    x[0], x[1], x[16] = 1.0, 1e-6, 1e-6
    y = po.bfp_round(x)
    assert y[1] == 0.0 and abs(y[16] - 1e-6) <= 1e-6 * 2.0 ** -13
    # subnormal floats: s = -126, kept exactly on the 2^-145 grid
    tiny = np.array([2.0 ** -140, -3 * 2.0 ** -145], np.float32)
    assert np.array_equal(po.bfp_round(tiny), tiny)
    # bounds over runs spanning 60 decades
    for _ in range(300):
        r = (rng.standard_normal(16) * 10.0 ** rng.uniform(-30, 30)).astype(np.float32)
        r[rng.random(16) < 0.3] *= np.float32(10.0 ** rng.uniform(-8, 0))
        s = int(np.frexp(np.abs(r).max())[1])
        err = np.abs(po.bfp_round(r).astype(np.float64) - r)
        assert np.all(err <= np.maximum(np.abs(r) * 2.0 ** -13, 2.0 ** (s - 20)))
