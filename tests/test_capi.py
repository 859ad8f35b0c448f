"""C ABI checks that need no GPU: the library loads, exports every function
include/ddx.h declares, and its host-only helpers agree with the restated
reference (decomposition, coefficients, wavelet, stability limit)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import ddx
import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "ddx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ddx_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(ddx.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(ddx.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


@pytest.mark.parametrize("n,pml,divide", [((256, 256, 200), 20, (2, 2, 1)), ((100, 100, 100), 10, (1, 1, 1)),
                                          ((101, 77, 50), 10, (3, 2, 2)), ((64, 64, 64), 8, (4, 1, 3)),
                                          ((676, 676, 210), 20, (2, 2, 1))])
def test_decomposition_matches_reference_rules(n, pml, divide):
    d = ddx.make_desc(*n, 10.0, pml, 0.001, 15.0, 10, divide=divide)
    subs = po.decompose(n, (10.0,) * 3, pml, divide)
    covered = np.zeros(3, int)
    for s in subs:
        g = ddx.subdomain(d, s["rank"])
        assert list(g["phy_n"]) == s["phy_n"]
        assert list(g["phy_off"]) == s["phy_off"]
        assert list(g["ext_n"]) == s["ext_n"]
        # margins (top,bottom,front,back,right,left) = (z lo, z hi, y lo, y hi, x lo, x hi)
        assert list(g["margin"]) == [s["mlo"][2], s["mhi"][2], s["mlo"][1], s["mhi"][1], s["mlo"][0], s["mhi"][0]]
        for a in range(3):
            if all(s["ci"][b] == 0 for b in range(3) if b != a):
                covered[a] += s["phy_n"][a]
    # the subdomains tile the global gridphy along each axis
    for a in range(3):
        assert covered[a] == n[a]


def test_config2_subdomains():
    d = ddx.make_desc(256, 256, 200, 10.0, 20, 0.001, 15.0, 10, divide=(2, 2, 1))
    for r in range(4):
        g = ddx.subdomain(d, r)
        assert g["phy_n"] == (128, 128, 200) and g["ext_n"] == (148, 148, 240)


def test_bad_decomposition_rejected():
    d = ddx.make_desc(20, 20, 20, 10.0, 6, 0.001, 15.0, 10, divide=(4, 1, 1))
    with pytest.raises(ddx.DdxError):
        ddx.subdomain(d, 0)  # 32/4 - 6 = 2 physical cells: too thin


@pytest.mark.parametrize("half", [1, 2, 3, 4])
def test_fd_coef_host_helper(half):
    assert np.array_equal(ddx.fd_coef(half), po.fd_coef(half).astype(np.float32))


@pytest.mark.parametrize("fm,dt,nt", [(15.0, 0.001, 1000), (12.0, 0.001, 3000), (100.0, 0.0001, 1600)])
def test_ricker_host_helper(fm, dt, nt):
    assert np.array_equal(ddx.ricker(fm, dt, nt), po.ricker(fm, dt, nt))
    assert ddx.ntime(1.0, 0.001) == 1000


def test_stable_dt():
    csum = 1225 / 1024 + 245 / 3072 + 49 / 5120 + 5 / 7168
    for v, h in ((3000.0, 10.0), (4500.0, 10.0), (4500.0, 12.5)):
        assert abs(ddx.stable_dt(v, h) - 1.0 / csum / (math.sqrt(3) / h * v)) < 1e-9


def _box_smooth_ref(a, radius):
    """Independent restatement: per axis, mean over the window [i-r, i+r]
    clipped to the grid, via cumulative sums in float64."""
    out = a.astype(np.float64)
    for ax, r in zip((2, 1, 0), radius):  # x, y, z axes of a [nz][ny][nx] array
        if r == 0:
            continue
        n = out.shape[ax]
        c = np.cumsum(out, axis=ax)
        c = np.concatenate([np.zeros_like(np.take(c, [0], axis=ax)), c], axis=ax)
        idx = np.arange(n)
        lo, hi = np.maximum(idx - r, 0), np.minimum(idx + r, n - 1)
        s = np.take(c, hi + 1, axis=ax) - np.take(c, lo, axis=ax)
        shape = [1, 1, 1]
        shape[ax] = n
        out = (s / (hi - lo + 1).reshape(shape)).astype(np.float32).astype(np.float64)
    return out.astype(np.float32)


@pytest.mark.parametrize("radius", [(0, 0, 0), (3, 0, 0), (0, 2, 0), (0, 0, 1), (4, 2, 1), (20, 4, 1)])
def test_smooth_model_matches_restatement(radius):
    rng = np.random.default_rng(5)
    a = (2000 + 500 * rng.random((9, 11, 13))).astype(np.float32)
    got = ddx.smooth_model(a, radius)
    ref = _box_smooth_ref(a, radius)
    np.testing.assert_allclose(got, ref, rtol=2e-7, atol=0)
    if radius == (0, 0, 0):
        assert np.array_equal(got, a)


def test_smooth_model_properties():
    c = np.full((6, 7, 8), 2500.0, np.float32)
    assert np.array_equal(ddx.smooth_model(c, (3, 3, 3)), c)  # constants are fixed points
    a = np.zeros((5, 5, 9), np.float32)
    a[2, 2, 4] = 9.0
    s = ddx.smooth_model(a, (1, 0, 0))  # a spike spreads over 3 cells along x, mean 3
    assert np.array_equal(s[2, 2, 3:6], [3.0, 3.0, 3.0]) and s.sum() == 9.0
