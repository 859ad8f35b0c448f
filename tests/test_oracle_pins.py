"""Pins of the CPU oracle (no GPU).

Each check derives the expected answer from something other than the
oracle's own formula: exact rationals, closed forms, symmetry, physics
(arrival times, absorption), exact invertibility (VOWR with R = 4) and
algebraic identities of the imaging condition.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import pyoracle as po
from ddx import synth
from conftest import rel_l2

# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
EXACT = {
    1: [Fraction(1)],
    2: [Fraction(9, 8), Fraction(-1, 24)],
    3: [Fraction(75, 64), Fraction(-25, 384), Fraction(3, 640)],
    4: [Fraction(1225, 1024), Fraction(-245, 3072), Fraction(49, 5120), Fraction(-5, 7168)],
}


@pytest.mark.parametrize("half", [1, 2, 3, 4])
def test_fd_coefficients_exact(half):
    c = po.fd_coef(half)
    assert np.allclose(c, [float(x) for x in EXACT[half]], rtol=0, atol=1e-15)
    # staggered first-derivative conditions: sum c_r (2r+1) = 1 and
    # sum c_r (2r+1)^(2k+1) = 0 for k = 1..half-1
    for k in range(half):
        s = sum(Fraction(x) * (2 * r + 1) ** (2 * k + 1) for r, x in enumerate(EXACT[half]))
        assert s == (1 if k == 0 else 0)


def test_fd_abs_sum_and_cfl():
    assert abs(np.abs(po.fd_coef(4)).sum() - 1.2863095238) < 1e-9
    assert abs(1 / (math.sqrt(3) * np.abs(po.fd_coef(4)).sum()) - 0.44884) < 1e-5


@pytest.mark.parametrize("T,dt,n", [(1.0, 0.001, 1000), (2.0, 0.001, 2000), (1.5, 0.001, 1500), (3.0, 0.001, 3000),
                                    (0.16, 0.0001, 1600)])
def test_ntime(T, dt, n):
    assert po.ntime(T, dt) == n


@pytest.mark.parametrize("fm,n0", [(15.0, 87), (12.0, 108), (25.0, 52)])
def test_ricker_closed_form(fm, n0):
    dt = 0.001
    w = po.ricker(fm, dt, 400)
    assert int(np.argmax(w)) == n0
    assert w[n0] == np.float32(1e4)
    t = (np.arange(400) - n0) * dt
    x = (math.pi * fm * t) ** 2
    ref = 1e4 * (1 - 2 * x) * np.exp(-x)
    assert np.max(np.abs(w - ref)) <= 1e-3 * 1e4 * 1e-3
    k = np.arange(1, n0)
    assert np.allclose(w[n0 + k], w[n0 - k], rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("l,r,d,n", [(-100.0, 150.0, 10.0, 25), (-200.0, 2000.0, 10.0, 220), (1280.0, 2000.0, 10.0, 72),
                                     (0.0, 0.0, 12.5, 0), (-250.0, 1237.5, 12.5, 119), (0.0, 9.9999, 10.0, 1)])
def test_gap_num(l, r, d, n):
    assert po.gap_num(l, r, d) == n


# ---------------------------------------------------------------------------
# model build and CPML coefficients
# ---------------------------------------------------------------------------
def _tiny(prec=4, **kw):
    nx, ny, nz, pml = 12, 10, 9, 6
    vp, vs, rho = synth.homogeneous_noise(nx, ny, nz, seed=3, rel=0.05)
    args = dict(nx=nx, ny=ny, nz=nz, dx=10.0, pml=pml, dt=0.001, fm=20.0, ntime=10, rec_width=1)
    args.update(kw)
    return args, vp, vs, rho, po.Oracle(args["nx"], args["ny"], args["nz"], args["dx"], args["pml"], args["dt"],
                                        args["fm"], args["ntime"], 1, vp, vs, rho, prec=prec)


def test_extension_and_lame():
    a, vp, vs, rho, o = _tiny()
    m = a["pml"]
    NZ, NY, NX = o.ext_shape
    k, j, i = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    ci = np.clip(i - m, 0, a["nx"] - 1)
    cj = np.clip(j - m, 0, a["ny"] - 1)
    ck = np.clip(k - m, 0, a["nz"] - 1)
    evp, evs, erh = vp[ck, cj, ci], vs[ck, cj, ci], rho[ck, cj, ci]
    lam = (erh * (evp * evp - np.float32(2.0) * evs * evs)).astype(np.float32)
    mu = (erh * (evs * evs)).astype(np.float32)
    assert np.array_equal(o.coef(12).astype(np.float32), lam)
    assert np.array_equal(o.coef(13).astype(np.float32), mu)
    assert np.array_equal(o.coef(14).astype(np.float32), erh)


@pytest.mark.parametrize("prec", [4, 8])
def test_cpml_profiles_from_first_principles(prec):
    """Invert (a, b) back to the damping d and alpha and compare with the
    textbook CPML profile d = log10(1/R) * 7 * vp / (2 L h) (dist/L)^2,
    alpha = pi fm (1 - dist/L), R = 1e-3, L = pml - 2, kappa = 1.  Low-side
    cells use the integer distance on the full grid and dist - 0.5 on the
    half grid, high-side cells the opposite."""
    a, vp, vs, rho, o = _tiny(prec)
    m, fm, h = a["pml"], a["fm"], a["dx"]
    dt = float(np.float32(a["dt"]))  # the run description carries dt as float32
    L = m - 2
    NZ, NY, NX = o.ext_shape
    A, B, AH, BH = o.coef(0), o.coef(1), o.coef(2), o.coef(3)  # x axis
    j, k = NY // 2, NZ // 2
    evp = vp[np.clip(k - m, 0, a["nz"] - 1), np.clip(j - m, 0, a["ny"] - 1)]
    tol = 2e-4 if prec == 4 else 1e-9
    for i in range(NX):
        v = float(evp[np.clip(i - m, 0, a["nx"] - 1)])
        lo, hi = 2 <= i < m, NX - m <= i < NX - 2
        if not (lo or hi):
            assert A[k, j, i] == 0 and AH[k, j, i] == 0
            continue
        dist_f = (m - i) if lo else (i - (NX - m - 1) - 0.5)
        dist_h = (m - i - 0.5) if lo else (i - (NX - m - 1))
        for aa, bb, dist in ((A[k, j, i], B[k, j, i], dist_f), (AH[k, j, i], BH[k, j, i], dist_h)):
            s = -math.log(bb) / dt          # d + alpha
            d = aa * s / (bb - 1.0)
            alpha = s - d
            d_ref = 3.0 * 7.0 * v / (2.0 * L * h) * (dist / L) ** 2
            a_ref = max(0.0, math.pi * fm * (1.0 - dist / L))
            assert abs(d - d_ref) <= tol * max(d_ref, 1.0), (i, d, d_ref)
            assert abs(alpha - a_ref) <= tol * max(a_ref, 1.0) + (1e-2 if prec == 4 else 1e-8), (i, alpha, a_ref)


# ---------------------------------------------------------------------------
# propagation physics
# ---------------------------------------------------------------------------
def _homog(n=40, pml=10, ntime=300, vp0=3000.0, fm=15.0, prec=4, rec_width=1):
    vp = np.full((n, n, n), vp0, np.float32)
    vs = (vp / np.sqrt(np.float32(3))).astype(np.float32)
    rho = np.full_like(vp, 2000.0)
    return po.Oracle(n, n, n, 10.0, pml, 0.001, fm, ntime, rec_width, vp, vs, rho, prec=prec)


def test_symmetry_diagonal_and_mirror():
    # source at the exact centre of a 37^3 grid, window closing before the
    # boundary reflections reach the receivers
    o = _homog(n=37, ntime=100)
    c = 18 * 10.0
    src = o.node(c, c, c)
    recs = [o.node(c + 80, c + 80, c), o.node(c - 80, c, c), o.node(c + 70, c, c)]
    w = po.ricker(15.0, 0.001, 100)
    s = o.model_shot(src, w, recs)
    # vx == vy on the x = y diagonal of an explosive source, bit for bit
    assert np.array_equal(s[0, 0], s[1, 0])
    # vx is staggered at i+1/2: node c-80 (x = -7.5 cells) mirrors node c+70
    # (x = +7.5); mirrored stencils round differently in FP32
    assert rel_l2(-s[0, 1], s[0, 2]) < 1e-4
    s8 = _homog(n=37, ntime=100, prec=8).model_shot(src, w, recs)
    assert rel_l2(-s8[0, 1], s8[0, 2]) < 1e-6


def test_p_arrival_time():
    """Peak of the radial velocity arrives at n0*dt + r/vp (within 2 samples)."""
    o = _homog(n=48, ntime=220)
    c = 24 * 10.0
    src = o.node(c, c, c)
    w = po.ricker(15.0, 0.001, 220)
    n0 = int(np.argmax(w))
    for r in (80.0, 150.0):
        rec = o.node(c + r - 5.0, c, c)  # vx node sits half a cell right: x = r
        s = o.model_shot(src, w, [rec])
        vx = s[0, 0]
        # velocity is the time derivative of the displacement pulse; its
        # envelope peak sits between its two extrema
        t_pos, t_neg = int(np.argmax(vx)), int(np.argmin(vx))
        t_mid = 0.5 * (t_pos + t_neg)
        expect = n0 + r / 3000.0 / 0.001
        assert abs(t_mid - expect) <= 3.0, (r, t_mid, expect)


def test_cpml_absorbs():
    """A receiver near a boundary sees (almost) no reflection: compared with
    the same receiver in a much larger domain over the same time window."""
    nt = 260
    w = po.ricker(15.0, 0.001, nt)
    small = _homog(n=30, pml=12, ntime=nt)
    big = _homog(n=110, pml=12, ntime=nt)
    # source and receiver at the same offsets from the low corner of the small grid
    xs, xr = 60.0, 40.0
    s1 = small.model_shot(small.node(xs, xs, xs), w, [small.node(xr, xs, xs)])
    off = 400.0
    s2 = big.model_shot(big.node(xs + off, xs + off, xs + off), w, [big.node(xr + off, xs + off, xs + off)])
    err = rel_l2(s1[0, 0], s2[0, 0])
    assert err < 0.05, err
    # and the wavefield decays after the pulse left the small grid
    late = np.abs(small.get(0, po.VX)).max()
    assert late < 0.05 * np.abs(s1[0, 0]).max()


def test_fp32_matches_fp64():
    a = synth.layered(20, 18, 16, 5, nlayer=(3, 4), vp_range=(2000.0, 3000.0), anomaly_radius=3)
    res = []
    for prec in (4, 8):
        o = po.Oracle(20, 18, 16, 10.0, 6, 0.001, 25.0, 150, 1, *a, prec=prec)
        res.append(o.model_shot(o.node(100.0, 90.0, 40.0), po.ricker(25.0, 0.001, 150),
                                [o.node(x, 50.0, 20.0) for x in (20.0, 80.0, 150.0)]))
    assert rel_l2(res[0], res[1]) < 1e-5


# ---------------------------------------------------------------------------
# VOWR reconstruction
# ---------------------------------------------------------------------------
def _recon_error(R, prec, nt=120):
    """max over snapshots of |backward - forward| / max over snapshots |forward|
    (L2 over gridphy)."""
    vp, vs, rho = synth.layered(22, 20, 18, 9, nlayer=(3, 4), vp_range=(2200.0, 3000.0), anomaly_radius=3)
    o = po.Oracle(22, 20, 18, 10.0, 6, 0.001, 25.0, nt, R, vp, vs, rho, prec=prec)
    rec = [o.node(50.0, 50.0, 20.0)]
    w = po.ricker(25.0, 0.001, nt)
    shot = o.node(110.0, 100.0, 90.0)
    obs = np.zeros((3, 1, nt), np.float32)
    fs, bs = o.rtm_shot(shot, w, rec, obs, snap_every=20)
    m = 6
    f = fs[1:, m:-m, m:-m, m:-m]
    b = bs[1:, m:-m, m:-m, m:-m]
    fmax = max(np.linalg.norm(x) for x in f)
    return max(np.linalg.norm(x - y) for x, y in zip(b, f)) / fmax


def test_vowr_exact_at_full_width():
    """With R = phy_fdorder_half the backward pass inverts the forward pass
    exactly (up to rounding): the reconstructed source wavefield equals the
    stored forward snapshots over gridphy."""
    assert _recon_error(4, 8) < 1e-10
    assert _recon_error(4, 4) < 1e-4


def test_vowr_error_falls_with_width():
    e = [_recon_error(R, 8) for R in (1, 2, 4)]
    assert e[0] > e[1] > e[2], e
    assert e[0] < 0.2, e


# ---------------------------------------------------------------------------
# imaging condition and post-processing
# ---------------------------------------------------------------------------
def test_imaging_identities():
    a, vp, vs, rho, o = _tiny(prec=8)
    rng = np.random.default_rng(0)
    shp = o.ext_shape
    fields = {f: rng.standard_normal(shp) for f in (po.VX, po.VY, po.VZ, po.VXP, po.VYP, po.VZP)}
    for s in (0, 1):
        for f, v in fields.items():
            o.set(s, f, v)
    o.reset_images()
    o.image_accumulate()
    # identical source and receiver wavefields: Ipp == Ip, Ips == Isp
    assert np.array_equal(o.image(1), o.image(0))
    assert np.array_equal(o.image(2), o.image(3))
    # pure-S source field (P part zero): every P-source image vanishes
    for f in (po.VXP, po.VYP, po.VZP):
        o.set(0, f, np.zeros(shp))
    o.reset_images()
    o.image_accumulate()
    for q in (0, 1, 2):
        assert np.abs(o.image(q)).max() == 0
    m = a["pml"]
    vx, vy, vz = fields[po.VX], fields[po.VY], fields[po.VZ]
    ss = (vx * (vx - fields[po.VXP]) + vy * (vy - fields[po.VYP]) + vz * (vz - fields[po.VZP]))[m:-m, m:-m, m:-m]
    assert np.allclose(o.image(4), ss, rtol=1e-12, atol=1e-12)


def test_laplacian_and_compensation():
    a, vp, vs, rho, o = _tiny(prec=8)
    nx, ny, nz = a["nx"], a["ny"], a["nz"]
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    quad = (i * i + 2.0 * j * j + 3.0 * k * k).astype(np.float64)
    for q in range(1, 5):
        o.set_image(q, quad * q)
    o.set_image(0, np.full((nz, ny, nx), 2.0))
    img, lap = o.finalize_box(0, 0, 0, nx, ny, nz)
    for q in range(4):
        inner = lap[q][2:-2, 2:-2, 2:-2]
        assert np.allclose(inner, 12.0 * (q + 1))  # Laplacian of x^2 + 2y^2 + 3z^2 is 12
        border = lap[q].copy()
        border[2:-2, 2:-2, 2:-2] = 0
        assert np.abs(border).max() == 0
        assert np.allclose(img[q], quad * (q + 1) / 2.0)
    # a subdomain box applies its own 2-cell border
    img, lap = o.finalize_box(3, 2, 1, 6, 6, 6)
    assert np.allclose(lap[0][2:-2, 2:-2, 2:-2], 12.0)
    assert lap[0][0].max() == 0 and lap[0][:, :, 1].max() == 0


def test_compensated_image_is_ill_conditioned_in_fp32():
    """The FP32 and FP64 oracles agree on the uncompensated image to FP32
    round-off, and the compensated image (divided by the illumination Ip,
    tiny far from the source) differs more: the division amplifies the
    round-off of low-illumination cells.  The GPU engine's distance from the
    FP64 oracle is stated per result with this in mind."""
    a = synth.layered(20, 18, 16, 5, nlayer=(3, 4), vp_range=(2000.0, 3000.0), anomaly_radius=3)
    raw, comp = [], []
    for prec in (4, 8):
        o = po.Oracle(20, 18, 16, 10.0, 6, 0.001, 25.0, 60, 2, *a, prec=prec)
        shot = o.node(100.0, 90.0, 40.0)
        rec = [o.node(x, y, 20.0) for x in (20.0, 80.0, 150.0) for y in (30.0, 120.0)]
        w = po.ricker(25.0, 0.001, 60)
        obs = o.model_shot(shot, w, rec).astype(np.float32)
        o.rtm_shot(shot, w, rec, obs)
        raw.append(o.image(1))
        comp.append(o.finalize_box(0, 0, 0, 20, 18, 16)[0][0])
    e_raw, e_comp = rel_l2(raw[0], raw[1]), rel_l2(comp[0], comp[1])
    assert e_raw < 1e-4
    assert e_comp > e_raw


# ---------------------------------------------------------------------------
# FWI hooks (PHY/fdtd3d_phy_mpi_2_vel_kernel.hpp:70-78,
# BASE/fdtd3d_base_4_seisrecord_meat.hpp:109-147)
# ---------------------------------------------------------------------------
def test_fwi_mode_accumulates_displacement_and_increment():
    """FWI mode propagates exactly as pure forward modelling; on gridphy it
    accumulates u += dt * v (after the update) and stores dt times the
    velocity increment of the step; nothing else changes."""
    a = synth.layered(16, 14, 12, 3, nlayer=(2, 3), vp_range=(2000.0, 2800.0), anomaly_radius=2)
    runs = {}
    for mode in (po.PURE_FORWARD, po.FWI):
        o = po.Oracle(16, 14, 12, 10.0, 5, 0.001, 25.0, 10, 1, *a, prec=4)
        rng = np.random.default_rng(4)
        shape = o.ext_shape
        for f in range(po.SXX, po.SAU):  # velocities start at 0, stresses random
            o.set(0, f, rng.standard_normal(shape) * 1e-3)
        hist = []
        for _ in range(3):
            before = o.get(0, po.VX).astype(np.float32)
            o.vel(0, mode)
            hist.append((before, o.get(0, po.VX).astype(np.float32), o.get(0, po.UX).astype(np.float32),
                         o.get(0, po.VXRT).astype(np.float32)))
            o.stress(0, mode)
        runs[mode] = (o, hist)
    op, of = runs[po.PURE_FORWARD][0], runs[po.FWI][0]
    for f in range(po.VZP + 1):
        assert np.array_equal(op.get(0, f), of.get(0, f)), f
    m = 5
    phy = (slice(m, -m),) * 3
    dt = np.float32(0.001)
    u = np.zeros_like(runs[po.FWI][1][0][0])
    for step, (before, after, ux, vxrt) in enumerate(runs[po.FWI][1]):
        u = (u + dt * after).astype(np.float32)
        assert np.array_equal(ux[phy], u[phy])  # u += dt * v, in float, after the update
        if step == 0:  # v starts at 0, so the increment is the new v exactly
            assert np.array_equal(vxrt[phy], dt * after[phy])
        else:
            assert rel_l2(vxrt[phy], dt * (after - before)[phy]) < 1e-4
        assert not ux[:m].any() and not vxrt[:m].any()  # gridphy only


def test_residual_injection():
    a = synth.homogeneous_noise(12, 12, 12, seed=2)
    o = po.Oracle(12, 12, 12, 10.0, 4, 0.001, 25.0, 6, 1, *a, prec=4)
    nodes = [o.node(30.0, 40.0, 20.0), o.node(60.0, 40.0, 20.0)]
    rng = np.random.default_rng(1)
    obs = rng.standard_normal((3, 2, 6)).astype(np.float32)
    syn = rng.standard_normal((3, 2, 6)).astype(np.float32)
    o.inject_residual(1, nodes, obs, syn, 4)
    for q in range(3):
        v = o.get(1, po.VX + q).ravel()
        for r, n in enumerate(nodes):
            assert np.float32(v[n]) == obs[q, r, 4] - syn[q, r, 4]
        assert np.count_nonzero(v) == 2
