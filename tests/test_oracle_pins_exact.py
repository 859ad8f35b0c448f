"""Exact pins of the CPU oracle (no GPU), each shown to bite.

Every check derives its expected answer from the mathematics, not from the
oracle's own code, and is run twice: on the faithful oracle, where it must
pass, and on the oracle with a planted mistake (pyoracle.planted_fault),
where it must fail.

- reverse receiver injection (BASE/fdtd3d_base_4_seisrecord_meat.hpp:79-107,
  called at RTM/rtm3d_blend_elastic_model_mpi_meat.hpp:67 before the reverse
  velocity update): a unit spike in the observed data appears in the
  reverse set at the receiver node, with its sign and component, at exactly
  its sample's iteration, and nothing appears before it;
- the vertical-stress source variant (PHY/fdtd3d_phy_mpi_3_stress_kernel.hpp:
  1308-1364 with szz only): one step writes szz at the node and nothing
  else; in a homogeneous medium it radiates an S wave that arrives at r/vs,
  where the explosive source radiates none;
- the interior 8th-order scheme, normal and shear terms
  (PHY/fdtd3d_phy_mpi_2_vel_kernel.hpp:7-80, _3_stress_kernel.hpp:7-60): a
  plane P or S wave sampled on the staggered grid is an exact solution of the
  leapfrog scheme when its frequency obeys the scheme's closed-form
  dispersion relation  (2/dt) sin(w dt/2) = c (2/h) sum_r c_r sin((2r+1)kh/2);
- the CPML half/full-grid assignment (CPML/fdtd3d_cpml_mpi_1_host_meat.hpp:
  186-189 low side, :304-307 high side; CPML/..._device_{x,y,z}_meat.hpp):
  every damped derivative uses the damping profile evaluated at the
  staggered position of the variable it updates.  Both sides reduce to one
  rule in positions: distance m - p on the low side and p - (N - m - 1/2)
  on the high side for a variable at position p (cells: integer p; half
  nodes: p = i + 1/2).  A 1D restatement built from that rule is compared
  with the oracle on a plane pulse crossing both CPMLs.  (The two sides are
  not mirror images of each other: the high-side profile sits half a cell
  off the mirror of the low side, so a mirror-symmetry check cannot pin this.)
"""
import math

import numpy as np
import pytest

import pyoracle as po

VP, VS, RHO = 3000.0, 1500.0, 2000.0  # lambda = 2 mu: a lambda-for-mu mistake shows
H, DT = 10.0, 0.001
DT32 = float(np.float32(DT))  # the time step the oracle steps with (a float argument)


def _homog_oracle(n, pml, ntime, prec=8, vs=VS, fm=15.0, rec_width=1):
    shape = (n[2], n[1], n[0])
    vp = np.full(shape, VP, np.float32)
    vsa = np.full(shape, vs, np.float32)
    rho = np.full(shape, RHO, np.float32)
    return po.Oracle(n[0], n[1], n[2], H, pml, DT, fm, ntime, rec_width, vp, vsa, rho, prec=prec)


# ---------------------------------------------------------------------------
# reverse receiver injection
# ---------------------------------------------------------------------------
def _injection_errors():
    """(largest reverse-set value before the spike's iteration, largest
    deviation from the expected spike state after it)."""
    nt, t0, amp = 24, 9, 0.75
    o = _homog_oracle((20, 18, 16), 6, nt)
    recs = [o.node(60.0, 70.0, 50.0), o.node(120.0, 90.0, 80.0)]
    obs = np.zeros((3, 2, nt), np.float32)
    q0, r0 = 1, 1
    obs[q0, r0, t0] = amp
    w = np.zeros(nt, np.float32)  # no source: the source side stays zero
    before, at = 0.0, 0.0
    for q in range(3):
        o.set_snap(1, po.VX + q)
        _, bs = o.rtm_shot(o.node(100.0, 90.0, 80.0), w, recs, obs, snap_every=1)
        # bs[j] = reverse set after reverse iteration it = nt - j (j >= 1)
        for j in range(1, nt - t0):  # iterations it > t0
            before = max(before, float(np.abs(bs[j]).max()))
        # after iteration t0 the spike has been added before the velocity
        # update, whose stress terms are still zero: v_q0 = amp at the node
        want = np.zeros(o.NX * o.NY * o.NZ)
        if q == q0:
            want[recs[r0]] = amp
        at = max(at, float(np.abs(bs[nt - t0].ravel() - want).max()))
    return before, at


def test_reverse_injection_spike():
    before, at = _injection_errors()
    assert before == 0.0 and at == 0.0, (before, at)


@pytest.mark.parametrize("fault", [po.FAULT_INJ_TIME, po.FAULT_INJ_SIGN, po.FAULT_INJ_COMP])
def test_reverse_injection_spike_bites(fault):
    with po.planted_fault(fault):
        before, at = _injection_errors()
    assert before > 0.1 or at > 0.1, (fault, before, at)


# ---------------------------------------------------------------------------
# vertical-stress source
# ---------------------------------------------------------------------------
def _source_step_error(src_mode):
    o = _homog_oracle((12, 12, 12), 4, 4)
    node = o.node(50.0, 60.0, 70.0)
    o.reset(0)
    o.source(0, node, 2.5, +1, src_mode)
    want = {po.SXX: 0.0 if src_mode == 1 else 2.5, po.SYY: 0.0 if src_mode == 1 else 2.5, po.SZZ: 2.5}
    err = 0.0
    for f in (po.VX, po.VY, po.VZ, po.SXX, po.SYY, po.SZZ, po.SXY, po.SXZ, po.SYZ, po.SAU):
        a = o.get(0, f).ravel()
        e = np.zeros_like(a)
        e[node] = want.get(f, 0.0)
        err = max(err, float(np.abs(a - e).max()))
    o.source(0, node, 2.5, -1, src_mode)  # the subtraction removes it again
    for f in (po.SXX, po.SYY, po.SZZ):
        err = max(err, float(np.abs(o.get(0, f)).max()))
    return err


@pytest.mark.parametrize("src_mode", [0, 1])
def test_source_writes_only_its_components(src_mode):
    assert _source_step_error(src_mode) == 0.0


def test_source_pin_bites():
    with po.planted_fault(po.FAULT_SRC_SZZ):
        assert _source_step_error(1) > 1.0


def test_s_wave_arrival_from_vertical_stress_source():
    """Receiver at 45 degrees in the x-z plane: the szz-only source radiates
    an S wave (transverse motion) that peaks at n0*dt + r/vs; the explosive
    source radiates none (x <-> z symmetry of the grid makes it vanish to
    rounding)."""
    n, nt = 56, 280
    w = po.ricker(15.0, DT, nt)
    n0 = int(np.argmax(w))
    c = n // 2 * H
    d = 14  # cells along x and z
    r = math.hypot(d * H, d * H)
    res = {}
    for mode in (0, 1):
        o = _homog_oracle((n, n, n), 10, nt, prec=4)
        src = o.node(c, c, c)
        rec = o.node(c + d * H, c, c + d * H)
        s = o.model_shot(src, w, [rec], src_mode=mode)
        radial = (s[0, 0] + s[2, 0]) / math.sqrt(2.0)
        trans = (s[0, 0] - s[2, 0]) / math.sqrt(2.0)
        res[mode] = (radial, trans)
    rad0, tr0 = res[0]
    assert np.abs(tr0).max() < 1e-4 * np.abs(rad0).max()
    rad1, tr1 = res[1]
    t_s = 0.5 * (int(np.argmax(tr1)) + int(np.argmin(tr1)))
    t_p = 0.5 * (int(np.argmax(rad1)) + int(np.argmin(rad1)))
    assert abs(t_s - (n0 + r / VS / DT)) <= 6.0, (t_s, n0 + r / VS / DT)
    assert abs(t_p - (n0 + r / VP / DT)) <= 6.0, (t_p, n0 + r / VP / DT)
    assert np.abs(tr1).max() > 0.3 * np.abs(rad1).max()


# ---------------------------------------------------------------------------
# plane waves and the scheme's dispersion relation
# ---------------------------------------------------------------------------
NORMAL = {0: po.SXX, 1: po.SYY, 2: po.SZZ}
SHEAR = {(0, 1): po.SXY, (1, 0): po.SXY, (0, 2): po.SXZ, (2, 0): po.SXZ, (1, 2): po.SYZ, (2, 1): po.SYZ}
VEL = {0: po.VX, 1: po.VY, 2: po.VZ}


def _k_mod(kh):
    """Modified wavenumber times h of the 8th-order staggered difference:
    2 sum_r c_r sin((2r+1) kh / 2)."""
    c = po.fd_coef(4)
    return 2.0 * sum(c[r] * math.sin((2 * r + 1) * kh / 2.0) for r in range(4))


def _coord(shape, axis):
    """Index along `axis` (0 x, 1 y, 2 z) of every cell of a (z, y, x) grid."""
    return np.indices(shape)[2 - axis].astype(np.float64)


def _plane_wave_error(kind, axis, pol, nsteps=3):
    """Max deviation (relative to each field's amplitude) of the oracle from
    the discrete plane wave after nsteps, over the cells that no boundary or
    CPML cell can have influenced."""
    n, pml = 64, 4
    o = _homog_oracle((n, n, n), pml, 8)
    lam, mu = RHO * (VP * VP - 2 * VS * VS), RHO * VS * VS
    c = VP if kind == "P" else VS
    kh = 2.0 * math.pi / 12.0
    w = 2.0 / DT32 * math.asin(c * DT32 / H * _k_mod(kh) / 2.0)
    p = _coord(o.ext_shape, axis)
    V = -1.0 / (RHO * c)
    if kind == "P":
        # normal stresses at integer positions, the velocity half a cell on
        sig = {NORMAL[axis]: (1.0, 0.0)}
        for b in range(3):
            if b != axis:
                sig[NORMAL[b]] = (lam / (lam + 2 * mu), 0.0)
        vel = {VEL[axis]: (V, 0.5)}
    else:
        # the velocity at integer positions, the shear stress half a cell on
        sig = {SHEAR[(axis, pol)]: (1.0, 0.5)}
        vel = {VEL[pol]: (V, 0.0)}
    o.reset(0)
    for f, (amp, off) in sig.items():
        o.set(0, f, amp * np.cos(kh * (p + off)))
    for f, (amp, off) in vel.items():  # velocities at t = -dt/2
        o.set(0, f, amp * np.cos(kh * (p + off) + w * DT32 / 2.0))
    for _ in range(nsteps):
        o.vel(0, po.PURE_FORWARD)
        o.stress(0, po.PURE_FORWARD)
    lo, hi = pml + 8 * nsteps + 1, n + pml - 8 * nsteps - 1
    box = (slice(lo, hi),) * 3
    err = 0.0
    t_sig, t_vel = nsteps * DT32, (nsteps - 0.5) * DT32
    for f, (amp, off) in sig.items():
        want = amp * np.cos(kh * (p + off) - w * t_sig)
        err = max(err, float(np.abs(o.get(0, f) - want)[box].max()) / abs(amp))
    for f, (amp, off) in vel.items():
        want = amp * np.cos(kh * (p + off) - w * t_vel)
        err = max(err, float(np.abs(o.get(0, f) - want)[box].max()) / abs(amp))
    others = set(range(po.SAU)) - set(sig) - set(vel)
    for f in others:
        err = max(err, float(np.abs(o.get(0, f))[box].max()))
    return err


PLANE = [("P", a, None) for a in range(3)] + [("S", a, b) for a in range(3) for b in range(3) if a != b]


@pytest.mark.parametrize("kind,axis,pol", PLANE)
def test_plane_wave_obeys_dispersion_relation(kind, axis, pol):
    assert _plane_wave_error(kind, axis, pol) < 1e-9


@pytest.mark.parametrize("fault,case", [(po.FAULT_FD_COEF, ("P", 1, None)),
                                        (po.FAULT_SHEAR_LAMBDA, ("S", 0, 1)),
                                        (po.FAULT_SHEAR_LAMBDA, ("S", 1, 0))])
def test_plane_wave_pin_bites(fault, case):
    with po.planted_fault(fault):
        assert _plane_wave_error(*case) > 1e-5


# ---------------------------------------------------------------------------
# CPML: damping profiles at the staggered positions
# ---------------------------------------------------------------------------
def _profile(pos, N, m, vp, fm):
    """a, b of the CPML memory update for variables at positions pos (cell
    units) on an axis of N cells with margin m (2 rigid layers included):
    distance m - p into the low-side layer, p - (N - m - 1/2) into the high
    side; d = d0 (dist/L)^2, alpha = pi fm (1 - dist/L) >= 0,
    b = exp(-(d + alpha) dt), a = d (b - 1) / (d + alpha)."""
    L = m - 2.0
    d0 = math.log10(1.0 / 0.001) * 7.0 * vp / (2.0 * L * H)
    dist = np.where(pos < m, m - pos, pos - (N - m - 0.5))
    d = d0 * (dist / L) ** 2
    al = np.maximum(0.0, math.pi * fm * (1.0 - dist / L))
    b = np.exp(-(d + al) * DT32)
    a = d * (b - 1.0) / (d + al)
    return a, b


def _cpml_1d(kind, N, m, nsteps, s0, lam, mu, fm):
    """1D restatement along one axis: integer-position field F and
    half-position field Hf.  P wave: Hf is the velocity (D+ of F), F the
    normal stress (D- of Hf, modulus lambda + 2 mu).  S wave: F is the
    velocity (D- of Hf), Hf the shear stress (D+ of F, modulus mu).  Cells in
    [2, m) and [N - m, N - 2) are CPML cells (4th order, memory variables);
    cells in [m, N - m) the interior (8th order); the 2 outer cells are never
    updated."""
    c8, c4 = po.fd_coef(4), po.fd_coef(2)
    i = np.arange(N)
    full = _profile(i.astype(np.float64), N, m, VP, fm)
    half = _profile(i + 0.5, N, m, VP, fm)
    inner = (i >= m) & (i < N - m)
    pml = ((i >= 2) & (i < m)) | ((i >= N - m) & (i < N - 2))
    F = s0.copy()
    Hf = np.zeros(N)
    psiF, psiH = np.zeros(N), np.zeros(N)
    mod = lam + 2 * mu if kind == "P" else mu

    def dplus(f, c):  # sum_r c_r (f[i+r+1] - f[i-r]), i.e. at i + 1/2 from integer samples
        g = np.pad(f, 4)
        out = np.zeros(N)
        for r, cr in enumerate(c):
            out += cr * (g[4 + r + 1:4 + r + 1 + N] - g[4 - r:4 - r + N])
        return out

    def dminus(f, c):  # sum_r c_r (f[i+r] - f[i-r-1]), i.e. at i from half samples
        g = np.pad(f, 4)
        out = np.zeros(N)
        for r, cr in enumerate(c):
            out += cr * (g[4 + r:4 + r + N] - g[4 - r - 1:4 - r - 1 + N])
        return out

    def upd(val, d8, d4, psi, prof, scale):
        a, b = prof
        dd = d4 / H
        psi[pml] = b[pml] * psi[pml] + a[pml] * dd[pml]
        dd = dd + np.where(pml, psi, 0.0)
        out = val.copy()
        out[inner] = val[inner] + scale / H * d8[inner]
        out[pml] = val[pml] + scale * dd[pml]
        return out

    for _ in range(nsteps):
        if kind == "P":
            Hf = upd(Hf, dplus(F, c8), dplus(F, c4), psiH, half, DT32 / RHO)
            F = upd(F, dminus(Hf, c8), dminus(Hf, c4), psiF, full, DT32 * mod)
        else:
            F = upd(F, dminus(Hf, c8), dminus(Hf, c4), psiF, full, DT32 / RHO)
            Hf = upd(Hf, dplus(F, c8), dplus(F, c4), psiH, half, DT32 * mod)
    return F, Hf


CPML_CASES = [("P", 0, None), ("P", 1, None), ("P", 2, None), ("S", 0, 1), ("S", 1, 2), ("S", 2, 0)]


def _cpml_error(kind, axis, pol, nsteps=80):
    """Relative max deviation of the oracle's centre line from the 1D
    restatement after a Gaussian pulse crossed both CPMLs of `axis`."""
    fm = 15.0
    nprop, m, ntr = 24, 8, 72
    n = [ntr, ntr, ntr]
    n[axis] = nprop
    o = _homog_oracle(tuple(n), m, 8, fm=fm)
    N = nprop + 2 * m
    lam, mu = RHO * (VP * VP - 2 * VS * VS), RHO * VS * VS
    i = np.arange(N, dtype=np.float64)
    s0 = np.exp(-((i - (N - 1) / 2.0) / 2.5) ** 2)
    F, Hf = _cpml_1d(kind, N, m, nsteps, s0, lam, mu, fm)
    p = _coord(o.ext_shape, axis)
    s0_3d = np.exp(-((p - (N - 1) / 2.0) / 2.5) ** 2)
    o.reset(0)
    if kind == "P":
        fF, fH = NORMAL[axis], VEL[axis]
        o.set(0, fF, s0_3d)
        for b in range(3):
            if b != axis:
                o.set(0, NORMAL[b], s0_3d * lam / (lam + 2 * mu))
    else:
        fF, fH = VEL[pol], SHEAR[(axis, pol)]
        o.set(0, fF, s0_3d)
    for _ in range(nsteps):
        o.vel(0, po.PURE_FORWARD)
        o.stress(0, po.PURE_FORWARD)
    mid = [o.NZ // 2, o.NY // 2, o.NX // 2]  # array axes (z, y, x)
    sel = tuple(slice(None) if ax == 2 - axis else mid[ax] for ax in range(3))
    gF, gH = o.get(0, fF)[sel], o.get(0, fH)[sel]
    eF = np.abs(gF - F).max() / np.abs(F).max()
    eH = np.abs(gH - Hf).max() / np.abs(Hf).max()
    # the pulse must have reached the damping layers for the check to mean anything
    assert np.abs(Hf[: m]).max() > 1e-3 * np.abs(Hf).max()
    return max(eF, eH)


@pytest.mark.parametrize("kind,axis,pol", CPML_CASES)
def test_cpml_profiles_at_staggered_positions(kind, axis, pol):
    assert _cpml_error(kind, axis, pol) < 1e-8


@pytest.mark.parametrize("fault,case", [(po.FAULT_PML_VX_X, ("P", 0, None)), (po.FAULT_PML_VY_Y, ("P", 1, None)),
                                        (po.FAULT_PML_VZ_Z, ("P", 2, None)), (po.FAULT_PML_SXX_X, ("P", 0, None)),
                                        (po.FAULT_PML_SXY_X, ("S", 0, 1)), (po.FAULT_PML_VY_SXY_X, ("S", 0, 1))])
def test_cpml_pin_bites(fault, case):
    with po.planted_fault(fault):
        assert _cpml_error(*case) > 1e-6


# ---------------------------------------------------------------------------
# the oracle's vectorized row runs are its per-cell code
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", [4, 8])
def test_vectorized_runs_equal_per_cell_code(prec):
    """A modelling shot and an RTM shot (every mode, CPML and VOWR edges) give
    bit-identical seismograms, snapshots and images with the vectorized runs
    and with the per-cell code."""
    from cases import case, oracle_for
    c = case(nx=26, ny=21, nz=19, ntime=50, rec_width=2)
    out = []
    for scalar in (False, True):
        ctx = po.scalar_paths() if scalar else None
        if ctx:
            ctx.__enter__()
        try:
            o = oracle_for(c, prec=prec)
            w = po.ricker(c["fm"], c["dt"], c["ntime"])
            shot = o.node(*c["shots"][0])
            recs = [o.node(*r) for r in c["recv"]]
            seis = o.model_shot(shot, w, recs)
            fs, bs = o.rtm_shot(shot, w, recs, seis.astype(np.float32), snap_every=10)
            out.append([seis, fs, bs] + [o.image(q) for q in range(5)])
        finally:
            if ctx:
                ctx.__exit__(None, None, None)
    for a, b in zip(*out):
        assert np.array_equal(a, b)
