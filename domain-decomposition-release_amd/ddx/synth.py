"""Synthetic, seeded inputs for the workloads the engine is measured on.

Random numbers come from splitmix64 (a counter-based generator, so any
element can be regenerated from (seed, index)) followed by Box-Muller.
Arrays are float32 with shape (nz, ny, nx): x fastest in memory, as the
reference's Field (FLD/jarvis_0_frame_bones.hpp:9-51).
"""
import math

import numpy as np

_M64 = (1 << 64) - 1


def splitmix64(seed, n):
    """n consecutive splitmix64 outputs for `seed` (uint64)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & _M64) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n):
    """U(0,1) doubles in (0, 1)."""
    z = splitmix64(seed, n)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / (1 << 53))


def normal(seed, n):
    """N(0,1) doubles by Box-Muller."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(2 * math.pi * u2)
    out[1::2] = r * np.sin(2 * math.pi * u2)
    return out[:n]


def gardner(vp):
    """rho = 310 vp^0.25 (kg/m^3, vp in m/s)."""
    return 310.0 * np.power(vp, 0.25)


def homogeneous_noise(nx, ny, nz, vp0=3000.0, rel=0.01, seed=1):
    """vp = vp0 (1 + rel N), vs = vp/sqrt(3) (1 + rel N'), rho = Gardner (1 + rel N''),
    three independent streams derived from `seed`."""
    n = nx * ny * nz
    vp = vp0 * (1.0 + rel * normal(seed * 3 + 0, n))
    vs = vp / math.sqrt(3.0) * (1.0 + rel * normal(seed * 3 + 1, n))
    rho = gardner(vp) * (1.0 + rel * normal(seed * 3 + 2, n))
    shp = (nz, ny, nx)
    return (vp.reshape(shp).astype(np.float32), vs.reshape(shp).astype(np.float32),
            rho.reshape(shp).astype(np.float32))


def layered(nx, ny, nz, seed, nlayer=(6, 12), vp_range=(1800.0, 4500.0), vs_ratio=1.8,
            anomaly_radius=30, anomaly_amp=0.10, anomaly_center=None, hetero_rel=0.0, hetero_len=0):
    """Flat layers (count uniform in nlayer, interfaces uniform in depth, vp
    uniform in vp_range), vs = vp / vs_ratio, Gardner density, plus a smooth
    Gaussian vp anomaly of the given radius (cells) and optional correlated
    Gaussian heterogeneity (relative std hetero_rel, correlation length
    hetero_len cells, by separable Gaussian filtering of white noise)."""
    u = uniform(seed, 64)
    nl = int(nlayer[0] + math.floor(u[0] * (nlayer[1] - nlayer[0] + 1)))
    nl = min(nl, nlayer[1])
    cuts = np.sort(np.floor(u[1:nl] * nz).astype(int))
    vps = vp_range[0] + u[32:32 + nl] * (vp_range[1] - vp_range[0])
    prof = np.empty(nz)
    edges = [0] + list(cuts) + [nz]
    for li in range(nl):
        prof[edges[li]:edges[li + 1]] = vps[li]
    vp = np.broadcast_to(prof[:, None, None], (nz, ny, nx)).astype(np.float64).copy()
    if anomaly_radius > 0 and anomaly_amp != 0:
        c = anomaly_center or (nx / 2.0, ny / 2.0, nz / 2.0)
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
        vp *= 1.0 + anomaly_amp * np.exp(-r2 / (2.0 * (anomaly_radius / 2.0) ** 2)) * (r2 <= anomaly_radius ** 2 * 4)
    if hetero_rel > 0 and hetero_len > 0:
        w = normal(seed + 7919, nx * ny * nz).reshape(nz, ny, nx)
        w = _gauss_filter3(w, hetero_len)
        w /= max(w.std(), 1e-30)
        vp *= 1.0 + hetero_rel * w
    vs = vp / vs_ratio
    rho = gardner(vp)
    return vp.astype(np.float32), vs.astype(np.float32), rho.astype(np.float32)


def _gauss_filter3(a, L):
    """Separable Gaussian smoothing with standard deviation L cells."""
    try:
        from scipy.ndimage import gaussian_filter
        return gaussian_filter(a, L, mode="constant", truncate=3.0)
    except ImportError:
        pass
    rad = int(3 * L)
    t = np.arange(-rad, rad + 1)
    k = np.exp(-0.5 * (t / L) ** 2)
    k /= k.sum()
    out = a
    for ax in range(3):
        out = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), ax, out)
    return out


def clip_for_stability(vp, vs, rho, dt, dx, safety=0.97):
    """Clip vp (and vs, rho consistently) so that vmax * dt <= safety * dt_cal,
    with dt_cal the 8th-order staggered limit 1 / (sum|c| * sqrt(3)/dx * vmax).
    Returns (vp, vs, rho, number_of_clipped_cells)."""
    csum = 1225 / 1024 + 245 / 3072 + 49 / 5120 + 5 / 7168
    vlim = safety / (csum * math.sqrt(3.0) / dx * dt)
    mask = vp > vlim
    n = int(mask.sum())
    if n:
        scale = np.where(mask, vlim / vp, 1.0).astype(np.float32)
        vp = (vp * scale).astype(np.float32)
        vs = (vs * scale).astype(np.float32)
        rho = np.where(mask, gardner(vp.astype(np.float64)), rho).astype(np.float32)
    return vp, vs, rho, n


# ---- the measured workloads ----
def config1():
    """100^3 homogeneous-plus-noise, 10 CPML layers, 2 shots, 8 receivers."""
    vp, vs, rho = homogeneous_noise(100, 100, 100, seed=1)
    shots = np.array([[150.0, 150.0, 150.0], [250.0, 250.0, 250.0]], np.float32)
    rec = np.array([[x, 150.0, 150.0] for x in range(150, 851, 100)], np.float32)
    return dict(nx=100, ny=100, nz=100, dx=10.0, pml=10, dt=0.001, fm=15.0, ntime=1000, divide=(1, 1, 1),
                vp=vp, vs=vs, rho=rho, shots=shots, recv=rec)


def config2(rec_width=1, seed=2):
    """256x256x200 layered RTM model with a Gaussian anomaly, 20 CPML layers,
    2x2x1 subdomains, 2000 steps, one surface shot, 128x128 receivers."""
    nx, ny, nz, dx, dt = 256, 256, 200, 10.0, 0.001
    vp, vs, rho = layered(nx, ny, nz, seed)
    vp, vs, rho, nclip = clip_for_stability(vp, vs, rho, dt, dx)
    shots = np.array([[128 * dx, 128 * dx, 5 * dx]], np.float32)
    g = np.arange(128) * 2 * dx
    xx, yy = np.meshgrid(g, g, indexing="xy")
    rec = np.stack([xx.ravel(), yy.ravel(), np.full(xx.size, 5 * dx)], axis=1).astype(np.float32)
    return dict(nx=nx, ny=ny, nz=nz, dx=dx, pml=20, dt=dt, fm=15.0, ntime=2000, divide=(2, 2, 1),
                rec_width=rec_width, vp=vp, vs=vs, rho=rho, shots=shots, recv=rec, clipped=nclip)


def config4(crop=0, seed=4):
    """676x676x210 layered model plus 2% correlated Gaussian heterogeneity
    (correlation length 8 cells), dx = 12.5 m, 20 CPML layers, 3000 steps,
    dt = 1 ms, fm = 12 Hz, one surface shot at the centre, receivers every
    16 cells at z = 5.  crop = n: the central n x n block of the model, its
    top n cells in z (the grid on which the FP64 oracle is compared), with
    the shot and receivers at the same places relative to the block.
    Images accumulate in FP64."""
    nx, ny, nz, dx, dt = 676, 676, 210, 12.5, 0.001
    vp, vs, rho = layered(nx, ny, nz, seed, nlayer=(6, 12), vp_range=(1800.0, 4000.0), anomaly_radius=40,
                          hetero_rel=0.02, hetero_len=8)
    if crop:
        x0, y0 = (nx - crop) // 2, (ny - crop) // 2
        sl = (slice(0, min(crop, nz)), slice(y0, y0 + crop), slice(x0, x0 + crop))
        vp, vs, rho = (np.ascontiguousarray(a[sl]) for a in (vp, vs, rho))
        nz, ny, nx = vp.shape
    vp, vs, rho, nclip = clip_for_stability(vp, vs, rho, dt, dx)
    shots = np.array([[(nx // 2) * dx, (ny // 2) * dx, 5 * dx]], np.float32)
    gx = np.arange(8, nx - 8, 16) * dx
    gy = np.arange(8, ny - 8, 16) * dx
    xx, yy = np.meshgrid(gx, gy, indexing="xy")
    rec = np.stack([xx.ravel(), yy.ravel(), np.full(xx.size, 5 * dx)], axis=1).astype(np.float32)
    return dict(nx=nx, ny=ny, nz=nz, dx=dx, pml=20, dt=dt, fm=12.0, ntime=3000, divide=(1, 1, 1),
                rec_width=1, image_fp64=1, vp=vp, vs=vs, rho=rho, shots=shots, recv=rec, clipped=nclip)


def config3(nshot=64, seed=3):
    """Batch of independent modelling shots: per-shot grids with each side
    uniform in [96, 320] cells, dx = 10 m, 10-20 CPML layers, layered models
    as in config2 (stability-clipped), 1500 steps, dt = 1 ms, fm uniform in
    [10, 20] Hz, one surface shot at the centre and a receiver line along x."""
    u = uniform(seed, 8 * nshot).reshape(nshot, 8)
    shots = []
    for i in range(nshot):
        nx, ny, nz = (int(96 + math.floor(u[i, a] * 225)) for a in range(3))
        pml = int(10 + math.floor(u[i, 3] * 11))
        fm = float(10.0 + 10.0 * u[i, 4])
        dx, dt = 10.0, 0.001
        vp, vs, rho = layered(nx, ny, nz, seed * 1000 + i, anomaly_radius=max(4, min(nx, ny, nz) // 8))
        vp, vs, rho, _ = clip_for_stability(vp, vs, rho, dt, dx)
        shot = np.array([[(nx // 2) * dx, (ny // 2) * dx, 5 * dx]], np.float32)
        xs = np.arange(4, nx - 4, 8) * dx
        rec = np.stack([xs, np.full(xs.size, (ny // 2) * dx), np.full(xs.size, 5 * dx)], 1).astype(np.float32)
        shots.append(dict(nx=nx, ny=ny, nz=nz, dx=dx, pml=pml, dt=dt, fm=fm, ntime=1500, vp=vp, vs=vs, rho=rho,
                          shots=shot, recv=rec))
    return shots
