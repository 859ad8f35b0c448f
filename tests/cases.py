"""Small seeded cases shared by the CPU and GPU tests."""
import struct

import numpy as np

import pyoracle as po
from ddx import synth

F32 = np.float32


def small_model(nx, ny, nz, seed=11, layered=True):
    if layered:
        vp, vs, rho = synth.layered(nx, ny, nz, seed, nlayer=(3, 5), vp_range=(2000.0, 3200.0),
                                    anomaly_radius=max(2, min(nx, ny, nz) // 6), anomaly_amp=0.1)
    else:
        vp, vs, rho = synth.homogeneous_noise(nx, ny, nz, seed=seed)
    return vp, vs, rho


def case(nx=24, ny=22, nz=20, dx=10.0, pml=6, dt=0.001, fm=25.0, ntime=60, rec_width=1, seed=11,
         layered=True):
    vp, vs, rho = small_model(nx, ny, nz, seed, layered)
    shot = np.array([[(nx // 2) * dx, (ny // 2 - 1) * dx, 3 * dx]], F32)
    rx = np.arange(1, nx - 1, 3) * dx
    ry = np.arange(1, ny - 1, 4) * dx
    xx, yy = np.meshgrid(rx, ry, indexing="xy")
    recv = np.stack([xx.ravel(), yy.ravel(), np.full(xx.size, 2 * dx)], 1).astype(F32)
    return dict(nx=nx, ny=ny, nz=nz, dx=dx, pml=pml, dt=dt, fm=fm, ntime=ntime, rec_width=rec_width,
                vp=vp, vs=vs, rho=rho, shots=shot, recv=recv)


def oracle_for(c, prec=4):
    return po.Oracle(c["nx"], c["ny"], c["nz"], c["dx"], c["pml"], c["dt"], c["fm"], c["ntime"],
                     c["rec_width"], c["vp"], c["vs"], c["rho"], prec=prec)


def oracle_rows(c, divide):
    """(rank, input, global ext flat index) of the kept receivers, output order."""
    d = (c["dx"],) * 3
    subs = po.decompose((c["nx"], c["ny"], c["nz"]), d, c["pml"], divide)
    rows = po.kept_receivers(subs, c["recv"], d, c["pml"])
    NX, NY = c["nx"] + 2 * c["pml"], c["ny"] + 2 * c["pml"]
    return [(r, i, po.ext_index(n, NX, NY)) for (r, i, n) in rows], subs


def random_fields(shape, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return [(rng.standard_normal(shape) * scale).astype(F32) for _ in range(13)]


def gms_bytes(n, d, l, names, fields_xfast):
    """GMS file written from the layout alone: 240-byte header with
    para_num @74 (int16), l @76, n @88 (uint32), d @100 (float32); then per
    parameter a 64-byte name and the data z fastest, then x, then y."""
    h = bytearray(240)
    struct.pack_into("<h", h, 0, 1)
    h[2:2 + 5] = b"model"
    struct.pack_into("<h", h, 74, len(fields_xfast))
    struct.pack_into("<3f", h, 76, *l)
    struct.pack_into("<3I", h, 88, *n)
    struct.pack_into("<3f", h, 100, *d)
    out = bytes(h)
    for nm, f in zip(names, fields_xfast):
        out += nm.encode().ljust(64, b"\0")
        # f is [nz][ny][nx]; the file wants [ny][nx][nz] (z fastest, then x, then y)
        out += np.ascontiguousarray(np.transpose(f, (1, 2, 0))).astype("<f4").tobytes()
    return out
