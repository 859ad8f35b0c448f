"""ddx_model_batch_noalloc (no reference counterpart; for callers that
cannot allocate device memory) against ddx_model_batch through the C ABI:
two calls in a row, the second with a smaller batch of smaller grids, each
member's seismograms bit for bit equal."""
import ctypes

import numpy as np
import pytest

import ddx
from ddx import synth

pytestmark = pytest.mark.gpu


def engine(c, ntime):
    d = ddx.make_desc(c["nx"], c["ny"], c["nz"], c["dx"], c["pml"], c["dt"], c["fm"], ntime, cfl_policy=1)
    e = ddx.Engine(d)
    e.set_model(c["vp"], c["vs"], c["rho"])
    e.set_wavelet(ddx.ricker(c["fm"], c["dt"], ntime))
    e.set_geometry(c["shots"], c["recv"])
    return e


def noalloc(engines):
    L = ddx.lib()
    f = L.ddx_model_batch_noalloc
    f.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    n = len(engines)
    hs = (ctypes.c_void_p * n)(*[e.h.value for e in engines])
    sh = (ctypes.c_int * n)(*([0] * n))
    rc = f(hs, sh, n)
    if rc != 0:
        raise RuntimeError(f"ddx_model_batch_noalloc: {rc} {L.ddx_last_error(engines[0].h).decode()}")


def uneven(sizes, seed):
    out = []
    for q, (nx, ny, nz) in enumerate(sizes):
        vp, vs, rho = synth.homogeneous_noise(nx, ny, nz, seed=seed + q)
        shots = np.array([[nx // 2 * 10.0, ny // 2 * 10.0, 30.0]], np.float32)
        recv = np.array([[x * 10.0, ny // 3 * 10.0, 20.0] for x in range(2, nx - 2, 5)], np.float32)
        out.append(dict(nx=nx, ny=ny, nz=nz, dx=10.0, pml=8, dt=0.001, fm=20.0, vp=vp, vs=vs, rho=rho, shots=shots,
                        recv=recv))
    return out


@pytest.mark.parametrize("sizes", [[(40, 36, 30), (52, 30, 28), (34, 34, 34)], [(30, 26, 24), (28, 32, 22)]])
def test_noalloc_batch_equals_batch(sizes):
    nt = 60
    cs = uneven(sizes, seed=len(sizes))
    es = [engine(c, nt) for c in cs]
    try:
        for e in es:  # seismogram buffers and brick tables exist after one shot
            e.model_shot(0)
        ddx.model_batch(es)
        ref = [e.seismograms(0) for e in es]
        for e in es:
            e.reset_wave(0)
        noalloc(es)
        got = [e.seismograms(0) for e in es]
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
    finally:
        for e in es:
            e.close()


def test_noalloc_refuses_unprepared_context():
    c = uneven([(30, 26, 24)], seed=9)[0]
    e = engine(c, 20)
    try:
        with pytest.raises(RuntimeError, match="seismogram buffer"):
            noalloc([e])
    finally:
        e.close()
