"""Python binding of libddx (the MI355X elastic FDTD / RTM engine).

Thin ctypes layer over the C ABI in include/ddx.h.  There is no CPU
fallback: if libddx.so is missing or no GPU is present, the calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG, "lib", "libddx.so")

VX, VY, VZ, SXX, SYY, SZZ, SXY, SXZ, SYZ, SAU, VXP, VYP, VZP, UX, UY, UZ, VXRT, VYRT, VZRT = range(19)
FIELD_NAMES = ["vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz", "sau", "vxp", "vyp", "vzp",
               "ux", "uy", "uz", "vxrt", "vyrt", "vzrt"]
NFIELD = 19
PURE_FORWARD, RTM_FORWARD, RTM_BACKWARD, RTM_REVERSE, FWI = range(5)
IP, IPP, IPS, ISP, ISS, IPP_LAP, IPS_LAP, ISP_LAP, ISS_LAP = range(9)
K_VEL, K_STRESS, K_VEL_PML, K_STRESS_PML, K_CMM, K_IMAGE = range(6)
STORE_FP32, STORE_FP16, STORE_BF16, STORE_BFP16 = range(4)
# bfp16: 16-bit codes with per-run exponents, for ring_storage only
STORAGE_NAMES = {"fp32": STORE_FP32, "fp16": STORE_FP16, "bf16": STORE_BF16, "bfp16": STORE_BFP16}
CLASS_VELOCITY, CLASS_STRESS = range(2)

EXPORTED = [
    "ddx_create", "ddx_destroy", "ddx_last_error", "ddx_set_model", "ddx_set_wavelet", "ddx_set_geometry",
    "ddx_receivers", "ddx_model_shot", "ddx_get_seismograms", "ddx_set_observed", "ddx_rtm_shot",
    "ddx_finalize_images", "ddx_reset_images", "ddx_get_image", "ddx_reset_wave", "ddx_step_vel",
    "ddx_model_steps", "ddx_rtm_steps", "ddx_check_finite", "ddx_step_stress", "ddx_update_vel", "ddx_update_stress", "ddx_halo_exchange", "ddx_vowr_save",
    "ddx_vowr_inject", "ddx_recv_inject", "ddx_image_step", "ddx_source", "ddx_get_field", "ddx_set_field", "ddx_synchronize", "ddx_profile",
    "ddx_kernel_time", "ddx_kernel_class_time", "ddx_cell_updates", "ddx_device_bytes", "ddx_subdomain", "ddx_fd_coef", "ddx_ntime",
    "ddx_ricker", "ddx_stable_dt", "ddx_set_tiling", "ddx_set_tiling_mode",
    "ddx_set_snapshots", "ddx_get_snapshot", "ddx_set_migration_model", "ddx_smooth_model",
    "ddx_record_synthetic", "ddx_get_synthetic", "ddx_inject_residual", "ddx_estimate_bytes",
    "ddx_set_storage_scale", "ddx_get_storage_scale", "ddx_set_ring_scaling", "ddx_set_brick", "ddx_get_brick", "ddx_autotune",
    "ddx_set_graphs", "ddx_model_batch", "ddx_rtm_batch", "ddx_batch_create", "ddx_batch_add", "ddx_batch_run", "ddx_batch_size", "ddx_batch_remove",
    "ddx_batch_last_error", "ddx_batch_destroy", "ddx_set_exchange", "ddx_get_exchange",
    "ddx_batch_add_rtm", "ddx_model_batch_noalloc",
]


class Desc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("dz", ctypes.c_float),
        ("pml_num", ctypes.c_int),
        ("x_divide_num", ctypes.c_int), ("y_divide_num", ctypes.c_int), ("z_divide_num", ctypes.c_int),
        ("dt", ctypes.c_float), ("fm", ctypes.c_float), ("ntime", ctypes.c_int),
        ("rec_width", ctypes.c_int), ("image_fp64", ctypes.c_int), ("source_mode", ctypes.c_int),
        ("device", ctypes.c_int), ("cfl_policy", ctypes.c_int), ("storage", ctypes.c_int),
        ("ring_storage", ctypes.c_int),
    ]


class DdxError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libddx.so (raises if it has not been built).  DDX_LIBRARY may
    name another build of it (A/B measurements of kernel variants)."""
    global _lib
    if _lib is None:
        path = os.environ.get("DDX_LIBRARY", LIB_PATH)
        if not os.path.exists(path):
            raise DdxError(f"{path} not found: build it with domain-decomposition-release_amd/build.py")
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        ip = ctypes.POINTER(ctypes.c_int)
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        i = ctypes.c_int
        L.ddx_create.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(vp)]
        L.ddx_destroy.argtypes = [vp]
        L.ddx_destroy.restype = None
        L.ddx_last_error.argtypes = [vp]
        L.ddx_last_error.restype = ctypes.c_char_p
        L.ddx_set_model.argtypes = [vp, fp, fp, fp]
        L.ddx_set_wavelet.argtypes = [vp, fp]
        L.ddx_set_geometry.argtypes = [vp, fp, i, fp, i]
        L.ddx_receivers.argtypes = [vp, ip, ip, ip]
        L.ddx_model_shot.argtypes = [vp, i]
        L.ddx_get_seismograms.argtypes = [vp, i, fp]
        L.ddx_set_observed.argtypes = [vp, i, fp]
        L.ddx_rtm_shot.argtypes = [vp, i]
        L.ddx_finalize_images.argtypes = [vp]
        L.ddx_reset_images.argtypes = [vp]
        L.ddx_get_image.argtypes = [vp, i, i, fp]
        L.ddx_reset_wave.argtypes = [vp, i]
        L.ddx_step_vel.argtypes = [vp, i, i]
        L.ddx_step_stress.argtypes = [vp, i, i]
        for name in ("ddx_update_vel", "ddx_update_stress", "ddx_halo_exchange"):
            getattr(L, name).argtypes = [vp, i, i]
        L.ddx_model_steps.argtypes = [vp, i, i, i]
        L.ddx_rtm_steps.argtypes = [vp, i, i, i]
        L.ddx_check_finite.argtypes = [vp, ip]
        L.ddx_vowr_save.argtypes = [vp, i]
        L.ddx_vowr_inject.argtypes = [vp, i, i]
        L.ddx_recv_inject.argtypes = [vp, i, i, i]
        L.ddx_image_step.argtypes = [vp]
        L.ddx_source.argtypes = [vp, i, i, i, i]
        L.ddx_get_field.argtypes = [vp, i, i, fp]
        L.ddx_set_field.argtypes = [vp, i, i, fp]
        L.ddx_synchronize.argtypes = [vp]
        L.ddx_profile.argtypes = [vp, i]
        L.ddx_kernel_time.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_longlong), dp]
        L.ddx_kernel_class_time.argtypes = [vp, i, dp, ctypes.POINTER(ctypes.c_longlong), dp]
        L.ddx_cell_updates.argtypes = [vp, dp]
        L.ddx_device_bytes.argtypes = [vp, dp]
        L.ddx_subdomain.argtypes = [ctypes.POINTER(Desc), i, ip]
        L.ddx_fd_coef.argtypes = [i, fp]
        L.ddx_ntime.argtypes = [ctypes.c_float, ctypes.c_float]
        L.ddx_ricker.argtypes = [ctypes.c_float, ctypes.c_float, i, fp]
        L.ddx_stable_dt.argtypes = [ctypes.c_float] * 4
        L.ddx_stable_dt.restype = ctypes.c_float
        L.ddx_set_tiling.argtypes = [vp, i, i, i, i]
        L.ddx_set_tiling_mode.argtypes = [vp, i, i, i, i, i, i]
        L.ddx_set_snapshots.argtypes = [vp, i, i]
        L.ddx_set_migration_model.argtypes = [vp, fp, fp, fp]
        L.ddx_smooth_model.argtypes = [fp, fp, i, i, i, i, i, i]
        L.ddx_record_synthetic.argtypes = [vp, i, i, i]
        L.ddx_get_synthetic.argtypes = [vp, i, fp]
        L.ddx_inject_residual.argtypes = [vp, i, i, i]
        L.ddx_estimate_bytes.argtypes = [ctypes.POINTER(Desc), i, i, i, i, ctypes.POINTER(ctypes.c_double)]
        L.ddx_get_snapshot.argtypes = [vp, i, i, fp]
        L.ddx_set_storage_scale.argtypes = [vp, i, i]
        L.ddx_get_storage_scale.argtypes = [vp, i, ip]
        L.ddx_set_brick.argtypes = [vp, i, i, i, i, i, i, i]
        L.ddx_get_brick.argtypes = [vp, i, i, ip]
        L.ddx_autotune.argtypes = [vp, i, ip]
        L.ddx_set_graphs.argtypes = [vp, i]
        L.ddx_set_exchange.argtypes = [vp, i]
        L.ddx_get_exchange.argtypes = [vp, ip]
        L.ddx_model_batch.argtypes = [ctypes.POINTER(vp), ip, i]
        L.ddx_rtm_batch.argtypes = [ctypes.POINTER(vp), ip, i]
        L.ddx_model_batch_noalloc.argtypes = [ctypes.POINTER(vp), ip, i]
        L.ddx_set_ring_scaling.argtypes = [vp, i]
        L.ddx_batch_create.argtypes = [ctypes.POINTER(vp)]
        L.ddx_batch_add.argtypes = [vp, vp, i]
        L.ddx_batch_add_rtm.argtypes = [vp, vp, i]
        L.ddx_batch_run.argtypes = [vp, i, ctypes.POINTER(vp), i, ip]
        L.ddx_batch_size.argtypes = [vp, ip]
        L.ddx_batch_remove.argtypes = [vp, vp]
        L.ddx_batch_last_error.argtypes = [vp]
        L.ddx_batch_last_error.restype = ctypes.c_char_p
        L.ddx_batch_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _i(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def storage_id(storage):
    """Storage format id from a name ("fp32", "fp16", "bf16", "bfp16") or an id."""
    if isinstance(storage, str):
        if storage not in STORAGE_NAMES:
            raise ValueError(f"storage must be one of {sorted(STORAGE_NAMES)}")
        return STORAGE_NAMES[storage]
    return int(storage)


def make_desc(nx, ny, nz, dx, pml, dt, fm, ntime, divide=(1, 1, 1), rec_width=1, image_fp64=0,
              source_mode=0, device=0, cfl_policy=0, dy=None, dz=None, storage=STORE_FP32,
              ring_storage=STORE_FP32):
    """ring_storage: VOWR boundary store format over FP32 wavefields
    (STORE_FP32 = as `storage`; "fp16", "bf16", or "bfp16": 16-bit codes
    with a 3-bit exponent and a 12-bit mantissa relative to one exponent per
    16 consecutive store entries)."""
    return Desc(nx, ny, nz, dx, dy or dx, dz or dx, pml, divide[0], divide[1], divide[2], dt, fm, ntime,
                rec_width, image_fp64, source_mode, device, cfl_policy, storage_id(storage),
                storage_id(ring_storage))


# ---- host-only helpers (no GPU) ----
def subdomain(desc, rank):
    out = np.zeros(21, np.int32)
    rc = lib().ddx_subdomain(ctypes.byref(desc), rank, _i(out))
    if rc:
        raise DdxError(f"ddx_subdomain failed ({rc})")
    return {
        "phy_n": tuple(out[0:3]), "phy_off": tuple(out[3:6]), "ext_n": tuple(out[6:9]),
        "margin": tuple(out[9:15]), "neighbour": tuple(out[15:21]),
    }


def fd_coef(half):
    out = np.zeros(half, np.float32)
    if lib().ddx_fd_coef(half, _f(out)):
        raise DdxError("ddx_fd_coef failed")
    return out


def ntime(T, dt):
    return lib().ddx_ntime(T, dt)


def model_batch(engines, ishots=None):
    """Modelling shots of several engines in one launch sequence
    (ddx_model_batch): shot ishots[q] (default 0) of engines[q]; the
    seismograms equal each engine's own model_shot."""
    n = len(engines)
    if n == 0:
        return
    hs = (ctypes.c_void_p * n)(*[e.h.value for e in engines])
    sh = np.asarray(ishots if ishots is not None else [0] * n, np.int32)
    if sh.shape != (n,):
        raise ValueError("one shot index per engine")
    rc = lib().ddx_model_batch(hs, _i(sh), n)
    if rc:
        raise DdxError(f"ddx_model_batch failed ({rc}): {lib().ddx_last_error(engines[0].h).decode()}")


def rtm_batch(engines, ishots=None):
    """RTM shots of several engines in one launch sequence (ddx_rtm_batch):
    shot ishots[q] (default 0) of engines[q], observed data set; each
    engine's images equal its own rtm_shot's."""
    n = len(engines)
    if n == 0:
        return
    hs = (ctypes.c_void_p * n)(*[e.h.value for e in engines])
    sh = np.asarray(ishots if ishots is not None else [0] * n, np.int32)
    if sh.shape != (n,):
        raise ValueError("one shot index per engine")
    rc = lib().ddx_rtm_batch(hs, _i(sh), n)
    if rc:
        raise DdxError(f"ddx_rtm_batch failed ({rc}): {lib().ddx_last_error(engines[0].h).decode()}")


class Batch:
    """Continuous batch of modelling and RTM shots (ddx_batch):
    add(engine, ishot) (a modelling shot) or add(engine, ishot, rtm=True)
    (an RTM shot whose observed data are set) at any time; run(max_steps)
    advances every member and returns the engines whose shots finished
    (their seismograms, or images, ready).  An engine belongs to the batch
    from add() until run() returns it."""

    def __init__(self):
        self.h = ctypes.c_void_p()
        rc = lib().ddx_batch_create(ctypes.byref(self.h))
        if rc:
            raise DdxError(f"ddx_batch_create failed ({rc})")
        self._members = {}

    def _check(self, rc, what):
        if rc:
            raise DdxError(f"{what} failed ({rc}): {lib().ddx_batch_last_error(self.h).decode()}")

    def add(self, engine, ishot=0, rtm=False):
        if rtm:
            self._check(lib().ddx_batch_add_rtm(self.h, engine.h, int(ishot)), "ddx_batch_add_rtm")
        else:
            self._check(lib().ddx_batch_add(self.h, engine.h, int(ishot)), "ddx_batch_add")
        self._members[engine.h.value] = engine

    def run(self, max_steps):
        cap = max(len(self._members), 1)
        done = (ctypes.c_void_p * cap)()
        nd = ctypes.c_int()
        self._check(lib().ddx_batch_run(self.h, int(max_steps), done, cap, ctypes.byref(nd)), "ddx_batch_run")
        return [self._members.pop(done[q]) for q in range(nd.value)]

    def remove(self, engine):
        """Take a member out before its shot ended; it is the caller's again."""
        self._check(lib().ddx_batch_remove(self.h, engine.h), "ddx_batch_remove")
        self._members.pop(engine.h.value, None)

    def __len__(self):
        n = ctypes.c_int()
        self._check(lib().ddx_batch_size(self.h, ctypes.byref(n)), "ddx_batch_size")
        return n.value

    def close(self):
        if self.h:
            lib().ddx_batch_destroy(self.h)
            self.h = None
            self._members = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def estimate_bytes(desc, nrec, nshot=1, rtm=True, migration=False):
    """Device bytes a context for desc will hold (ddx_estimate_bytes; no GPU)."""
    out = ctypes.c_double()
    rc = lib().ddx_estimate_bytes(ctypes.byref(desc), int(nrec), int(nshot), int(bool(rtm)), int(bool(migration)),
                                  ctypes.byref(out))
    if rc:
        raise DdxError(f"ddx_estimate_bytes failed ({rc}): {lib().ddx_last_error(None).decode()}")
    return out.value


def smooth_model(a, radius):
    """Box smoothing of a [nz][ny][nx] float32 grid with half-widths
    radius = (rx, ry, rz) (ddx_smooth_model; no GPU needed)."""
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty_like(a)
    nz, ny, nx = a.shape
    rc = lib().ddx_smooth_model(_f(a), _f(out), nx, ny, nz, *[int(r) for r in radius])
    if rc:
        raise DdxError(f"ddx_smooth_model failed ({rc})")
    return out


def ricker(fm, dt, nt):
    out = np.zeros(nt, np.float32)
    if lib().ddx_ricker(fm, dt, nt, _f(out)):
        raise DdxError("ddx_ricker failed")
    return out


def stable_dt(vmax, dx, dy=None, dz=None):
    return lib().ddx_stable_dt(vmax, dx, dy or dx, dz or dx)


class Engine:
    """One engine context (all subdomains of the decomposition on one GPU)."""

    def __init__(self, desc):
        self.desc = desc
        self.h = ctypes.c_void_p()
        rc = lib().ddx_create(ctypes.byref(desc), ctypes.byref(self.h))
        if rc:
            raise DdxError(f"ddx_create failed ({rc}): {lib().ddx_last_error(None).decode()}")
        self.nsub = desc.x_divide_num * desc.y_divide_num * desc.z_divide_num
        m = desc.pml_num
        self.ext_shape = (desc.nz + 2 * m, desc.ny + 2 * m, desc.nx + 2 * m)

    def _check(self, rc, what):
        if rc:
            raise DdxError(f"{what} failed ({rc}): {lib().ddx_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            lib().ddx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_model(self, vp, vs, rho):
        d = self.desc
        arrs = [np.ascontiguousarray(a, np.float32) for a in (vp, vs, rho)]
        for a in arrs:
            assert a.shape == (d.nz, d.ny, d.nx), a.shape
        self._check(lib().ddx_set_model(self.h, *[_f(a) for a in arrs]), "ddx_set_model")

    def set_wavelet(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.shape == (self.desc.ntime,)
        self._check(lib().ddx_set_wavelet(self.h, _f(w)), "ddx_set_wavelet")

    def set_geometry(self, shots, recv):
        s = np.ascontiguousarray(np.asarray(shots, np.float32).reshape(-1, 3))
        r = np.ascontiguousarray(np.asarray(recv, np.float32).reshape(-1, 3))
        self._check(lib().ddx_set_geometry(self.h, _f(s), len(s), _f(r), len(r)), "ddx_set_geometry")
        self.nshot = len(s)

    def receivers(self):
        n = ctypes.c_int()
        lib().ddx_receivers(self.h, ctypes.byref(n), None, None)
        rank = np.zeros(max(n.value, 1), np.int32)
        inp = np.zeros(max(n.value, 1), np.int32)
        lib().ddx_receivers(self.h, None, _i(rank), _i(inp))
        return rank[: n.value], inp[: n.value]

    @property
    def nrec(self):
        n = ctypes.c_int()
        lib().ddx_receivers(self.h, ctypes.byref(n), None, None)
        return n.value

    def model_shot(self, ishot):
        self._check(lib().ddx_model_shot(self.h, ishot), "ddx_model_shot")

    def seismograms(self, ishot):
        out = np.zeros((3, self.nrec, self.desc.ntime), np.float32)
        self._check(lib().ddx_get_seismograms(self.h, ishot, _f(out)), "ddx_get_seismograms")
        return out

    def set_observed(self, ishot, obs):
        o = np.ascontiguousarray(obs, np.float32)
        assert o.shape == (3, self.nrec, self.desc.ntime)
        self._check(lib().ddx_set_observed(self.h, ishot, _f(o)), "ddx_set_observed")

    def rtm_shot(self, ishot):
        self._check(lib().ddx_rtm_shot(self.h, ishot), "ddx_rtm_shot")

    def finalize_images(self):
        self._check(lib().ddx_finalize_images(self.h), "ddx_finalize_images")

    def reset_images(self):
        self._check(lib().ddx_reset_images(self.h), "ddx_reset_images")

    def image(self, which, rank):
        sd = subdomain(self.desc, rank)
        nx, ny, nz = sd["phy_n"]
        out = np.zeros((nz, ny, nx), np.float32)
        self._check(lib().ddx_get_image(self.h, which, rank, _f(out)), "ddx_get_image")
        return out

    def reset_wave(self, s):
        self._check(lib().ddx_reset_wave(self.h, s), "ddx_reset_wave")

    def step_vel(self, s, mode):
        """Velocity half step of set s, then its CMM halo exchange."""
        self._check(lib().ddx_step_vel(self.h, s, mode), "ddx_step_vel")

    def step_stress(self, s, mode):
        self._check(lib().ddx_step_stress(self.h, s, mode), "ddx_step_stress")

    def model_steps(self, ishot, s0, n):
        """Stages s0 .. s0+n-1 of modelling shot ishot (its time steps)."""
        self._check(lib().ddx_model_steps(self.h, int(ishot), int(s0), int(n)), "ddx_model_steps")

    def rtm_steps(self, ishot, s0, n):
        """Stages s0 .. s0+n-1 of RTM shot ishot (2 ntime stages: forward
        steps, then reverse iterations)."""
        self._check(lib().ddx_rtm_steps(self.h, int(ishot), int(s0), int(n)), "ddx_rtm_steps")

    def finite(self):
        """False when a velocity of either wavefield set is infinite or NaN."""
        out = ctypes.c_int()
        self._check(lib().ddx_check_finite(self.h, ctypes.byref(out)), "ddx_check_finite")
        return bool(out.value)

    def update_vel(self, s, mode):
        """Velocity half step without the CMM exchange."""
        self._check(lib().ddx_update_vel(self.h, s, mode), "ddx_update_vel")

    def update_stress(self, s, mode):
        """Stress half step without the CMM exchange."""
        self._check(lib().ddx_update_stress(self.h, s, mode), "ddx_update_stress")

    def halo_exchange(self, s, batch):
        """CMM exchange of one batch (0 velocities, 1 stresses, 2 stresses + sau)."""
        self._check(lib().ddx_halo_exchange(self.h, s, batch), "ddx_halo_exchange")

    def vowr_save(self, it):
        self._check(lib().ddx_vowr_save(self.h, it), "ddx_vowr_save")

    def vowr_inject(self, which, it):
        """which 1: velocities, 2: stresses of boundary slice it."""
        self._check(lib().ddx_vowr_inject(self.h, which, it), "ddx_vowr_inject")

    def recv_inject(self, s, ishot, it):
        self._check(lib().ddx_recv_inject(self.h, s, ishot, it), "ddx_recv_inject")

    def image_step(self):
        self._check(lib().ddx_image_step(self.h), "ddx_image_step")

    def source(self, s, ishot, it, sign=1):
        self._check(lib().ddx_source(self.h, s, ishot, it, sign), "ddx_source")

    def get(self, s, field):
        out = np.zeros(self.ext_shape, np.float32)
        self._check(lib().ddx_get_field(self.h, s, field, _f(out)), "ddx_get_field")
        return out

    def set(self, s, field, arr):
        a = np.ascontiguousarray(arr, np.float32)
        assert a.shape == self.ext_shape
        self._check(lib().ddx_set_field(self.h, s, field, _f(a)), "ddx_set_field")

    def synchronize(self):
        self._check(lib().ddx_synchronize(self.h), "ddx_synchronize")

    def set_tiling(self, tx, ty, cy, kz):
        self._check(lib().ddx_set_tiling(self.h, tx, ty, cy, kz), "ddx_set_tiling")

    def set_migration_model(self, vp, vs, rho):
        """Model for the RTM passes (None removes it)."""
        if vp is None:
            self._check(lib().ddx_set_migration_model(self.h, None, None, None), "ddx_set_migration_model")
            return
        a = [np.ascontiguousarray(x, np.float32) for x in (vp, vs, rho)]
        self._check(lib().ddx_set_migration_model(self.h, *[_f(x) for x in a]), "ddx_set_migration_model")

    def record_synthetic(self, s, ishot, it):
        self._check(lib().ddx_record_synthetic(self.h, s, ishot, it), "ddx_record_synthetic")

    def synthetic(self, ishot):
        out = np.zeros((3, self.nrec, self.desc.ntime), np.float32)
        self._check(lib().ddx_get_synthetic(self.h, ishot, _f(out)), "ddx_get_synthetic")
        return out

    def inject_residual(self, s, ishot, it):
        self._check(lib().ddx_inject_residual(self.h, s, ishot, it), "ddx_inject_residual")

    def set_snapshots(self, field, every):
        """Keep `field` of the source-side set every `every` steps of the next
        RTM shots (forward and backward pass); every = 0 frees them."""
        self._check(lib().ddx_set_snapshots(self.h, field, every), "ddx_set_snapshots")
        self._snap = (field, every)

    def snapshot(self, pass_, index):
        """Snapshot `index` of pass 0 (forward) or 1 (backward), [nz][ny][nx] gridphy."""
        d = self.desc
        out = np.zeros((d.nz, d.ny, d.nx), np.float32)
        self._check(lib().ddx_get_snapshot(self.h, pass_, index, _f(out)), "ddx_get_snapshot")
        return out

    def set_storage_scale(self, field_class, log2_scale):
        """Power-of-two scale of a field class (CLASS_VELOCITY / CLASS_STRESS)
        in the 16-bit storage formats (over FP32 wavefields with a 16-bit
        ring_storage: the scale of the VOWR store's values)."""
        self._check(lib().ddx_set_storage_scale(self.h, field_class, int(log2_scale)), "ddx_set_storage_scale")

    def set_ring_scaling(self, mode):
        """16-bit VOWR store over FP32 wavefields: 0 class scales, 1 one
        scale per field and time slice."""
        self._check(lib().ddx_set_ring_scaling(self.h, int(mode)), "ddx_set_ring_scaling")

    def storage_scale(self, field_class):
        out = ctypes.c_int()
        self._check(lib().ddx_get_storage_scale(self.h, field_class, ctypes.byref(out)), "ddx_get_storage_scale")
        return out.value

    def set_brick(self, kernel, mode, tx, ty, cy, kz, px=1):
        """Brick shape of one kernel (0 velocity, 1 stress) and mode; px is
        the lane form: 1 one cell per thread, 2 paired bricks (two x cells
        per thread), 3 one cell per thread with the x arm from neighbouring
        lanes (tx = 64, FP32 storage), 4 form 3 with the y arm from an LDS
        tile of the block's rows."""
        self._check(lib().ddx_set_brick(self.h, kernel, mode, tx, ty, cy, kz, px), "ddx_set_brick")

    def brick(self, kernel, mode):
        out = np.zeros(5, np.int32)
        self._check(lib().ddx_get_brick(self.h, kernel, mode, _i(out)), "ddx_get_brick")
        return tuple(int(v) for v in out)

    def autotune(self, reps=3):
        """Pick the fastest brick per kernel and mode on this context (resets
        the wavefields); returns the number of shapes timed."""
        n = ctypes.c_int()
        self._check(lib().ddx_autotune(self.h, int(reps), ctypes.byref(n)), "ddx_autotune")
        return n.value

    def set_graphs(self, on=True):
        """Replay the modelling time step from a captured HIP graph."""
        self._check(lib().ddx_set_graphs(self.h, int(bool(on))), "ddx_set_graphs")

    def set_exchange(self, mode):
        """CMM exchange mode: 0 direct copy into the ghosts, 1 contiguous
        message blocks (pack, then unpack)."""
        self._check(lib().ddx_set_exchange(self.h, int(mode)), "ddx_set_exchange")

    def exchange(self):
        m = ctypes.c_int()
        self._check(lib().ddx_get_exchange(self.h, ctypes.byref(m)), "ddx_get_exchange")
        return m.value

    def set_tiling_mode(self, kernel, mode, tx, ty, cy, kz):
        """Brick shape of one kernel (0 velocity, 1 stress) in one mode."""
        self._check(lib().ddx_set_tiling_mode(self.h, kernel, mode, tx, ty, cy, kz), "ddx_set_tiling_mode")

    def profile(self, kernel):
        self._check(lib().ddx_profile(self.h, kernel), "ddx_profile")

    def kernel_time(self):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        b = ctypes.c_double()
        self._check(lib().ddx_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b)),
                    "ddx_kernel_time")
        return ms.value, n.value, b.value

    def class_time(self, kernel):
        """(ms, launches, algorithmic bytes) of one kernel class since
        profile() (profile(-1) times every class at once)."""
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        b = ctypes.c_double()
        self._check(lib().ddx_kernel_class_time(self.h, int(kernel), ctypes.byref(ms), ctypes.byref(n),
                                                ctypes.byref(b)), "ddx_kernel_class_time")
        return ms.value, n.value, b.value

    def cell_updates(self):
        c = ctypes.c_double()
        lib().ddx_cell_updates(self.h, ctypes.byref(c))
        return c.value

    def device_bytes(self):
        c = ctypes.c_double()
        lib().ddx_device_bytes(self.h, ctypes.byref(c))
        return c.value
