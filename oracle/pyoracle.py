"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  The engine never does.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# DDX_ORACLE_ISA=v4 loads the AVX-512 build (same results; bench.py's CPU
# baseline uses it on hosts that have AVX-512)
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle_v4.so" if os.environ.get("DDX_ORACLE_ISA") == "v4"
                         else "liboracle.so")

# field ids (shared with the engine)
VX, VY, VZ, SXX, SYY, SZZ, SXY, SXZ, SYZ, SAU, VXP, VYP, VZP, UX, UY, UZ, VXRT, VYRT, VZRT = range(19)
NFIELD = 19
PURE_FORWARD, RTM_FORWARD, RTM_BACKWARD, RTM_REVERSE, FWI = range(5)


class OrcDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("dz", ctypes.c_float),
        ("pml", ctypes.c_int),
        ("dt", ctypes.c_float), ("fm", ctypes.c_float),
        ("ntime", ctypes.c_int),
        ("rec_width", ctypes.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        lp = ctypes.POINTER(ctypes.c_long)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.POINTER(OrcDesc), ctypes.c_int, fp, fp, fp]
        L.orc_destroy.argtypes = [vp]
        L.orc_fd_coef.argtypes = [ctypes.c_int, dp]
        L.orc_gap_num.restype = ctypes.c_int
        L.orc_gap_num.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.orc_ntime.restype = ctypes.c_int
        L.orc_ntime.argtypes = [ctypes.c_float, ctypes.c_float]
        L.orc_ricker.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int, fp]
        L.orc_node.restype = ctypes.c_long
        L.orc_node.argtypes = [ctypes.POINTER(OrcDesc), ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.orc_reset.argtypes = [vp, ctypes.c_int]
        L.orc_reset_images.argtypes = [vp]
        L.orc_vel.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.orc_stress.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.orc_source.argtypes = [vp, ctypes.c_int, ctypes.c_long, ctypes.c_float, ctypes.c_int, ctypes.c_int]
        L.orc_get_field.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.orc_set_field.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.orc_get_coef.argtypes = [vp, ctypes.c_int, dp]
        L.orc_model_shot.restype = ctypes.c_int
        L.orc_model_shot.argtypes = [vp, ctypes.c_long, fp, lp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.orc_ring_slice_size.restype = ctypes.c_size_t
        L.orc_ring_slice_size.argtypes = [vp]
        L.orc_rtm_shot.restype = ctypes.c_int
        L.orc_rtm_shot.argtypes = [vp, ctypes.c_long, fp, lp, ctypes.c_int, fp, ctypes.c_int,
                                   ctypes.c_int, dp, dp]
        L.orc_get_image.argtypes = [vp, ctypes.c_int, dp]
        L.orc_set_image.argtypes = [vp, ctypes.c_int, dp]
        L.orc_image_accumulate.argtypes = [vp]
        L.orc_finalize_box.argtypes = [vp] + [ctypes.c_int] * 6 + [dp, dp]
        L.orc_inject_residual.argtypes = [vp, ctypes.c_int, lp, ctypes.c_int, fp, fp, ctypes.c_int]
        L.orc_sample.argtypes = [vp, ctypes.c_int, lp, ctypes.c_int, ctypes.c_int, dp]
        L.orc_set_storage.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_set_ring_storage.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_quantize.argtypes = [ctypes.c_int, ctypes.c_int, fp, fp, ctypes.POINTER(ctypes.c_uint16),
                                   ctypes.c_long]
        L.orc_plant_fault.argtypes = [ctypes.c_int]
        L.orc_scalar_paths.argtypes = [ctypes.c_int]
        L.orc_set_ring_dynamic.argtypes = [vp, ctypes.c_int]
        L.orc_set_ring_bfp.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_bfp_round.argtypes = [fp, fp, ctypes.c_long]
        L.orc_ring_exponent.restype = ctypes.c_int
        L.orc_ring_exponent.argtypes = [ctypes.c_float]
        L.orc_set_snap.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


# Planted mistakes of the oracle (see Fault in fdtd_oracle.cpp): a pin is
# shown to bite by failing with one of them switched on.
FAULT_NONE = 0
FAULT_INJ_TIME, FAULT_INJ_SIGN, FAULT_INJ_COMP = 1, 2, 3
FAULT_SRC_SZZ, FAULT_SHEAR_LAMBDA, FAULT_FD_COEF = 4, 5, 6
FAULT_PML_VX_X, FAULT_PML_VY_Y, FAULT_PML_VZ_Z = 7, 8, 9
FAULT_PML_SXX_X, FAULT_PML_SXY_X, FAULT_PML_VY_SXY_X = 10, 11, 12


class scalar_paths:
    """Context manager: every oracle update takes the per-cell code (the
    vectorized row runs are off)."""

    def __enter__(self):
        lib().orc_scalar_paths(1)
        return self

    def __exit__(self, *exc):
        lib().orc_scalar_paths(0)
        return False


class planted_fault:
    """Context manager: oracles created and stepped inside it carry the
    planted mistake `code`."""

    def __init__(self, code):
        self.code = code

    def __enter__(self):
        lib().orc_plant_fault(self.code)
        return self

    def __exit__(self, *exc):
        lib().orc_plant_fault(FAULT_NONE)
        return False


def _f(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _l(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_long))


def fd_coef(half):
    out = np.zeros(half, np.float64)
    lib().orc_fd_coef(half, _d(out))
    return out


def gap_num(l, r, d):
    return lib().orc_gap_num(l, r, d)


def ntime(T, dt):
    return lib().orc_ntime(T, dt)


def ricker(fm, dt, nt):
    out = np.zeros(nt, np.float32)
    lib().orc_ricker(fm, dt, nt, _f(out))
    return out


STORE_FP32, STORE_FP16, STORE_BF16, STORE_BFP16 = 0, 1, 2, 3


def bfp_round(x):
    """float32 x as the block-scaled 16-bit store keeps it (runs of 16
    from x[0])."""
    x = np.ascontiguousarray(x, np.float32).ravel()
    out = np.zeros_like(x)
    lib().orc_bfp_round(_f(x), _f(out), x.size)
    return out


def ring_exponent(m):
    """Scale exponent of a store slice whose largest |value| is m."""
    return lib().orc_ring_exponent(float(m))


def quantize(fmt, x, lg=0):
    """(values, 16-bit codes) of float32 x stored in format fmt at scale 2^lg."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    bits = np.zeros(x.shape, np.uint16)
    lib().orc_quantize(fmt, lg, _f(x), _f(out), bits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), x.size)
    return out, bits


class Oracle:
    """Global-grid restatement of the reference propagators.

    vp, vs, rho: float32 arrays on the physical grid, shape (nz, ny, nx)
    (x fastest in memory, as the reference's Field).
    """

    def __init__(self, nx, ny, nz, dx, pml, dt, fm, ntime, rec_width, vp, vs, rho, prec=4, dy=None, dz=None):
        self.desc = OrcDesc(nx, ny, nz, dx, dy or dx, dz or dx, pml, dt, fm, ntime, rec_width)
        self.nx, self.ny, self.nz, self.pml, self.ntime = nx, ny, nz, pml, ntime
        self.NX, self.NY, self.NZ = nx + 2 * pml, ny + 2 * pml, nz + 2 * pml
        vp = np.ascontiguousarray(vp, np.float32)
        vs = np.ascontiguousarray(vs, np.float32)
        rho = np.ascontiguousarray(rho, np.float32)
        assert vp.shape == (nz, ny, nx)
        self.h = lib().orc_create(ctypes.byref(self.desc), prec, _f(vp), _f(vs), _f(rho))
        if not self.h:
            raise ValueError("oracle: invalid description")

    def set_storage(self, fmt, lg_v=0, lg_s=0):
        """Store wavefields in format fmt (FP32 mirror only) with scales
        2^lg_v (velocity class) and 2^lg_s (stress class)."""
        lib().orc_set_storage(self.h, fmt, lg_v, lg_s)

    def set_ring_storage(self, fmt, lg_v=0, lg_s=0):
        """VOWR boundary store in format fmt over FP32 wavefields (FP32
        mirror only): values quantized when saved, class scales 2^lg_v
        (velocities) and 2^lg_s (stresses)."""
        lib().orc_set_ring_storage(self.h, fmt, lg_v, lg_s)

    def set_ring_bfp(self, subs):
        """Block-scaled 16-bit VOWR store over FP32 wavefields, stored per
        subdomain part of each ring box as the engine stores it (subs: the
        decomposition, decompose())."""
        cuts = [sorted({int(s["phy_off"][a]) for s in subs}) for a in range(3)]
        flat = np.array(cuts[0] + cuts[1] + cuts[2], np.int32)
        lib().orc_set_ring_bfp(self.h, flat.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(cuts[0]),
                               len(cuts[1]), len(cuts[2]))

    def set_ring_dynamic(self, on=True):
        """16-bit VOWR store with one power-of-two scale per field and time
        slice (the slice's largest |value| in [2^14, 2^15)) instead of the
        class scales."""
        lib().orc_set_ring_dynamic(self.h, int(bool(on)))

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def ext_shape(self):
        return (self.NZ, self.NY, self.NX)

    def node(self, x, y, z):
        return lib().orc_node(ctypes.byref(self.desc), x, y, z)

    def reset(self, s=0):
        lib().orc_reset(self.h, s)

    def reset_images(self):
        lib().orc_reset_images(self.h)

    def vel(self, s, mode):
        lib().orc_vel(self.h, s, mode)

    def stress(self, s, mode):
        lib().orc_stress(self.h, s, mode)

    def source(self, s, node, val, sign=1, src_mode=0):
        lib().orc_source(self.h, s, node, val, sign, src_mode)

    def inject_residual(self, s, rec_nodes, obs, syn, it):
        """v += obs - syn at the receivers, sample it (FWI adjoint source)."""
        rec = np.ascontiguousarray(rec_nodes, np.int64)
        o = np.ascontiguousarray(obs, np.float32)
        y = np.ascontiguousarray(syn, np.float32)
        lib().orc_inject_residual(self.h, s, _l(rec), len(rec), _f(o), _f(y), it)

    def sample(self, s, rec_nodes, it, seis):
        """Receiver values of set s into seis[3][nrec][ntime] at sample it."""
        rec = np.ascontiguousarray(rec_nodes, np.int64)
        assert seis.dtype == np.float64 and seis.flags.c_contiguous
        lib().orc_sample(self.h, s, _l(rec), len(rec), it, _d(seis))

    def get(self, s, field):
        out = np.zeros(self.ext_shape, np.float64)
        lib().orc_get_field(self.h, s, field, _d(out))
        return out

    def set(self, s, field, arr):
        a = np.ascontiguousarray(arr, np.float64)
        assert a.shape == self.ext_shape
        lib().orc_set_field(self.h, s, field, _d(a))

    def coef(self, which):
        out = np.zeros(self.ext_shape, np.float64)
        lib().orc_get_coef(self.h, which, _d(out))
        return out

    def model_shot(self, shot_node, wavelet, rec_nodes, nsteps=None, src_mode=0):
        nsteps = nsteps or self.ntime
        rec = np.ascontiguousarray(rec_nodes, np.int64)
        w = np.ascontiguousarray(wavelet, np.float32)
        out = np.zeros((3, len(rec), nsteps), np.float64)
        lib().orc_model_shot(self.h, shot_node, _f(w), _l(rec), len(rec), nsteps, src_mode, _d(out))
        return out

    def set_snap(self, s, field):
        """rtm_shot snapshots hold `field`: of set 0 in the forward pass and
        of set s in the reverse loop (default set 0, VX)."""
        lib().orc_set_snap(self.h, s, field)

    def rtm_shot(self, shot_node, wavelet, rec_nodes, obs, src_mode=0, snap_every=0):
        rec = np.ascontiguousarray(rec_nodes, np.int64)
        w = np.ascontiguousarray(wavelet, np.float32)
        o = np.ascontiguousarray(obs, np.float32)
        assert o.shape == (3, len(rec), self.ntime)
        fs = bs = None
        if snap_every > 0:
            nsnap = (self.ntime - 1) // snap_every + 1
            fs = np.zeros((nsnap,) + self.ext_shape, np.float64)
            bs = np.zeros((nsnap,) + self.ext_shape, np.float64)
        lib().orc_rtm_shot(self.h, shot_node, _f(w), _l(rec), len(rec), _f(o), src_mode, snap_every,
                           _d(fs) if fs is not None else None, _d(bs) if bs is not None else None)
        return fs, bs

    def image(self, which):
        out = np.zeros((self.nz, self.ny, self.nx), np.float64)
        lib().orc_get_image(self.h, which, _d(out))
        return out

    def set_image(self, which, arr):
        a = np.ascontiguousarray(arr, np.float64)
        assert a.shape == (self.nz, self.ny, self.nx)
        lib().orc_set_image(self.h, which, _d(a))

    def image_accumulate(self):
        lib().orc_image_accumulate(self.h)

    def finalize_box(self, i0, j0, k0, ni, nj, nk):
        img = np.zeros((4, nk, nj, ni), np.float64)
        lap = np.zeros((4, nk, nj, ni), np.float64)
        lib().orc_finalize_box(self.h, i0, j0, k0, ni, nj, nk, _d(img), _d(lap))
        return img, lap


# ---------------------------------------------------------------------------
# Decomposition bookkeeping (pure Python restatement)
# ---------------------------------------------------------------------------
def decompose(n, d, pml, ndiv):
    """Subdomains of the reference's decomposition, rank order x fastest.

    n = (nx, ny, nz) global gridphy, d = (dx, dy, dz), ndiv = (x, y, z).
    Restates BASE/fdtd3d_base_3_geomodel_meat.hpp:579-704 (set_gridphy,
    set_gridext) and OP/jarvis_2_mpicu_halo_comm_operator_meat.hpp:26-101
    (rank -> (i, j, k)).  Frame l / r values are computed in float32 as the
    reference's Frame.
    """
    f32 = np.float32
    subs = []
    nsub = ndiv[0] * ndiv[1] * ndiv[2]
    for rank in range(nsub):
        ci = (rank % (ndiv[0] * ndiv[1]) % ndiv[0], rank % (ndiv[0] * ndiv[1]) // ndiv[0],
              rank // (ndiv[0] * ndiv[1]))
        s = {"rank": rank, "ci": ci, "phy_n": [], "phy_off": [], "ext_n": [], "mlo": [], "mhi": [],
             "lphy": [], "rphy": [], "lext": [], "rext": []}
        for a in range(3):
            ext = n[a] + 2 * pml
            i, nd = ci[a], ndiv[a]
            if nd == 1:
                pn = n[a]
            elif i == 0:
                pn = ext // nd - pml
            elif i < nd - 1:
                pn = ext // nd
            else:
                pn = ext // nd + ext % nd - pml
            off = 0 if i == 0 else i * (ext // nd) - pml
            lphy = f32(0) if i == 0 else f32(f32(i * (ext // nd)) * f32(d[a]) - f32(pml) * f32(d[a]))
            rphy = f32(lphy + f32(pn - 1) * f32(d[a]))
            mlo = pml if i == 0 else 0
            mhi = pml if i == nd - 1 else 0
            lext = f32(f32(-mlo) * f32(d[a]) + lphy)
            en = pn + mlo + mhi
            rext = f32(lext + f32(en - 1) * f32(d[a]))
            for k, v in (("phy_n", pn), ("phy_off", off), ("ext_n", en), ("mlo", mlo), ("mhi", mhi),
                         ("lphy", lphy), ("rphy", rphy), ("lext", lext), ("rext", rext)):
                s[k].append(v)
        subs.append(s)
    return subs


def kept_receivers(subs, recv, d, pml):
    """Rows of the per-rank seismograms in output order: (rank, input index,
    global extended node (i, j, k)).  A receiver belongs to each subdomain
    whose gridphy contains it, inclusive, in metres
    (UTIL/.jarvis_2_geometry/jarvis_device_geometry_meat.hpp:101-104); its
    node is get_gap_num from the subdomain's gridext origin
    (BASE/fdtd3d_base_4_seisrecord_meat.hpp:67-70)."""
    f32 = np.float32
    rows = []
    for s in subs:
        for ri, p in enumerate(np.asarray(recv, np.float32)):
            if all(p[a] >= s["lphy"][a] and p[a] <= s["rphy"][a] for a in range(3)):
                node = []
                for a in range(3):
                    loc = gap_num(float(s["lext"][a]), float(p[a]), float(d[a]))
                    node.append(loc + s["phy_off"][a] + pml - s["mlo"][a])
                rows.append((s["rank"], ri, tuple(node)))
    return rows


def ext_index(node, NX, NY):
    i, j, k = node
    return i + NX * (j + NY * k)
