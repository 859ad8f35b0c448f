"""Host C++ layer (domain-decomposition-release_amd/host/): argfile reader,
shot/receiver reader, GMS reader and writer, Field raw/fld/txt I/O.

Pinned against the reference's shipped project files (tests/golden/host/,
copied by tests/golden/make_host_fixtures.py) and against an independent
Python restatement of the binary layouts (GMS header offsets, .fld header).
"""
import json
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import ddx
from cases import gms_bytes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "domain-decomposition-release_amd")
GOLD = os.path.join(ROOT, "tests", "golden", "host")
TOOL = os.path.join(PKG, "lib", "ddx_hosttool")


@pytest.fixture(scope="module")
def tool():
    srcs = [os.path.join(PKG, "host", f) for f in ("json_lite.cpp", "jarvis_host.cpp", "hosttool.cpp",
                                                    "jarvis_host.hpp", "json_lite.hpp")]
    if not os.path.exists(TOOL) or os.path.getmtime(TOOL) < max(os.path.getmtime(s) for s in srcs):
        sys.path.insert(0, PKG)
        import build as ddx_build
        ddx_build.build_host_tool()
    return TOOL


def run(tool, *args):
    r = subprocess.run([tool] + [str(a) for a in args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def test_argfile_shipped_cube100(tool):
    proj = os.path.join(GOLD, "cube100") + "/"
    a = json.loads(run(tool, "argfile", proj))
    assert a["proj_path"] == proj
    assert a["model_path"] == proj + "cube100.gms"
    assert a["shot_path"] == proj + "shot.txt"
    assert a["recv_path"] == proj + "rece.txt"
    assert a["output_path"] == proj + "out/"
    assert (a["x_divide_num"], a["y_divide_num"], a["z_divide_num"]) == (2, 2, 1)
    assert a["slots_per_node"] == 4 and a["pml_num"] == 10 and a["t_snap"] == -1 and a["is_output"] == 0
    assert np.float32(a["fm"]) == np.float32(100) and np.float32(a["dt"]) == np.float32(0.0001)
    assert np.float32(a["T"]) == np.float32(0.16)
    # engine keys absent from the shipped file take their defaults
    assert a["rec_width"] == 1 and a["image_fp64"] == 0 and a["source_mode"] == 0


def test_ntime_matches_shipped_seismograms(tool):
    """The shipped cube100 seismograms hold ceil(T/dt) = 1600 samples per
    receiver for T = 0.16, dt = 0.0001 taken from the shipped argfile."""
    a = json.loads(run(tool, "argfile", os.path.join(GOLD, "cube100") + "/"))
    meta = json.load(open(os.path.join(GOLD, "seis_fixture_meta.json")))
    nt = ddx.ntime(a["T"], a["dt"])
    for comp in ("vx", "vy", "vz"):
        assert meta[comp]["cols"] == nt == 1600
        assert meta[comp]["all_rows_same_cols"] and meta[comp]["ends_with_tab_newline"]


def test_argfile_shipped_salt(tool):
    proj = os.path.join(GOLD, "salt") + "/"
    a = json.loads(run(tool, "argfile", proj))
    assert a["model_path"] == proj + "cube1001.gms"
    assert np.float32(a["fm"]) == np.float32(20) and np.float32(a["dt"]) == np.float32(0.002)
    assert np.float32(a["T"]) == np.float32(4) and a["pml_num"] == 10
    assert ddx.ntime(a["T"], a["dt"]) == 2000


def test_argfile_conversions_and_engine_keys(tool, tmp_path):
    """jsoncpp-style conversions: real -> int truncates, int -> float,
    missing keys read as 0 / "" / false, both comment styles accepted."""
    (tmp_path / "argfile.json").write_text("""{
      /* block comment */
      "model_path": "m.gms", // line comment
      "pml_num": 12.9,
      "fm": 15,
      "dt": 1e-3,
      "T": 2,
      "is_output": 1,
      "x_divide_num": 2, "y_divide_num": 1,
      "rec_width": 4, "image_fp64": 1, "source_mode": 1, "cfl_policy": 1, "storage": "bf16", "ring_storage": "1",
      "n_xyz": [256, 255, 200], "d_xyz": [10, 10.5, 12.5],
      "vp_path": "vp.raw"
    }""")
    a = json.loads(run(tool, "argfile", str(tmp_path) + "/"))
    assert a["pml_num"] == 12 and np.float32(a["fm"]) == 15 and a["is_output"] == 1
    assert a["z_divide_num"] == 0 and a["slots_per_node"] == 0 and a["shot_path"] == str(tmp_path) + "/"
    assert (a["rec_width"], a["image_fp64"], a["source_mode"], a["cfl_policy"]) == (4, 1, 1, 1)
    assert a["storage"] == ddx.STORE_BF16 and a["ring_storage"] == ddx.STORE_FP16
    assert a["n_xyz"] == [256, 255, 200] and a["d_xyz"] == [10, 10.5, 12.5]
    assert ddx.ntime(a["T"], a["dt"]) == 2000


def test_argfile_missing_aborts(tool, tmp_path):
    r = subprocess.run([tool, "argfile", str(tmp_path) + "/nothing/"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "arg_file open error" in r.stdout


def test_shot_and_receiver_lines(tool):
    shots = json.loads(run(tool, "points", os.path.join(GOLD, "cube100", "shot.txt")))
    assert shots == [[25, 25, 25, 0], [25, 25, 50, 1]]  # the file has no trailing newline
    recv = json.loads(run(tool, "points", os.path.join(GOLD, "cube100", "rece.txt")))
    assert recv == [[75, 75, 75, 0]]


def test_gms_reader_layout(tool, tmp_path):
    nx, ny, nz = 5, 4, 3
    rng = np.random.default_rng(3)
    fields = [rng.standard_normal((nz, ny, nx)).astype(np.float32) for _ in range(3)]
    p = tmp_path / "m.gms"
    p.write_bytes(gms_bytes((nx, ny, nz), (10.0, 12.5, 20.0), (100.0, -50.0, 0.0), ["vp", "vs", "rho"], fields))
    info = json.loads(run(tool, "gms2raw", p, str(tmp_path) + "/p"))
    assert info["para_num"] == 3 and info["names"] == ["vp", "vs", "rho"]
    fr = info["frame"]
    assert fr["n"] == [nx, ny, nz] and fr["d"] == [10.0, 12.5, 20.0] and fr["l"] == [100.0, -50.0, 0.0]
    assert fr["r"] == [140.0, -12.5, 40.0] and fr["type"] == 30  # grid
    assert fr["size_info"] == "_5_4_3_100.000000_-50.000000_0.000000"
    for q in range(3):
        got = np.fromfile(info["files"][q], np.float32).reshape(nz, ny, nx)
        assert info["files"][q].endswith(f"p{q}_5_4_3_100.000000_-50.000000_0.000000.raw")
        assert np.array_equal(got, fields[q])


def test_gms_writer_matches_layout(tool, tmp_path):
    nx, ny, nz = 6, 3, 4
    rng = np.random.default_rng(4)
    fields = [rng.standard_normal((nz, ny, nx)).astype(np.float32) for _ in range(2)]
    raws = []
    for q, f in enumerate(fields):
        raws.append(tmp_path / f"f{q}.raw")
        f.tofile(raws[-1])
    out = tmp_path / "w.gms"
    run(tool, "writegms", out, nx, ny, nz, 10, 10, 10, 0, 0, 0, *raws)
    ref = gms_bytes((nx, ny, nz), (10.0, 10.0, 10.0), (0.0, 0.0, 0.0), ["para0", "para1"], fields)
    got = out.read_bytes()
    assert len(got) == len(ref)
    assert got[74:112] == ref[74:112]  # para_num, l, n, d
    assert got[240:] == ref[240:]


def test_seismogram_txt_matches_shipped_format(tool, tmp_path):
    """Parse the excerpt of the reference's vx_seis_0_0.txt, write it back
    through Field::save(ascii_txt) and compare byte for byte."""
    ex = open(os.path.join(GOLD, "vx_seis_excerpt.txt"), "rb").read()
    rows = [ln for ln in ex.split(b"\n") if ln]
    vals = np.array([[float(v) for v in ln.split(b"\t") if v] for ln in rows], np.float32)
    raw = tmp_path / "s.raw"
    vals.tofile(raw)  # [n_recv][ntime]: time fastest
    run(tool, "txt", raw, vals.shape[1], vals.shape[0], tmp_path / "s")
    assert (tmp_path / "s.txt").read_bytes() == ex


def test_fld_header_layout(tool, tmp_path):
    nx, ny, nz = 7, 5, 3
    a = np.arange(nx * ny * nz, dtype=np.float32).reshape(nz, ny, nx)
    raw = tmp_path / "a.raw"
    a.tofile(raw)
    r = json.loads(run(tool, "fld", raw, nx, ny, nz, 2, 3, 4, 1, -1, 0.5, tmp_path / "a"))
    assert r["same"]
    b = (tmp_path / "a.fld").read_bytes()
    assert len(b) == 72 + a.nbytes
    n_elem, n_rows, n_slice = struct.unpack_from("<3Q", b, 0)
    n_cols, n_slices = struct.unpack_from("<2I", b, 24)
    d = struct.unpack_from("<3f", b, 32)
    l = struct.unpack_from("<3f", b, 44)
    rr = struct.unpack_from("<3f", b, 56)
    (typ,) = struct.unpack_from("<i", b, 68)
    assert (n_elem, n_rows, n_slice, n_cols, n_slices) == (nx * ny * nz, nx, nx * ny, ny, nz)
    assert d == (2.0, 3.0, 4.0) and l == (1.0, -1.0, 0.5) and rr == (13.0, 11.0, 8.5) and typ == 30
    assert np.array_equal(np.frombuffer(b[72:], np.float32).reshape(nz, ny, nx), a)


def test_argfile_rejects_unknown_ring_storage(tool, tmp_path):
    (tmp_path / "argfile.json").write_text('{"ring_storage": "fp8"}')
    r = subprocess.run([tool, "argfile", str(tmp_path) + "/"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "ring_storage" in (r.stdout + r.stderr)


def test_argfile_block_scaled_store(tool, tmp_path):
    """ring_storage "bfp16" (or 3) selects the block-scaled store; as the
    wavefield storage it is refused (a store format only)."""
    for val in ('"bfp16"', '"3"'):
        (tmp_path / "argfile.json").write_text('{"ring_storage": %s}' % val)
        a = json.loads(run(tool, "argfile", str(tmp_path) + "/"))
        assert a["ring_storage"] == ddx.STORE_BFP16 and a["storage"] == ddx.STORE_FP32
    (tmp_path / "argfile.json").write_text('{"storage": "bfp16"}')
    r = subprocess.run([tool, "argfile", str(tmp_path) + "/"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "storage" in (r.stdout + r.stderr)
