"""Copy the reference's shipped project files into tests/golden/host/ as
fixtures for the host-format tests (run in a container that has the
reference checkout: python tests/golden/make_host_fixtures.py).

Copied as data: app/.main/fdtd3d_blend_elastic_model_mpi/fd3d_proj_small/
{cube100,salt}/argfile.json, cube100/shot.txt, cube100/rece.txt, and an
excerpt of cube100/out/vx_seis_0_0.txt (the first 3 lines, first 48 values
each, byte for byte) together with the row and column counts of the three
shipped seismogram files.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/app/.main/fdtd3d_blend_elastic_model_mpi/fd3d_proj_small"
OUT = os.path.join(HERE, "host")


def copy(src, dst):
    with open(src, "rb") as f:
        data = f.read()
    os.makedirs(os.path.dirname(dst), exist_ok=True)
    with open(dst, "wb") as f:
        f.write(data)


def main():
    for proj in ("cube100", "salt"):
        copy(os.path.join(SRC, proj, "argfile.json"), os.path.join(OUT, proj, "argfile.json"))
    for name in ("shot.txt", "rece.txt"):
        copy(os.path.join(SRC, "cube100", name), os.path.join(OUT, "cube100", name))
    meta = {}
    for comp in ("vx", "vy", "vz"):
        with open(os.path.join(SRC, "cube100", "out", f"{comp}_seis_0_0.txt"), "rb") as f:
            lines = f.read().split(b"\n")
        body = [ln for ln in lines if ln]
        meta[comp] = {"rows": len(body), "cols": [len(ln.split(b"\t")) - 1 for ln in body][0],
                      "all_rows_same_cols": len({len(ln.split(b"\t")) for ln in body}) == 1,
                      "ends_with_tab_newline": all(ln.endswith(b"\t") for ln in body)}
        if comp == "vx":
            ex = b"".join(b"".join(v + b"\t" for v in ln.split(b"\t")[:48]) + b"\n" for ln in body[:3])
            with open(os.path.join(OUT, "vx_seis_excerpt.txt"), "wb") as f:
                f.write(ex)
    with open(os.path.join(OUT, "seis_fixture_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()
