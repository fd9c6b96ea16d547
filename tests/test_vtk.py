"""VTK snapshots (include/init.hpp:160-254): the library's host writer
hv_vtk_write_fields against files the reference's own writer produced
(tests/golden/snapshot_5x4x3_*.vtk, from tests/golden/make_golden.py vtk),
byte for byte, and against the live reference when its harness is built.
CPU only: the writer needs no GPU."""
import ctypes
import sys
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent / "hyper-vlasov_amd"))
sys.path.insert(0, str(HERE))

import hvamd  # noqa: E402
from hvamd._lib import HvWorld  # noqa: E402
from oracle_lib import Reference  # noqa: E402

GOLD = HERE / "golden"


def _write(path, n, d, f, single):
    lib = hvamd.load()
    w = HvWorld((ctypes.c_int * 3)(*n), (ctypes.c_double * 3)(*d))
    arr = [np.ascontiguousarray(f[k], dtype=np.float64) for k in ("density", "potentials", "fx", "fy", "fz")]
    ptr = [a.ctypes.data_as(ctypes.c_void_p) for a in arr]
    return lib.hv_vtk_write_fields(str(path).encode(), ctypes.byref(w), *ptr, 1 if single else 0)


def _fixture():
    z = np.load(GOLD / "snapshot_5x4x3.npz")
    return tuple(int(v) for v in z["n"]), tuple(float(v) for v in z["d_max"]), {k: z[k] for k in z.files}


@pytest.mark.parametrize("tag,single", [("f64", 0), ("f32", 1)])
def test_vtk_writer_matches_reference_file(tmp_path, tag, single):
    n, d, f = _fixture()
    out = tmp_path / "out.vtk"
    assert _write(out, n, d, f, single) == 0
    assert out.read_bytes() == (GOLD / f"snapshot_5x4x3_{tag}.vtk").read_bytes()


def test_vtk_layout_closed_form(tmp_path):
    """Header text, section order, big-endian float32 and x-fastest order,
    read back independently of the writer."""
    n, d = (3, 2, 4), (1.0, 2.0, 3.0)
    G = 24
    q = np.arange(2 * G, dtype=np.float64)
    f = {"density": q + 0.5, "potentials": -q, "fx": q * 2, "fy": q * 3, "fz": q * 4}
    out = tmp_path / "s.vtk"
    assert _write(out, n, d, f, 0) == 0
    b = out.read_bytes()
    head = (b"# vtk DataFile Version 3.0\nVolumetric data\nBINARY\nDATASET STRUCTURED_POINTS\n"
            b"DIMENSIONS 3 2 4\nORIGIN -1.000000 -2.000000 -3.000000\nSPACING 0.666667 2.000000 1.500000\n"
            b"POINT_DATA 24\n")
    assert b.startswith(head)
    pos = len(head)

    def take(label, count):
        nonlocal pos
        assert b[pos:pos + len(label)] == label, (label, b[pos:pos + 40])
        pos += len(label)
        v = np.frombuffer(b[pos:pos + 4 * count], dtype=">f4")
        pos += 4 * count
        return v

    # node (i, j, k) sits at (i*ny + j)*nz + k in the fields, at (k*ny + j)*nx + i in the file
    i, j, k = np.meshgrid(np.arange(3), np.arange(2), np.arange(4), indexing="ij")
    src = ((i * 2 + j) * 4 + k).ravel()
    dst = ((k * 2 + j) * 3 + i).ravel()
    for sp, tag in ((0, b"p"), (1, b"n")):
        v = take(b"VECTORS forces_" + tag + b" float\n", 3 * G).reshape(G, 3)
        for a, key in enumerate(("fx", "fy", "fz")):
            assert np.array_equal(v[dst, a], f[key][sp * G + src].astype(np.float32))
    for name in (b"density", b"potentials"):
        fld = f[name.decode()]
        for part, tag in ((0, b"p"), (1, b"n"), (2, b"t")):
            v = take(b"SCALARS " + name + b"_" + tag + b" float 1\nLOOKUP_TABLE default\n", G)
            want = (fld[src] + fld[G + src]) if part == 2 else fld[part * G + src]
            assert np.array_equal(v[dst], want.astype(np.float32))
    assert pos == len(b)


@pytest.mark.skipif(not Reference.available, reason="reference harness not built")
@pytest.mark.parametrize("single", [0, 1])
def test_vtk_writer_matches_live_reference(tmp_path, single):
    rng = np.random.default_rng(5)
    n, d = (7, 6, 5), (3.3, 2.9, 4.1)
    G2 = 2 * int(np.prod(n))
    f = {k: rng.normal(0, 10, G2) for k in ("fx", "fy", "fz", "potentials")}
    f["density"] = rng.uniform(0, 0.3, G2)
    Reference().write_vtk(tmp_path / "ref.vtk", n, d, f["fx"], f["fy"], f["fz"], f["density"], f["potentials"],
                          single=bool(single))
    assert _write(tmp_path / "ours.vtk", n, d, f, single) == 0
    assert (tmp_path / "ours.vtk").read_bytes() == (tmp_path / "ref.vtk").read_bytes()


def test_vtk_writer_reports_unwritable_path(tmp_path):
    n, d, f = _fixture()
    assert _write(tmp_path / "missing_dir" / "x.vtk", n, d, f, 0) == hvamd._lib.HV_ERR_IO
