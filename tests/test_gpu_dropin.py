"""The drop-in gpu_simulate (hyper-vlasov_amd/dropin/hv_gpu_simulate.hpp)
compiled against the reference's own headers (oracle/_ref/libhvdropin.so,
built by `make -C oracle dropin` where the reference checkout exists): the
function a maintainer adds at src/hyper_vlasov.cpp:120-123, run on the
reference's types.  Its stats.txt and final particles equal the CPU oracle's
loop bit for bit (FP64; the same stats text), stay within the oracle's own
tolerance of the reference's loop (oracle/_ref/libhvref.so), and its
snapshots are the reference writer's files."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

from hvamd import configs
from oracle_lib import Reference

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
DROPIN = ROOT / "oracle" / "_ref" / "libhvdropin.so"
KEYS = ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz")


def _dropin():
    if not DROPIN.exists():
        pytest.skip("oracle/_ref/libhvdropin.so not built (needs the reference checkout at build time)")
    L = ctypes.CDLL(str(DROPIN))
    P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    L.dropin_gpu_simulate.argtypes = [P] * 9 + [I, I, P, P] + [D] * 5 + [I, I, D, D, I, I, D, I, ctypes.c_char_p, I]
    L.dropin_gpu_simulate.restype = None
    return L


def _case():
    """The A=40 nucleus with 4,000 test particles on a 32^3 grid, 24 steps of
    0.5 fm/c, stats and snapshots every 4 steps, the reference's boost at
    t_exc = 2 fm/c (step 4) and its SOR settings."""
    w = configs.c0()
    x, y, z, kx, ky, kz = w.initial_state()
    st = dict(x=x, y=y, z=z, kx=kx, ky=ky, kz=kz)
    st.update({k: np.zeros(w.total) for k in ("fx", "fy", "fz")})
    return w, st, dict(steps=24, substeps=4, t_f=12.0, t_exc=2.0, eta=0.1014)


def _run(L, w, st, c, outdir, single):
    a = {k: np.ascontiguousarray(v, dtype=np.float64).copy() for k, v in st.items()}
    n = (ctypes.c_int * 3)(*w.world().n)
    d = (ctypes.c_double * 3)(*w.world().d_max)
    sk = w.sky()
    prm = w.params()
    L.dropin_gpu_simulate(*[a[k].ctypes.data for k in KEYS], w.n_protons, w.n_neutrons, n, d, sk.A, sk.B, sk.C,
                          sk.gamma, c["t_f"], c["steps"], c["substeps"], c["t_exc"], c["eta"], w.z, w.n,
                          prm.sigma_r, prm.part_per_nucleon, str(outdir).encode(), single)
    return a


def test_dropin_fp64_equals_oracle_and_tracks_reference(oracle, tmp_path):
    L = _dropin()
    w, st, c = _case()
    print("dropin gpu_simulate<double> on", w.name, flush=True)
    got = _run(L, w, st, c, tmp_path, 0)
    prm = w.params()
    sk = w.sky()
    n, d = w.world().n, w.world().d_max
    dt = c["t_f"] / c["steps"]
    st_o, stats_o, its_o, _, warn = oracle.simulate(st, w.n_protons, w.n_neutrons, n, d, sk.A, sk.B, sk.C, sk.gamma,
                                                    dt, c["steps"], c["t_exc"], c["eta"], 0, w.z, w.n, prm.sigma_r,
                                                    prm.part_per_nucleon, 1.5, 1e-4, 512)
    for k in KEYS:
        assert np.array_equal(got[k], st_o[k]), k
    # stats.txt: the reference's line format of the oracle's values, as text
    want = "".join("%e %e %e %e %e\n" % (s * dt, np.sqrt(stats_o[s, 2]), np.sqrt(stats_o[s, 3]), stats_o[s, 0],
                                        stats_o[s, 1]) for s in range(0, c["steps"], c["substeps"]))
    assert (tmp_path / "stats.txt").read_text() == want
    # the reference's own loop (its kernels, compiled from its headers)
    if Reference.available:
        ref = Reference()
        st_r, stats_r, _ = ref.simulate(st, w.n_protons, w.n_neutrons, n, d, sk.A, sk.B, sk.C, sk.gamma, c["t_f"],
                                        c["steps"], c["t_exc"], c["eta"], w.z, w.n, prm.sigma_r, prm.part_per_nucleon)
        for k in ("x", "y", "z", "kx", "ky", "kz"):
            scale = np.abs(st_r[k]).max()
            assert np.abs(got[k] - st_r[k]).max() <= 1e-9 * scale, k
        assert np.abs(stats_o - stats_r).max() <= 1e-9 * np.abs(stats_r).max()
    # one snapshot per stats line, each of the size the reference's writer gives
    files = sorted(p.name for p in tmp_path.glob("out-*.vtk"))
    assert files == ["out-%04d.vtk" % i for i in range(c["steps"] // c["substeps"])]
    if Reference.available:
        G = int(np.prod(n))
        ref = Reference()
        probe = tmp_path / "ref_layout.vtk"
        z2 = np.zeros(2 * G)
        ref.write_vtk(str(probe), n, d, z2, z2, z2, z2, z2)
        head = probe.read_bytes()
        body = (tmp_path / files[0]).read_bytes()
        assert len(body) == len(head)
        assert body[:200].split(b"BINARY")[0] == head[:200].split(b"BINARY")[0]


def test_dropin_float_runs_the_mixed_form(tmp_path):
    """gpu_simulate<float> (the reference's --float build) runs the mixed form
    and tracks the double run's dipole over the same 24 steps."""
    L = _dropin()
    w, st, c = _case()
    print("dropin gpu_simulate<float> vs <double>", flush=True)
    (tmp_path / "d").mkdir()
    (tmp_path / "f").mkdir()
    _run(L, w, st, c, tmp_path / "d", 0)
    _run(L, w, st, c, tmp_path / "f", 1)
    a = np.loadtxt(tmp_path / "d" / "stats.txt")
    b = np.loadtxt(tmp_path / "f" / "stats.txt")
    assert a.shape == b.shape == (6, 5)
    da = (a[:, 3] - a[:, 4])
    db = (b[:, 3] - b[:, 4])
    assert np.abs(da - db).max() <= 1e-4 * np.abs(da).max() + 1e-6
    assert len(list((tmp_path / "f").glob("out-*.vtk"))) == 6
