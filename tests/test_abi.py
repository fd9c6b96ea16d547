"""The C-ABI library: it loads without a GPU, exports every entry point the
public headers declare, and its host-only parts (the seeded sampler) behave.
No device calls."""
import ctypes
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: Path):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hv_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import hvamd
    lib = hvamd.load()
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        names += _declared(h)
    assert len(names) > 30
    nm = subprocess.run(["nm", "-D", "--defined-only", str(hvamd._lib.LIB_PATH)], capture_output=True,
                        text=True, check=True).stdout
    exported = set(re.findall(r"\bT (hv_[a-z0-9_]+)$", nm, flags=re.M))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert getattr(lib, n) is not None


def test_python_binding_covers_the_abi():
    from hvamd import _lib
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        names += _declared(h)
    assert sorted(set(names) - set(_lib.SIGNATURES)) == []


def test_build_info_names_gfx950():
    import hvamd
    assert b"gfx950" in hvamd.load().hv_build_info()


def test_sampler_reproducible_and_uniform():
    import hvamd
    a = hvamd.sample_uniform_spheres(128, 20000, 30000, 7.11, 1.333)
    b = hvamd.sample_uniform_spheres(128, 20000, 30000, 7.11, 1.333)
    c = hvamd.sample_uniform_spheres(129, 20000, 30000, 7.11, 1.333)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    assert not np.array_equal(a[0], c[0])
    r = np.sqrt(a[0] ** 2 + a[1] ** 2 + a[2] ** 2)
    k = np.sqrt(a[3] ** 2 + a[4] ** 2 + a[5] ** 2)
    assert r.max() < 7.11 and k.max() < 1.333
    # uniform in a ball: P(r < s R) = s^3 ; <r^2> = 3/5 R^2
    assert abs((r < 0.5 * 7.11).mean() - 0.125) < 0.01
    assert abs((r ** 2).mean() / (0.6 * 7.11 ** 2) - 1) < 0.01
    assert abs((k ** 2).mean() / (0.6 * 1.333 ** 2) - 1) < 0.01
    # isotropy: mean position and momentum near zero
    assert abs(a[2].mean()) < 0.05 and abs(a[5].mean()) < 0.01


def test_sampler_prefix_stable():
    """particle i depends only on (seed, i): a larger draw starts with the same particles."""
    import hvamd
    a = hvamd.sample_uniform_spheres(7, 100, 0, 3.0, 1.0)
    b = hvamd.sample_uniform_spheres(7, 250, 0, 3.0, 1.0)
    assert np.array_equal(a[0], b[0][:100])


def test_fermi_momentum():
    import hvamd
    assert abs(hvamd.load().hv_fermi_momentum(0.16) - (3 * np.pi ** 2 * 0.16 / 2) ** (1 / 3)) < 1e-14


def test_sampler_rejects_bad_args():
    import hvamd
    with pytest.raises(hvamd.HvError):
        hvamd.sample_uniform_spheres(1, -1, 3, 1.0, 1.0)


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_oracle_header_only_used_by_tests():
    """The product never links or loads the oracle."""
    for p in (ROOT / "hyper-vlasov_amd").rglob("*"):
        if p.suffix in (".py", ".hip", ".cpp", ".h") or p.name == "Makefile":
            assert "liboracle" not in p.read_text() and "hv_oracle" not in p.read_text(), p


def test_stats_noalloc_rejects_sizes_beyond_its_scratch():
    """hv_stats_noalloc checks its arguments before any device call: more
    particles than its fixed scratch holds, an empty species or a null pointer
    return HV_ERR_ARG (no GPU needed for these)."""
    import hvamd
    from hvamd import _lib
    lib = hvamd.load()
    w = hvamd.World((16, 16, 16), (4.0, 4.0, 4.0)).c()
    dummy = ctypes.c_void_p(16)  # never dereferenced: the checks fail first
    big = _lib.HV_NOALLOC_MAX_PARTICLES
    for np_, nn in ((big, 1), (big // 2, big // 2 + 1), (0, 5), (5, 0)):
        assert lib.hv_stats_noalloc(ctypes.byref(w), 0, dummy, dummy, dummy, np_, nn, dummy, None) == _lib.HV_ERR_ARG
    assert lib.hv_stats_noalloc(ctypes.byref(w), 0, None, dummy, dummy, 5, 5, dummy, None) == _lib.HV_ERR_ARG
    assert lib.hv_stats_noalloc(ctypes.byref(w), 0, dummy, dummy, dummy, 5, 5, None, None) == _lib.HV_ERR_ARG
    assert lib.hv_stats_noalloc(ctypes.byref(w), 7, dummy, dummy, dummy, 5, 5, dummy, None) == _lib.HV_ERR_ARG
