"""The command-line driver hv_vlasov: the reference program's arguments,
input-file keys (include/init.hpp:276-337) and stats.txt lines
(src/hyper_vlasov.cpp:64-70), over the C-ABI."""
import json
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "hyper-vlasov_amd" / "hv_vlasov"

# the 21 keys in the order scripts/programs.py:42-66 writes them
INPUT = {"a": 0.6, "A": -356.0, "B": 303.0, "C": 32.0, "z": 20, "n": 20, "nx": 32, "V0": -50.0, "t_f": 10.0,
         "t_exc": 0.0, "eta_exc": 0.1, "steps": 20, "gamma": 7.0 / 6.0, "r_fwhm": 1.444, "k_fwhm": 0.3,
         "substeps": 5, "use_gpu": 1, "epsilon_p": -8.0, "epsilon_n": -8.0, "d_max_scale": 2.0,
         "num_test_part": 50}


def _write(path, values, extra=""):
    with open(path, "w") as f:
        for k, v in values.items():
            f.write(f"{k} {v}\n")
        f.write(extra)


def _check(path):
    return subprocess.run([str(DRIVER), "--check", str(path)], capture_output=True, text=True, timeout=60)


@pytest.fixture(scope="module")
def driver():
    if not DRIVER.exists():
        pytest.skip("hv_vlasov not built (make -C hyper-vlasov_amd)")
    return DRIVER


def test_check_derives_the_reference_run(driver, tmp_path):
    _write(tmp_path / "input.txt", INPUT)
    r = _check(tmp_path / "input.txt")
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    A = INPUT["z"] + INPUT["n"]
    assert d["n"] == 32
    assert d["d_max"] == pytest.approx(2.0 * 1.5 * A ** (1 / 3), rel=1e-15)  # include/init.hpp:329
    assert d["n_protons"] == 1000 and d["n_neutrons"] == 1000
    assert d["sigma_r"] == pytest.approx(1.444 / (2 * math.sqrt(2 * math.log(2))), rel=1e-15)
    assert d["dt"] == 0.5
    assert d["sor"] == [1.5, 1e-4, 512]  # include/physics.hpp:103-104
    assert d["boost"] == 0  # the reference's weighted split


def test_optional_keys(driver, tmp_path):
    _write(tmp_path / "input.txt", INPUT, "sor_omega 1.8\nsor_tol 1e-6\nboost 1\nseed 7\nvtk 0\nsome_unknown_token\n")
    d = json.loads(_check(tmp_path / "input.txt").stdout)
    assert d["sor"] == [1.8, 1e-6, 512] and d["boost"] == 1 and d["seed"] == 7 and d["vtk"] == 0


@pytest.mark.parametrize("edit", ["missing", "duplicate", "garbage"])
def test_invalid_input_exits_1(driver, tmp_path, edit):
    vals = dict(INPUT)
    extra = ""
    if edit == "missing":
        del vals["epsilon_n"]
    elif edit == "duplicate":
        extra = "steps 30\n"
    else:
        vals["nx"] = "thirty"
    _write(tmp_path / "input.txt", vals, extra)
    r = _check(tmp_path / "input.txt")
    assert r.returncode == 1
    assert "Invalid input file" in r.stderr


def test_bad_arguments(driver):
    r = subprocess.run([str(DRIVER)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "BAD ARGUMENTS" in r.stderr


@pytest.mark.gpu
def test_run_writes_reference_stats(driver, tmp_path):
    """A run's stats.txt, read as the reference's scripts read it (np.loadtxt,
    scripts/programs.py:92), equals the library's own step diagnostics."""
    import hvamd
    print("driver run", flush=True)
    _write(tmp_path / "input.txt", INPUT, "init 1\n")
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([str(DRIVER), str(tmp_path / "input.txt"), str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Simulation ended." in r.stdout
    data = np.loadtxt(out / "stats.txt", dtype=np.float64)
    d = json.loads(_check(tmp_path / "input.txt").stdout)
    world = hvamd.World((32,) * 3, (d["d_max"],) * 3)
    sim = hvamd.Simulation(world, d["n_protons"], d["n_neutrons"], hvamd.HV_PREC_F64)
    sim.upload_particles(*hvamd.sample_uniform_spheres(128, d["n_protons"], d["n_neutrons"], d["r_init"],
                                                       d["k_fermi"]))
    sk = hvamd.Skyrme(*d["skyrme"])
    prm = hvamd.Params(t_f=10.0, steps=20, substeps=5, t_exc=0.0, eta_exc=0.1,
                       boost_mode=hvamd.HV_BOOST_REF_WEIGHTED, z=20, n=20, sigma_r=d["sigma_r"],
                       part_per_nucleon=50, sor_omega=1.5, sor_tol=1e-4, sor_max_it=512)
    sim.run_init(sk, prm)
    s1, _ = sim.run_steps(sk, prm, 0, 10)
    snap = tmp_path / "snap.vtk"
    assert sim.lib.hv_write_vtk(sim.ctx, str(snap).encode()) == 0
    s2, _ = sim.run_steps(sk, prm, 10, 10)
    sim.close()
    stats = np.concatenate([s1, s2])
    # snapshots at the top of steps 0, 5, 10, 15 (src/hyper_vlasov.cpp:64-85)
    assert sorted(p.name for p in out.glob("*.vtk")) == [f"out-{i:04d}.vtk" for i in range(4)]
    assert (out / "out-0002.vtk").read_bytes() == snap.read_bytes()
    assert r.stdout.count("Processed step:") == 4
    rows = stats[::5]
    want = np.column_stack([0.5 * np.arange(0, 20, 5), np.sqrt(rows[:, 2]), np.sqrt(rows[:, 3]), rows[:, 0],
                            rows[:, 1]])
    assert data.shape == want.shape
    np.testing.assert_allclose(data, want, rtol=5e-6, atol=1e-12)  # %e keeps 7 significant digits


@pytest.mark.gpu
def test_run_from_the_reference_initial_state(driver, tmp_path):
    """Without `init 1` the driver starts from the reference's Woods-Saxon /
    Fermi-level nucleus (hv_init_woods_saxon with the input's V0, a,
    epsilon_p, epsilon_n, k_fwhm and seed 128), prints the reference's
    initialisation lines, and runs the loop from it."""
    print("driver run from the reference's initial state", flush=True)
    _write(tmp_path / "input.txt", INPUT, "vtk 0\n")
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([str(DRIVER), str(tmp_path / "input.txt"), str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    assert "MAX TEST PART" in r.stdout and "CHI SQUARED P" in r.stdout and "Simulation ended." in r.stdout
    data = np.loadtxt(out / "stats.txt", dtype=np.float64)
    assert data.shape[1] == 5 and np.isfinite(data).all()
    # sqrt<r^2> of a bound nucleus of A = 40: a few fm
    assert 2.0 < data[0, 1] < 5.0 and 2.0 < data[0, 2] < 5.0


@pytest.mark.gpu
def test_reference_initial_state_repeats_across_processes(driver, tmp_path):
    """Two runs of the driver from the reference's initial state (seed 128)
    in two processes print the same initialisation and the same stats.txt:
    the rounds' rand() sequence is private to the initialisation, so the GPU
    runtime's own threads cannot shift it."""
    print("driver: reference initial state in two processes", flush=True)
    runs = []
    for rep in range(2):
        d = tmp_path / f"run{rep}"
        d.mkdir()
        _write(d / "input.txt", INPUT, "vtk 0\n")
        (d / "out").mkdir()
        r = subprocess.run([str(DRIVER), str(d / "input.txt"), str(d / "out")], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
        lines = [ln for ln in r.stdout.splitlines() if not ln.startswith("Time taken")]
        runs.append((lines, (d / "out" / "stats.txt").read_bytes()))
    assert runs[0][0] == runs[1][0]
    assert runs[0][1] == runs[1][1]
