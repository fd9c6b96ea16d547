"""Every benchmark workload as a GPU test, through the C-ABI.

- c0 (A=40, 4,000 tp, 32^3): the whole 1,000-step FP64 run bit for bit
  against the CPU oracle; the mixed form against FP64.
- c1 (A=208, 104,000 tp, 64^3): 20 FP64 steps bit for bit against the oracle;
  mixed vs FP64 D(t) within 1e-4 of the amplitude over the first 500 steps and
  the dipole-response centroid within 0.1 MeV over the first 1,000 steps.
- c2 (A=208, 1.04M tp, 128^3): 2 FP64 steps (full SOR) bit for bit against
  the oracle; mixed vs FP64 D(t) within 1e-4 over 100 steps.
- c3 (ensemble): real instances of the 256-nucleus ensemble at their own
  grids, particle counts, Skyrme sets and boosts, batch vs solo runs bit for
  bit in both forms, and one FP64 instance against the oracle.

The D(t) tolerances follow the chaos of the test-particle dynamics: an FP64 run
whose initial positions move by one FP32 rounding step leaves 1e-4 of the
D(t) amplitude after a few hundred steps (c0: ~500, c1: ~800), so the mixed
form is held to 1e-4 over horizons inside that window."""
import numpy as np
import pytest

import hvamd
from hvamd import configs, spectrum
from hvamd.batch import Batch

pytestmark = pytest.mark.gpu

PARTICLE_KEYS = ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz")


def _eq(a, b, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    if not np.array_equal(a, b):
        bad = np.flatnonzero(a.ravel() != b.ravel())
        i = bad[0]
        raise AssertionError(f"{what}: {bad.size} of {a.size} differ; first at {i}: {a.ravel()[i]!r} vs "
                             f"{b.ravel()[i]!r}")


def _oracle_run(oracle, w, state, steps):
    p = w.params(steps)
    s = w.sky()
    n, d = w.world().n, w.world().d_max
    return oracle.simulate(state, w.n_protons, w.n_neutrons, n, d, s.A, s.B, s.C, s.gamma, p.dt, steps, p.t_exc,
                           p.eta_exc, p.boost_mode, w.z, w.n, p.sigma_r, p.part_per_nucleon, p.sor_omega,
                           p.sor_tol, p.sor_max_it)


def _gpu_run(w, steps, prec=hvamd.HV_PREC_F64, chunk=None):
    x, y, z, kx, ky, kz = w.initial_state()
    sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, prec)
    try:
        sim.upload_particles(x, y, z, kx, ky, kz)
        prm = w.params(steps)
        it0 = sim.run_init(w.sky(), prm)
        stats, its = [], []
        chunk = chunk or steps
        for first in range(0, steps, chunk):
            s, it = sim.run_steps(w.sky(), prm, first, min(chunk, steps - first))
            stats.append(s)
            its.append(it)
        return it0, np.concatenate(stats), np.concatenate(its), sim.download_particles()
    finally:
        sim.close()


def _fp64_vs_oracle(oracle, w, steps, chunk=None):
    x, y, z, kx, ky, kz = w.initial_state()
    it0, stats, its, got = _gpu_run(w, steps, chunk=chunk)
    st_o, stats_o, its_o, _, _ = _oracle_run(oracle, w, dict(x=x, y=y, z=z, kx=kx, ky=ky, kz=kz), steps)
    print(w.name, steps, "steps; SOR iterations: pre-loop", it0, "last", list(its[-5:]), flush=True)
    assert it0 == its_o[0]
    _eq(its, its_o[1:], "SOR iterations")
    _eq(stats, stats_o, "stats")
    for k in PARTICLE_KEYS:
        _eq(got[k], st_o[k], k)


def test_c0_full_run_fp64_bitwise(oracle):
    """All 1,000 steps of the A=40, 32^3 workload, FP64, bit for bit (in
    chunks of 250 steps, as a driver writing stats between calls would)."""
    print("c0 1000 steps fp64 vs oracle", flush=True)
    _fp64_vs_oracle(oracle, configs.c0(), 1000, chunk=250)


def test_c1_fp64_bitwise(oracle):
    """20 steps of the A=208, 64^3 workload, FP64, bit for bit."""
    print("c1 20 steps fp64 vs oracle", flush=True)
    _fp64_vs_oracle(oracle, configs.c1(), 20)


def test_c2_fp64_bitwise(oracle):
    """2 steps of the 1.04M-particle, 128^3 workload with the full SOR (the
    reference stop rule; a few hundred iterations per solve), bit for bit."""
    print("c2 2 steps fp64 vs oracle", flush=True)
    _fp64_vs_oracle(oracle, configs.c2(), 2)


def _rel_d(w, stats_a, stats_b, horizon):
    da = hvamd.dipole(stats_a[:horizon], w.z, w.n)
    db = hvamd.dipole(stats_b[:horizon], w.z, w.n)
    return float(np.abs(da - db).max() / np.abs(da).max())


@pytest.fixture(scope="module")
def c1_runs():
    w = configs.c1()
    print("c1 1000 steps fp64 and mixed", flush=True)
    f64 = _gpu_run(w, 1000, hvamd.HV_PREC_F64)
    mix = _gpu_run(w, 1000, hvamd.HV_PREC_MIXED)
    return w, f64, mix


def test_c1_mixed_dipole_within_1e4_over_500_steps(c1_runs):
    w, f64, mix = c1_runs
    rel = _rel_d(w, f64[1], mix[1], 500)
    print("c1 mixed vs fp64: max |dD| / max |D| over 500 steps", rel)
    assert rel < 1e-4


def test_c1_mixed_centroid_within_0p1_mev_over_1000_steps(c1_runs):
    w, f64, mix = c1_runs
    eta = w.dp / hvamd.HBARC
    c64 = spectrum.dipole_response(f64[1], w.dt, w.z, w.n, eta)[4]
    cmx = spectrum.dipole_response(mix[1], w.dt, w.z, w.n, eta)[4]
    print("c1 centroid over 1000 steps: fp64", c64, "mixed", cmx, "diff", cmx - c64)
    assert 5.0 < c64 < 40.0
    assert abs(cmx - c64) < 0.1


def test_c0_mixed_dipole_within_1e4_over_250_steps():
    w = configs.c0()
    print("c0 250 steps fp64 and mixed", flush=True)
    f64 = _gpu_run(w, 250, hvamd.HV_PREC_F64)
    mix = _gpu_run(w, 250, hvamd.HV_PREC_MIXED)
    rel = _rel_d(w, f64[1], mix[1], 250)
    print("c0 mixed vs fp64: max |dD| / max |D| over 250 steps", rel)
    assert rel < 1e-4


def test_c2_mixed_dipole_within_1e4_over_100_steps():
    """The headline form at its own size: mixed vs FP64 over the first 100
    steps (25 fm/c)."""
    w = configs.c2()
    print("c2 100 steps fp64 and mixed", flush=True)
    f64 = _gpu_run(w, 100, hvamd.HV_PREC_F64)
    mix = _gpu_run(w, 100, hvamd.HV_PREC_MIXED)
    rel = _rel_d(w, f64[1], mix[1], 100)
    print("c2 mixed vs fp64: max |dD| / max |D| over 100 steps", rel, "SOR iterations fp64",
          float(f64[2].mean()), "mixed", float(mix[2].mean()))
    assert rel < 1e-4


# instances of the 256-nucleus ensemble (configs.c3) at their own sizes:
# index -> (grid, test particles): 7 -> (32, 26811), 12 -> (64, 5712),
# 9 -> (96, 10144), 3 -> (64, 15040), 0 -> (96, 126672)
C3_PICK = (7, 12, 9, 3, 0)


def _c3_instances(steps):
    ens = configs.c3(max(C3_PICK) + 1)
    ws = [ens[i] for i in C3_PICK]
    return ws, [w.params(steps) for w in ws]


def _solo(w, prm, prec, steps, resident):
    sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, prec)
    try:
        sim.set_option(hvamd.HV_OPT_SOR_RESIDENT_F32, 1 if resident else 0)
        sim.upload_particles(*w.initial_state())
        it0 = sim.run_init(w.sky(), prm)
        stats, its = sim.run_steps(w.sky(), prm, 0, steps)
        return it0, stats, its, sim.download_particles()
    finally:
        sim.close()


@pytest.mark.parametrize("form", ["f64", "mixed_resident", "mixed_twopass_sorted"])
def test_c3_real_instances_batch_equals_solo(form):
    """Five real ensemble instances (32^3, 64^3 and 96^3 grids, 5.7k-127k test
    particles, their own Skyrme sets and boosts) in one batch with the
    ensemble's default options, or (mixed_twopass_sorted) with the options
    its autotuner picks on the full ensemble (the two-pass smearing with the
    widest tiles, the batched sort every 4 steps): each equals its solo run
    bit for bit."""
    steps = 12 if form == "f64" else 24
    prec = hvamd.HV_PREC_F64 if form == "f64" else hvamd.HV_PREC_MIXED
    ws, prms = _c3_instances(steps)
    print(f"c3 real instances form={form}: " + ", ".join(f"{w.name} {w.total} tp" for w in ws), flush=True)
    b = Batch([w.world() for w in ws], [w.n_protons for w in ws], [w.n_neutrons for w in ws], prec)
    try:
        if form == "mixed_twopass_sorted":
            b.set_option(hvamd.HV_OPT_SMEAR_FUSED, 0)
            b.set_option(hvamd.HV_OPT_SMEAR_TILE, 64)
            b.set_option(hvamd.HV_OPT_SORT_EVERY, 4)
        for i, w in enumerate(ws):
            b.instance(i).upload_particles(*w.initial_state())
        sks = [w.sky() for w in ws]
        it0 = b.run_init(sks, prms)
        stats, its = b.run_steps(sks, prms, 0, steps)
        parts = [b.instance(i).download_particles() for i in range(len(ws))]
        resident = b.get_option(hvamd.HV_OPT_SOR_RESIDENT_F32) == 1 if prec == hvamd.HV_PREC_MIXED else False
    finally:
        b.close()
    if prec == hvamd.HV_PREC_MIXED:
        assert resident  # the ensemble's LDS-resident FP32 launch ran
    for i, w in enumerate(ws):
        s_it0, s_stats, s_its, s_part = _solo(w, prms[i], prec, steps, resident)
        print(i, w.name, "batch iters", list(its[i][:6]), "solo", list(s_its[:6]), flush=True)
        assert it0[i] == s_it0
        _eq(its[i], s_its, f"{w.name} iterations")
        _eq(stats[i], s_stats, f"{w.name} stats")
        for k in PARTICLE_KEYS:
            _eq(parts[i][k], s_part[k], f"{w.name} {k}")


def test_c3_instance_fp64_bitwise(oracle):
    """One real ensemble instance (A=27, 26,811 test particles, 32^3 over an
    18.6 fm box, its own B, C and boost), 30 FP64 steps, bit for bit against
    the oracle."""
    w = configs.c3(8)[7]
    print(f"c3 instance {w.name} fp64 vs oracle", flush=True)
    _fp64_vs_oracle(oracle, w, 30)


def test_c2_extrapolated_start_keeps_the_stop_rule_and_cuts_iterations():
    """The mixed form's extrapolated start (HV_OPT_SOR_EXTRAPOLATE 2, the
    default) against the last-potential start (0) on 60 steps of the 128^3
    workload: from step 10 on every solve ends by the stop rule (an FP64
    iteration with max |dphi| below the tolerance) well before the iteration
    cap, with far fewer iterations per step.  (The first solves after the
    pre-loop solve, which hits the cap from a zero interior at 128^3, may hit
    it too, in both forms as in the FP64 one; the history counts converged
    solves only.)"""
    w = configs.c2()
    steps = 60
    its = {}
    for order in (0, 2):
        print(f"c2 mixed {steps} steps, extrapolation order {order}", flush=True)
        x, y, z, kx, ky, kz = w.initial_state()
        sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, hvamd.HV_PREC_MIXED)
        try:
            sim.set_option(hvamd.HV_OPT_SOR_EXTRAPOLATE, order)
            assert sim.get_option(hvamd.HV_OPT_SOR_EXTRAPOLATE) == order
            sim.upload_particles(x, y, z, kx, ky, kz)
            prm = w.params(steps)
            sim.run_init(w.sky(), prm)
            _, its[order] = sim.run_steps(w.sky(), prm, 0, steps)
        finally:
            sim.close()
    m0, m2 = float(its[0][10:].mean()), float(its[2][10:].mean())
    print("mean SOR iterations per step (steps 10-59): last-potential start", m0, "extrapolated", m2)
    print("iterations, last-potential start:", list(its[0]), "extrapolated:", list(its[2]))
    for order in (0, 2):
        assert its[order][10:].max() < w.max_it, order
    assert m2 < 0.5 * m0


def test_fp64_rejects_extrapolation():
    w = configs.c0()
    sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, hvamd.HV_PREC_F64)
    try:
        assert sim.get_option(hvamd.HV_OPT_SOR_EXTRAPOLATE) == 0
        with pytest.raises(hvamd.HvError):
            sim.set_option(hvamd.HV_OPT_SOR_EXTRAPOLATE, 2)
    finally:
        sim.close()
