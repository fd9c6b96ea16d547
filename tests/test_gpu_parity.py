"""GPU parity: libhvamd.so (through its C-ABI) against the CPU oracle.

FP64 is held to bit equality: every kernel keeps the oracle's operation order,
both sides are compiled with -ffp-contract=off, the CIC deposit and the
diagnostics sums are exact integer sums, the Skyrme power is the shared
deterministic hv_det_pow and the smearing weights / Coulomb faces come from the
same host-side tables.  The mixed-precision path is held to the tolerances
stated in each test.
"""
import ctypes

import numpy as np
import pytest

import hvamd
from hvamd import configs
from inputs import SIGMA_R, particles, smooth_field

pytestmark = pytest.mark.gpu

GRIDS = [
    ((32, 32, 32), (8.0, 8.0, 8.0)),      # c0 grid
    ((20, 16, 24), (6.0, 5.0, 7.5)),      # anisotropic, non-power-of-two
]


def _sim(n, d, np_, nn, prec=hvamd.HV_PREC_F64):
    return hvamd.Simulation(hvamd.World(n, d), np_, nn, prec)


def _eq(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape
    if not np.array_equal(a, b):
        bad = np.flatnonzero(a.ravel() != b.ravel())
        i = bad[0]
        raise AssertionError(f"{bad.size} of {a.size} differ; first at {i}: {a.ravel()[i]!r} vs {b.ravel()[i]!r}")


@pytest.mark.parametrize("n,d", GRIDS)
def test_deposit_bitwise(oracle, n, d):
    print(f"deposit grid={n}", flush=True)
    st = particles(11, 700, 900, 0.6 * min(d), spill=0.01)
    sim = _sim(n, d, 700, 900)
    sim.upload_particles(**st)
    warn = sim.distribute_volumetric_particles_cic()
    got = sim.download_field(hvamd.HV_FIELD_DENSITY)
    want, _, skipped = oracle.deposit(st["x"], st["y"], st["z"], 700, 900, n, d)
    _eq(got, want)
    assert skipped > 0 and (warn & 2)


# smearing forms: HV_OPT_SMEAR_FUSED
FORMS = {"two-pass": 0, "fused": 1}


def _set_form(sim, form):
    sim.set_option(hvamd.HV_OPT_SMEAR_FUSED, FORMS[form])
    assert sim.get_option(hvamd.HV_OPT_SMEAR_FUSED) == FORMS[form]


@pytest.mark.parametrize("form", ["two-pass", "fused"])
@pytest.mark.parametrize("n,d", GRIDS + [((64, 60, 64), (16.0, 15.0, 16.0)),   # chunks along i
                                         ((24, 20, 18), (6.0, 5.0, 4.5))])      # nz not a multiple of 4
def test_smear_bitwise(oracle, n, d, form):
    """Both smearing forms (two-pass; fused plane walk) against the oracle."""
    print(f"smear grid={n} form={form}", flush=True)
    st = particles(12, 500, 600, 0.5 * min(d))
    counts, _, _ = oracle.deposit(st["x"], st["y"], st["z"], 500, 600, n, d)
    sim = _sim(n, d, 500, 600)
    _set_form(sim, form)
    sim.upload_field(hvamd.HV_FIELD_DENSITY, counts)
    sim.compute_volumetric_densities(SIGMA_R, 100)
    got = sim.download_field(hvamd.HV_FIELD_DENSITY)
    _eq(got, oracle.smear(counts, n, d, SIGMA_R, 100))


@pytest.mark.parametrize("n,d", GRIDS + [((64, 60, 64), (16.0, 15.0, 16.0)), ((96, 96, 96), (19.2, 19.2, 19.2)),
                                         ((24, 20, 18), (6.0, 5.0, 4.5))])
def test_smear_mixed_near_oracle(oracle, n, d):
    """The mixed form smears in FP32 with one rounding per tap (a fused
    multiply-add): within 2e-5 of the largest value of the oracle's FP64
    smearing (a few hundred taps per node at FP32's 6e-8 unit roundoff), and
    every form gives the same bits."""
    print(f"smear mixed grid={n}", flush=True)
    st = particles(12, 500, 600, 0.5 * min(d))
    counts, _, _ = oracle.deposit(st["x"], st["y"], st["z"], 500, 600, n, d)
    want = oracle.smear(counts, n, d, SIGMA_R, 100)
    got = []
    for form in ("two-pass", "fused"):
        sim = _sim(n, d, 500, 600, hvamd.HV_PREC_MIXED)
        _set_form(sim, form)
        sim.upload_field(hvamd.HV_FIELD_DENSITY, counts)
        sim.compute_volumetric_densities(SIGMA_R, 100)
        got.append(sim.download_field(hvamd.HV_FIELD_DENSITY))
        sim.close()
    _eq(got[0], got[1])
    err = np.abs(got[0] - want).max() / np.abs(want).max()
    print("max error / max", err)
    assert err < 2e-5


@pytest.mark.parametrize("prec", [hvamd.HV_PREC_F64, hvamd.HV_PREC_MIXED])
@pytest.mark.parametrize("n,d,radius,spill,sort_every", [
    ((32, 32, 32), (8.0, 8.0, 8.0), 4.0, 0.0, 8),          # the c0 nucleus
    ((48, 40, 56), (12.0, 10.0, 14.0), 3.0, 0.01, 8),     # small nucleus, large anisotropic box, some outside
    ((48, 40, 56), (12.0, 10.0, 14.0), 3.0, 0.01, 0),     # unsorted: the deposit's per-particle atomics
    ((20, 16, 24), (6.0, 5.0, 7.5), 5.5, 0.0, 8),         # nucleus over the whole box, counts on its faces
    ((40, 36, 30), (10.0, 9.0, 7.5), 3.5, 0.0, 8),        # nz not a multiple of 4: the one-node x pass (FP32)
])
def test_smear_count_box_loop_equals_fused(n, d, radius, spill, sort_every, prec):
    """The two-pass form reads only the count box that the deposit records (its
    z and y pass visits the box's planes and tiles, its x pass writes zeros
    outside the box grown by the radius); the fused form reads every count: a
    few steps of the loop give the same bits in both."""
    w = configs.c0()
    st = particles(21, w.n_protons, w.n_neutrons, radius, spill=spill)
    outs = []
    for form in ("fused", "two-pass"):
        print(f"count-box loop grid={n} prec={prec} sort={sort_every} form={form}", flush=True)
        sim = _sim(n, d, w.n_protons, w.n_neutrons, prec)
        _set_form(sim, form)
        sim.set_option(hvamd._lib.HV_OPT_SORT_EVERY, sort_every)
        sim.upload_particles(**st)
        stats, its = sim.simulate(w.sky(), w.params(6))
        outs.append((stats, list(its), sim.download_particles(), sim.download_field(hvamd.HV_FIELD_DENSITY)))
        sim.close()
    assert np.abs(outs[0][3]).max() > 0
    for o in outs[1:]:
        _eq(o[0], outs[0][0])
        assert o[1] == outs[0][1]
        for k in ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz"):
            _eq(o[2][k], outs[0][2][k])
        _eq(o[3], outs[0][3])


@pytest.mark.parametrize("n,d", GRIDS)
def test_skyrme_merge_fdm_bitwise(oracle, n, d):
    print(f"skyrme/merge/fdm grid={n}", flush=True)
    G = int(np.prod(n))
    rho = np.concatenate([smooth_field(1, n, 0.09), smooth_field(2, n, 0.1)])
    phi = smooth_field(3, n, 20.0)
    sim = _sim(n, d, 10, 10)
    sim.upload_field(hvamd.HV_FIELD_DENSITY, rho)
    sim.upload_field(hvamd.HV_FIELD_COULOMB, phi)
    sim.compute_volumetric_skyrme_potentials(hvamd.Skyrme(-356.0, 303.0, 32.0, 7.0 / 6.0))
    U = sim.download_field(hvamd.HV_FIELD_POTENTIALS)
    U_o = oracle.skyrme(rho, n, d, -356.0, 303.0, 32.0, 7.0 / 6.0)
    _eq(U, U_o)
    sim.merge_volumetric_potentials()
    M_o = oracle.merge(U_o, phi, n, d)
    _eq(sim.download_field(hvamd.HV_FIELD_POTENTIALS), M_o)
    sim.compute_volumetric_forces_fdm()
    F_o = oracle.fdm(M_o, n, d)
    for which, want in zip((hvamd.HV_FIELD_FORCES_X, hvamd.HV_FIELD_FORCES_Y, hvamd.HV_FIELD_FORCES_Z), F_o):
        _eq(sim.download_field(which), want)
    assert G == len(phi)


@pytest.mark.parametrize("n,d", GRIDS)
@pytest.mark.parametrize("omega,tol", [(1.5, 1e-4), (1.8, 1e-6)])
def test_sor_bitwise(oracle, n, d, omega, tol):
    print(f"sor grid={n} omega={omega} tol={tol}", flush=True)
    G = int(np.prod(n))
    rho_p = smooth_field(4, n, 0.08)
    phi0 = oracle.coulomb_bc(np.zeros(G), n, d, 50)
    sim = _sim(n, d, 10, 10)
    sim.upload_field(hvamd.HV_FIELD_DENSITY, np.concatenate([rho_p, np.zeros(G)]))
    sim.set_initial_coulomb_boundaries(50)
    _eq(sim.download_field(hvamd.HV_FIELD_COULOMB), phi0)
    it, md = sim.compute_volumetric_coulomb_potentials_sor(omega, tol, 512)
    phi_o, it_o, md_o, _ = oracle.sor(phi0, rho_p, n, d, omega, tol, 512)
    assert it == it_o and md == md_o
    _eq(sim.download_field(hvamd.HV_FIELD_COULOMB), phi_o)
    # warm start: a second solve after a small change converges in fewer steps
    rho2 = rho_p * 1.001
    sim.upload_field(hvamd.HV_FIELD_DENSITY, np.concatenate([rho2, np.zeros(G)]))
    it2, _ = sim.compute_volumetric_coulomb_potentials_sor(omega, tol, 512)
    phi_o2, it_o2, _, _ = oracle.sor(phi_o, rho2, n, d, omega, tol, 512)
    assert it2 == it_o2 and it2 < it_o
    _eq(sim.download_field(hvamd.HV_FIELD_COULOMB), phi_o2)


def test_sor_cap_warning(oracle):
    print("sor cap", flush=True)
    n, d = (24, 24, 24), (6.0, 6.0, 6.0)
    G = int(np.prod(n))
    sim = _sim(n, d, 10, 10)
    sim.upload_field(hvamd.HV_FIELD_DENSITY, np.concatenate([smooth_field(5, n, 0.08), np.zeros(G)]))
    sim.set_initial_coulomb_boundaries(82)
    it, md = sim.compute_volumetric_coulomb_potentials_sor(1.5, 1e-12, 7)
    assert it == 7 and md > 1e-12
    assert sim.last_warnings & hvamd._lib.HV_WARN_SOR_NOT_CONVERGED


@pytest.mark.parametrize("n,d", GRIDS)
def test_particle_kernels_bitwise(oracle, n, d):
    print(f"gather/kick/drift/excite/stats grid={n}", flush=True)
    np_, nn = 800, 1100
    st = particles(13, np_, nn, 0.7 * min(d), spill=0.01)
    F = [smooth_field(6 + a, n, 3.0, positive=False) for a in range(3)]
    F = [np.concatenate([f, 0.5 * f]) for f in F]
    sim = _sim(n, d, np_, nn)
    sim.upload_particles(**st)
    for which, f in zip((hvamd.HV_FIELD_FORCES_X, hvamd.HV_FIELD_FORCES_Y, hvamd.HV_FIELD_FORCES_Z), F):
        sim.upload_field(which, f)
    sim.distribute_forces_to_particles_cic()
    got = sim.download_particles()
    f_o = oracle.gather(st["x"], st["y"], st["z"], np_, nn, F, n, d)
    for k, want in zip(("fx", "fy", "fz"), f_o):
        _eq(got[k], want)
    sim.update_momenta_half(0.5)
    k_o = oracle.kick([st["kx"], st["ky"], st["kz"]], f_o, 0.5)
    sim.update_positions_full(0.5)
    x_o = oracle.drift([st["x"], st["y"], st["z"]], k_o, 0.5)
    sim.nuclear_excitation(0.1014, 20, 28, hvamd.HV_BOOST_REF_WEIGHTED)
    kz_o = oracle.excite(k_o[2], np_, nn, 0.1014, 20, 28, 0)
    got = sim.download_particles()
    for k, want in zip(("kx", "ky", "kz", "x", "y", "z"), (k_o[0], k_o[1], kz_o, *x_o)):
        _eq(got[k], want)
    st_o = oracle.stats(x_o[0], x_o[1], x_o[2], np_, nn, n, d)
    got_st = [sim.center_of_mass(hvamd.HV_PROTONS), sim.center_of_mass(hvamd.HV_NEUTRONS),
              sim.mean_squared_radius(hvamd.HV_PROTONS), sim.mean_squared_radius(hvamd.HV_NEUTRONS)]
    _eq(np.array(got_st), st_o)


def _oracle_run(oracle, w, st, steps):
    p = w.params(steps)
    s = w.sky()
    n, d = w.world().n, w.world().d_max
    return oracle.simulate(st, w.n_protons, w.n_neutrons, n, d, s.A, s.B, s.C, s.gamma, p.dt, steps, p.t_exc,
                           p.eta_exc, p.boost_mode, w.z, w.n, p.sigma_r, p.part_per_nucleon, p.sor_omega,
                           p.sor_tol, p.sor_max_it)


@pytest.mark.parametrize("steps", [1, 12])
def test_fused_loop_bitwise_c0(oracle, steps):
    """The whole step loop of the A=40, 32^3 workload, FP64, bit for bit."""
    print(f"fused loop c0 steps={steps}", flush=True)
    w = configs.c0()
    x, y, z, kx, ky, kz = w.initial_state()
    st = dict(x=x, y=y, z=z, kx=kx, ky=ky, kz=kz)
    sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, hvamd.HV_PREC_F64)
    sim.upload_particles(**st)
    stats, its = sim.simulate(w.sky(), w.params(steps))
    got = sim.download_particles()
    st_o, stats_o, its_o, _, _ = _oracle_run(oracle, w, st, steps)
    assert list(its) == list(its_o[1:])
    _eq(stats, stats_o)
    for k in ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz"):
        _eq(got[k], st_o[k])


@pytest.mark.parametrize("prec", [hvamd.HV_PREC_F64, hvamd.HV_PREC_MIXED])
def test_total_energy_matches_oracle(oracle, prec):
    """hv_total_energy on the loop's state equals the oracle's exact fixed-point
    evaluation of the same (downloaded) momenta, density and phi, bit for bit
    in both forms (the mixed form's FP32 values widen exactly)."""
    print(f"total energy prec={prec}", flush=True)
    w = configs.c0()
    x, y, z, kx, ky, kz = w.initial_state()
    sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, prec)
    sim.upload_particles(x, y, z, kx, ky, kz)
    prm = w.params(6)
    sk = w.sky()
    sim.run_init(sk, prm)
    e0 = sim.total_energy(sk, prm.part_per_nucleon)
    sim.run_steps(sk, prm, 0, 6)
    e1 = sim.total_energy(sk, prm.part_per_nucleon)
    st = sim.download_particles()
    rho = sim.download_field(hvamd.HV_FIELD_DENSITY)
    phi = sim.download_field(hvamd.HV_FIELD_COULOMB)
    sim.close()
    n, d = w.world().n, w.world().d_max
    want = oracle.total_energy((st["kx"], st["ky"], st["kz"]), rho, phi, n, d, sk.A, sk.B, sk.C, sk.gamma,
                               prm.part_per_nucleon)
    print("E init", e0, "after 6 steps", e1)
    _eq(e1, want)
    assert e1[1] > 0 and e1[2] < 0 and e1[3] > 0
    assert abs(e1[0] - e0[0]) < 0.05 * abs(e0[1])


def test_fused_loop_chunked_equals_single(oracle):
    """run_steps in chunks gives the same bits as one call (the boost and the
    split half kicks are carried across calls)."""
    print("fused loop chunked", flush=True)
    w = configs.c0()
    x, y, z, kx, ky, kz = w.initial_state()
    prm = w.params(9)
    prm.t_exc = 2.0  # boost at step 4
    outs = []
    for chunks in ([9], [3, 1, 5]):
        sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, hvamd.HV_PREC_F64)
        sim.upload_particles(x, y, z, kx, ky, kz)
        sim.run_init(w.sky(), prm)
        first = 0
        allst = []
        for c in chunks:
            s, _ = sim.run_steps(w.sky(), prm, first, c)
            allst.append(s)
            first += c
        outs.append((np.concatenate(allst), sim.download_particles()))
    _eq(outs[0][0], outs[1][0])
    for k in ("x", "kz", "fz"):
        _eq(outs[0][1][k], outs[1][1][k])


def test_fused_loop_matches_per_kernel_sequence(oracle):
    """hv_simulate equals the reference's call sequence driven kernel by kernel."""
    print("fused vs per-kernel", flush=True)
    w = configs.c0()
    x, y, z, kx, ky, kz = w.initial_state()
    prm = w.params(3)
    sk = w.sky()
    a = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons)
    a.upload_particles(x, y, z, kx, ky, kz)
    stats_a, _ = a.simulate(sk, prm)
    b = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons)
    b.upload_particles(x, y, z, kx, ky, kz)

    def solve():
        b.distribute_volumetric_particles_cic()
        b.compute_volumetric_densities(prm.sigma_r, prm.part_per_nucleon)
        b.compute_volumetric_skyrme_potentials(sk)
        b.compute_volumetric_coulomb_potentials_sor(prm.sor_omega, prm.sor_tol, prm.sor_max_it)
        b.merge_volumetric_potentials()
        b.compute_volumetric_forces_fdm()
        b.distribute_forces_to_particles_cic()

    b.set_initial_coulomb_boundaries(prm.z)
    solve()
    excited = False
    stats_b = []
    for step in range(prm.steps):
        stats_b.append([b.center_of_mass(0), b.center_of_mass(1), b.mean_squared_radius(0),
                        b.mean_squared_radius(1)])
        if step * prm.dt >= prm.t_exc and not excited:
            excited = True
            b.nuclear_excitation(prm.eta_exc, prm.z, prm.n, prm.boost_mode)
        b.update_momenta_half(prm.dt)
        b.update_positions_full(prm.dt)
        solve()
        b.update_momenta_half(prm.dt)
    _eq(stats_a, np.array(stats_b))
    pa, pb = a.download_particles(), b.download_particles()
    for k in ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz"):
        _eq(pa[k], pb[k])
    _eq(a.download_field(hvamd.HV_FIELD_COULOMB), b.download_field(hvamd.HV_FIELD_COULOMB))


def test_mixed_precision_tracks_fp64():
    """FP32 particles/grid with FP64 Coulomb and reductions stay close to the
    FP64 path over a short horizon (statistical quantities)."""
    print("mixed vs fp64", flush=True)
    w = configs.c0()
    x, y, z, kx, ky, kz = w.initial_state()
    res = {}
    for prec in (hvamd.HV_PREC_F64, hvamd.HV_PREC_MIXED):
        sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, prec)
        sim.upload_particles(x, y, z, kx, ky, kz)
        stats, _ = sim.simulate(w.sky(), w.params(40))
        res[prec] = hvamd.dipole(stats, w.z, w.n)
    d64, d32 = res[hvamd.HV_PREC_F64], res[hvamd.HV_PREC_MIXED]
    rel = np.abs(d32 - d64).max() / np.abs(d64).max()
    print("mixed vs fp64 D(t) max rel", rel)
    assert rel < 1e-4


def test_determinism_large():
    """Two runs of the 1.04M-particle, 128^3 workload give identical bits."""
    print("determinism c2 2 steps", flush=True)
    w = configs.c2()
    x, y, z, kx, ky, kz = w.initial_state()
    outs = []
    for _ in range(2):
        sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, hvamd.HV_PREC_F64)
        sim.upload_particles(x, y, z, kx, ky, kz)
        prm = w.params(2)
        prm.sor_max_it = 64
        stats, its = sim.simulate(w.sky(), prm)
        outs.append((stats, sim.download_particles()["z"], sim.download_field(hvamd.HV_FIELD_COULOMB)))
        sim.close()
    _eq(outs[0][0], outs[1][0])
    _eq(outs[0][1], outs[1][1])
    _eq(outs[0][2], outs[1][2])


def test_resident_solve_with_defect_and_fold_equals_three_launches():
    """In the mixed loop (extrapolated starts) the deep-halo resident FP32
    correction computes its residual and folds its correction itself; the
    lane-owned body (HV_OPT_SOR_F32_OWNED 1) still runs the defect launch, the
    solve and the fold launch.  Both give the same bits, step by step."""
    w = configs.c1()
    steps = 24
    outs = []
    for owned in (2, 1):
        print("mixed loop, resident body", owned, flush=True)
        sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, hvamd.HV_PREC_MIXED)
        sim.set_option(hvamd.HV_OPT_SOR_F32_OWNED, owned)
        sim.upload_particles(*w.initial_state())
        sim.run_init(w.sky(), w.params(steps))
        stats, its = sim.run_steps(w.sky(), w.params(steps), 0, steps)
        assert sim.get_option(hvamd.HV_OPT_SOR_RESIDENT_F32) == 1
        outs.append((stats, its, sim.download_particles(), sim.download_field(hvamd.HV_FIELD_COULOMB),
                     sim.get_option(hvamd.HV_OPT_SOR_F32_OWNED)))
        sim.close()
    print("bodies that ran:", outs[0][4], outs[1][4])
    assert outs[0][4] == 2 and outs[1][4] != 2  # (1, or the per-row body where that does not fit)
    _eq(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    for k in ("x", "y", "z", "kx", "ky", "kz"):
        _eq(outs[0][2][k], outs[1][2][k])
    _eq(outs[0][3], outs[1][3])


@pytest.mark.parametrize("sort_every", [0, 1, 3])
def test_sort_cadence_does_not_change_bits(oracle, sort_every):
    """Particle sorting (any cadence) leaves every result bit unchanged, and
    downloads come back in the caller's order."""
    print(f"sort cadence {sort_every}", flush=True)
    w = configs.c0()
    x, y, z, kx, ky, kz = w.initial_state()
    st = dict(x=x, y=y, z=z, kx=kx, ky=ky, kz=kz)
    steps = 7
    sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, hvamd.HV_PREC_F64)
    sim.set_option(hvamd._lib.HV_OPT_SORT_EVERY, sort_every)
    assert sim.get_option(hvamd._lib.HV_OPT_SORT_EVERY) == sort_every
    sim.upload_particles(**st)
    stats, its = sim.simulate(w.sky(), w.params(steps))
    got = sim.download_particles()
    st_o, stats_o, its_o, _, _ = _oracle_run(oracle, w, st, steps)
    _eq(stats, stats_o)
    for k in ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz"):
        _eq(got[k], st_o[k])
    # a partial upload after sorting lands on the right particles
    sim.upload_particles(fx=np.arange(w.total, dtype=np.float64))
    again = sim.download_particles()
    _eq(again["fx"], np.arange(w.total, dtype=np.float64))
    _eq(again["x"], st_o["x"])


@pytest.mark.parametrize("n,d", GRIDS + [((64, 64, 64), (12.8, 12.8, 12.8)), ((128, 128, 128), (16.0, 16.0, 16.0))])
def test_sor_resident_equals_per_sweep_launches(oracle, n, d):
    """The LDS-resident cooperative SOR gives the same iterate and iteration
    count as one launch per red/black sweep."""
    print(f"sor resident vs launches grid={n}", flush=True)
    G = int(np.prod(n))
    rho_p = smooth_field(14, n, 0.09)
    outs = []
    for resident in (1, 0):
        sim = _sim(n, d, 10, 10)
        sim.set_option(hvamd._lib.HV_OPT_SOR_RESIDENT, resident)
        sim.upload_field(hvamd.HV_FIELD_DENSITY, np.concatenate([rho_p, np.zeros(G)]))
        sim.set_initial_coulomb_boundaries(82)
        it, md = sim.compute_volumetric_coulomb_potentials_sor(1.8, 1e-6, 60)
        it2, md2 = sim.compute_volumetric_coulomb_potentials_sor(1.8, 1e-6, 200)
        used = sim.get_option(hvamd._lib.HV_OPT_SOR_RESIDENT)
        print(f"  resident option={resident} active={used} iters={it},{it2}", flush=True)
        outs.append((it, md, it2, md2, sim.download_field(hvamd.HV_FIELD_COULOMB), used))
        sim.close()
    assert outs[0][5] == 1, "resident SOR not available on this device"
    assert outs[0][:4] == outs[1][:4]
    _eq(outs[0][4], outs[1][4])
    if G <= 64 ** 3:
        phi0 = oracle.coulomb_bc(np.zeros(G), n, d, 82)
        phi_o, it_o, md_o, _ = oracle.sor(phi0, rho_p, n, d, 1.8, 1e-6, 60)
        phi_o, it_o2, md_o2, _ = oracle.sor(phi_o, rho_p, n, d, 1.8, 1e-6, 200)
        assert (it_o, md_o, it_o2, md_o2) == outs[0][:4]
        _eq(outs[0][4], phi_o)


@pytest.mark.parametrize("n,d", GRIDS + [((24, 40, 34), (6.0, 8.0, 7.0)), ((128, 128, 128), (16.0, 16.0, 16.0))])
def test_sor_fused_equals_per_sweep(oracle, n, d):
    """One launch per iteration (red + black fused, out of place, ring
    recomputed) gives the same bits as one launch per sweep."""
    print(f"sor fused vs per-sweep grid={n}", flush=True)
    G = int(np.prod(n))
    rho_p = smooth_field(15, n, 0.09)
    outs = []
    for fused in (1, 0):
        sim = _sim(n, d, 10, 10)
        sim.set_option(hvamd._lib.HV_OPT_SOR_FUSED, fused)
        sim.upload_field(hvamd.HV_FIELD_DENSITY, np.concatenate([rho_p, np.zeros(G)]))
        sim.set_initial_coulomb_boundaries(82)
        r1 = sim.compute_volumetric_coulomb_potentials_sor(1.8, 1e-6, 37)
        r2 = sim.compute_volumetric_coulomb_potentials_sor(1.8, 1e-6, 300)
        outs.append((r1, r2, sim.download_field(hvamd.HV_FIELD_COULOMB)))
        sim.close()
    assert outs[0][:2] == outs[1][:2]
    _eq(outs[0][2], outs[1][2])


@pytest.mark.parametrize("form", ["rows", "rows_two_per_cu", "owned", "deep", "deep_two_per_cu"])
@pytest.mark.parametrize("n", [(48, 48, 48), (40, 36, 50), (64, 64, 64), (96, 96, 96), (128, 128, 128)])
def test_resident_f32_equals_sweeps_at_fixed_iterations(n, form):
    """With a tolerance no iteration reaches, the FP32 correction runs exactly
    max_it iterations in both forms, so the LDS-resident kernel (its tiles,
    row sets and lane mappings; the body that derives node indexes per sweep,
    one or two tiles per CU, the lane-owned body, or the deep-halo body with
    one exchange per iteration, one or two tiles per CU) and the per-sweep
    launches must give the same potential bit for bit."""
    print("resident FP32 vs sweeps, fixed iterations", n, form, flush=True)
    per_cu = 2 if form.endswith("two_per_cu") else 1
    owned = {"owned": 1, "deep": 2, "deep_two_per_cu": 2}.get(form, 0)
    d = tuple(0.2 * x for x in n)
    G = int(np.prod(n))
    rho_p = smooth_field(16, n, 0.09)
    out = {}
    for res32 in (0, 1):
        sim = _sim(n, d, 10, 10, hvamd.HV_PREC_MIXED)
        sim.set_option(hvamd.HV_OPT_SOR_RESIDENT_F32, res32)
        sim.set_option(hvamd.HV_OPT_SOR_F32_TILES_PER_CU, per_cu)
        sim.set_option(hvamd.HV_OPT_SOR_F32_OWNED, owned)
        sim.upload_field(hvamd.HV_FIELD_DENSITY, np.concatenate([rho_p, np.zeros(G)]))
        sim.set_initial_coulomb_boundaries(82)
        sim.compute_volumetric_coulomb_potentials_sor(1.8, 1e-6, 512)  # cold: FP64
        sim.upload_field(hvamd.HV_FIELD_DENSITY, np.concatenate([rho_p * 1.01, np.zeros(G)]))
        it, _ = sim.compute_volumetric_coulomb_potentials_sor(1.8, 1e-30, 40)
        assert sim.get_option(hvamd.HV_OPT_SOR_RESIDENT_F32) == res32
        if res32:
            assert sim.get_option(hvamd.HV_OPT_SOR_F32_OWNED) == owned
        out[res32] = (it, sim.download_field(hvamd.HV_FIELD_COULOMB))
        sim.close()
    assert out[0][0] == out[1][0] == 40 + 4
    _eq(out[0][1], out[1][1])


def test_resident_f32_tag_wraps():
    """The deep-halo body's exchange words carry the low 16 bits of the
    iteration tag.  Past 65,536 iterations the tag wraps to 0, the value the
    buffers are cleared to; a word still holds the tag of two iterations
    earlier then, so the solve must go on exactly as before: 65,600 FP32
    iterations (tolerance 0 never stops) equal the per-sweep launches bit for
    bit, with no wait running into its spin limit."""
    print("resident FP32 tag wrap, 48^3, 65600 iterations", flush=True)
    n = (48, 48, 48)
    d = tuple(0.2 * x for x in n)
    G = int(np.prod(n))
    rho_p = smooth_field(17, n, 0.09)
    out = {}
    for res32 in (0, 1):
        sim = _sim(n, d, 10, 10, hvamd.HV_PREC_MIXED)
        sim.set_option(hvamd.HV_OPT_SOR_RESIDENT_F32, res32)
        sim.upload_field(hvamd.HV_FIELD_DENSITY, np.concatenate([rho_p, np.zeros(G)]))
        sim.set_initial_coulomb_boundaries(82)
        sim.compute_volumetric_coulomb_potentials_sor(1.8, 1e-6, 512)  # cold: FP64
        sim.upload_field(hvamd.HV_FIELD_DENSITY, np.concatenate([rho_p * 1.02, np.zeros(G)]))
        it, _ = sim.compute_volumetric_coulomb_potentials_sor(1.8, 0.0, 65600)
        if res32:
            assert sim.get_option(hvamd.HV_OPT_SOR_F32_OWNED) == 2
        out[res32] = (it, sim.download_field(hvamd.HV_FIELD_COULOMB))
        sim.close()
    assert out[0][0] == out[1][0] == 65600 + 4
    _eq(out[0][1], out[1][1])


@pytest.mark.parametrize("res32", [0, 1])
def test_mixed_sor_matches_fp64_iterations(res32):
    """With the reference's omega (HV_OPT_SOR_F32_OMEGA = 1) the mixed form's
    FP32 defect-correction SOR follows the FP64 iteration: the FP32 correction
    carries the same iterates up to rounding, so the count matches the FP64
    solve within a few final FP64 sweeps (plus the lag of the LDS-resident
    FP32 iteration's convergence test), and the potential agrees within the
    stop tolerance's reach.  With the default factor (the grid rule) the
    correction takes another path to the same fixed point: the FP64 sweeps
    still end on the reference stop rule, and the potential is at least as
    close to the exact solution (a solve to tol 1e-11) as the FP64 solve's own
    (within a factor 2)."""
    print("mixed SOR vs fp64, resident FP32", res32, flush=True)
    n, d = (48, 48, 48), (9.6, 9.6, 9.6)
    G = int(np.prod(n))
    rho_p = smooth_field(16, n, 0.09)
    res = {}
    for key, prec, omega32, tol in (("f64", hvamd.HV_PREC_F64, 0, 1e-6), ("exact", hvamd.HV_PREC_F64, 0, 1e-11),
                                    ("ref", hvamd.HV_PREC_MIXED, 1, 1e-6), ("rule", hvamd.HV_PREC_MIXED, 0, 1e-6)):
        sim = _sim(n, d, 10, 10, prec)
        sim.set_option(hvamd.HV_OPT_SOR_RESIDENT_F32, res32)
        sim.set_option(hvamd.HV_OPT_SOR_F32_OMEGA, omega32)
        sim.upload_field(hvamd.HV_FIELD_DENSITY, np.concatenate([rho_p, np.zeros(G)]))
        sim.set_initial_coulomb_boundaries(82)
        it0, _ = sim.compute_volumetric_coulomb_potentials_sor(1.8, 1e-6, 512)  # cold: FP64 in both
        sim.upload_field(hvamd.HV_FIELD_DENSITY, np.concatenate([rho_p * 1.01, np.zeros(G)]))
        it1, md1 = sim.compute_volumetric_coulomb_potentials_sor(1.8, tol, 20000)  # warm: FP32 correction
        used = sim.get_option(hvamd.HV_OPT_SOR_RESIDENT_F32)
        res[key] = (it0, it1, md1, sim.download_field(hvamd.HV_FIELD_COULOMB))
        if prec == hvamd.HV_PREC_MIXED:
            assert used == res32
        sim.close()
    f, m, r, e = res["f64"], res["ref"], res["rule"], res["exact"]
    print("iters fp64", f[:2], "mixed (reference omega)", m[:2], "mixed (grid rule)", r[:2])
    assert f[0] == m[0] == r[0]
    assert abs(f[1] - m[1]) <= 12 + 2 * res32 and m[2] < 1e-6
    assert np.abs(f[3] - m[3]).max() < 2e-5
    ef, er = np.abs(f[3] - e[3]).max(), np.abs(r[3] - e[3]).max()
    print("max |phi - exact|: fp64", ef, "mixed (grid rule)", er)
    assert r[2] < 1e-6
    assert er <= 2 * ef + 1e-7


def test_f32_omega_option_values():
    """HV_OPT_SOR_F32_OMEGA 1 (the run's omega) and 1800 (omega / 1000) give
    the same bits on a run whose omega is 1.8; values outside 0, 1,
    1001..1999 are refused."""
    n, d = (32, 32, 32), (8.0, 8.0, 8.0)
    G = int(np.prod(n))
    rho_p = smooth_field(16, n, 0.09)
    out = []
    for v in (1, 1800):
        sim = _sim(n, d, 10, 10, hvamd.HV_PREC_MIXED)
        sim.set_option(hvamd.HV_OPT_SOR_F32_OMEGA, v)
        assert sim.get_option(hvamd.HV_OPT_SOR_F32_OMEGA) == v
        sim.upload_field(hvamd.HV_FIELD_DENSITY, np.concatenate([rho_p, np.zeros(G)]))
        sim.set_initial_coulomb_boundaries(82)
        sim.compute_volumetric_coulomb_potentials_sor(1.8, 1e-6, 512)
        sim.upload_field(hvamd.HV_FIELD_DENSITY, np.concatenate([rho_p * 1.01, np.zeros(G)]))
        it, _ = sim.compute_volumetric_coulomb_potentials_sor(1.8, 1e-6, 512)
        out.append((it, sim.download_field(hvamd.HV_FIELD_COULOMB)))
        for bad in (2, 1000, 2000, -1):
            with pytest.raises(RuntimeError):
                sim.set_option(hvamd.HV_OPT_SOR_F32_OMEGA, bad)
        sim.close()
    assert out[0][0] == out[1][0]
    _eq(out[0][1], out[1][1])


@pytest.mark.parametrize("prec", [hvamd.HV_PREC_F64, hvamd.HV_PREC_MIXED])
def test_write_vtk_snapshot_of_loop_state(oracle, tmp_path, prec):
    """hv_write_vtk after a few loop steps writes the reference's snapshot of
    the context's density, merged potentials and their FDM forces.  FP64: the
    forces are the oracle's FDM (bitwise the reference's) and the file equals
    the host writer's (byte-exact to the reference's writer on its own) and,
    where the reference harness is built, the reference writer's.  Mixed form:
    the same with the device's FP32 forces and the float instantiation's
    header and p + n sums."""
    print(f"vtk snapshot prec={prec}", flush=True)
    from oracle_lib import Reference
    w = configs.c0()
    sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, prec)
    sim.upload_particles(*w.initial_state())
    prm = w.params(4)
    sim.run_init(w.sky(), prm)
    sim.run_steps(w.sky(), prm, 0, 4)
    got = tmp_path / "got.vtk"
    assert sim.lib.hv_write_vtk(sim.ctx, str(got).encode()) == 0
    rho = sim.download_field(hvamd.HV_FIELD_DENSITY)
    U = sim.download_field(hvamd.HV_FIELD_POTENTIALS)
    n, d = w.world().n, w.world().d_max
    single = prec == hvamd.HV_PREC_MIXED
    if single:
        Fx, Fy, Fz = (sim.download_field(f) for f in (hvamd.HV_FIELD_FORCES_X, hvamd.HV_FIELD_FORCES_Y,
                                                        hvamd.HV_FIELD_FORCES_Z))
    else:
        Fx, Fy, Fz = oracle.fdm(U, n, d)
    sim.close()
    want = tmp_path / "want.vtk"
    wd = hvamd._lib.HvWorld((ctypes.c_int * 3)(*n), (ctypes.c_double * 3)(*d))
    arr = [np.ascontiguousarray(v, dtype=np.float64) for v in (rho, U, Fx, Fy, Fz)]
    assert hvamd.load().hv_vtk_write_fields(str(want).encode(), ctypes.byref(wd),
                                            *[a.ctypes.data_as(ctypes.c_void_p) for a in arr],
                                            1 if single else 0) == 0
    assert got.read_bytes() == want.read_bytes()
    if Reference.available and not single:
        ref = tmp_path / "ref.vtk"
        Reference().write_vtk(ref, n, d, Fx, Fy, Fz, rho, U)
        assert got.read_bytes() == ref.read_bytes()
