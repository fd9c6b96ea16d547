"""Pin the CPU oracle against the reference itself.

Two sources, both from the reference's own hot-path templates
(include/physics.hpp, include/tools.hpp) compiled by oracle/Makefile:
  - live: oracle/_ref/libhvref.so, where the reference checkout was present
    at build time (skipped otherwise);
  - committed golden vectors tests/golden/*.npz written by
    tests/golden/make_golden.py from the same library (always run).

Expected agreement (and why):
  - bit equality for Coulomb faces, SOR (omega 1.5, tol 1e-4), merge, FDM,
    gather, kick, drift and boost: the oracle keeps the reference's
    expressions and both are compiled with -ffp-contract=off;
  - deposit within 8 * 2^-41 per node: the oracle rounds each CIC weight to
    2^-40 fixed point (order-free sums);
  - smearing within 4e-15 of the field maximum: same tap set, factorised
    weights and nested summation order;
  - Skyrme within 1e-13 of max |U|: deterministic pow vs libm pow (a few ulp),
    then the reference's cancellations;
  - centre of mass / radius within 2^-30 quanta: fixed-point sums.
"""
from pathlib import Path

import numpy as np
import pytest

from inputs import SIGMA_R, particles, smooth_field

GOLDEN = Path(__file__).resolve().parent / "golden"


def _eq(a, b):
    assert np.array_equal(np.asarray(a), np.asarray(b))


def _kernel_checks(oracle, g, get):
    """g: dict of reference outputs; get: reference inputs."""
    n = tuple(int(v) for v in g["n"])
    d = tuple(float(v) for v in g["d_max"])
    np_, nn = int(g["np_"]), int(g["nn"])
    G = int(np.prod(n))
    sk = [float(v) for v in g["skyrme"]]
    x, y, z = get("in_x"), get("in_y"), get("in_z")
    counts, _, _ = oracle.deposit(x, y, z, np_, nn, n, d)
    assert np.abs(counts - g["counts"]).max() <= 8 * 2.0 ** -41
    rho = oracle.smear(g["counts"], n, d, float(g["sigma_r"]), int(g["ppn"]))
    assert np.abs(rho - g["rho"]).max() <= 4e-15 * np.abs(g["rho"]).max()
    rho_lit = oracle.smear(g["counts"], n, d, float(g["sigma_r"]), int(g["ppn"]), literal=True)
    assert np.abs(rho_lit - g["rho"]).max() <= 4e-15 * np.abs(g["rho"]).max()
    U = oracle.skyrme(g["rho"], n, d, *sk)
    assert np.abs(U - g["U"]).max() <= 1e-13 * np.abs(g["U"]).max()
    _eq(oracle.coulomb_bc(np.zeros(G), n, d, int(g["zprot"])), g["phi_bc"])
    phi, it, md, warn = oracle.sor(g["phi_bc"], g["rho"][:G], n, d, 1.5, 1e-4, 512)
    _eq(phi, g["phi"])
    assert md < 1e-4 and warn == 0 and 0 < it < 512
    _eq(oracle.merge(g["U"], g["phi"], n, d), g["U_merged"])
    F = oracle.fdm(g["U_merged"], n, d)
    for a, k in zip(F, ("Fx", "Fy", "Fz")):
        _eq(a, g[k])
    Fref = [g["Fx"], g["Fy"], g["Fz"]]
    f = oracle.gather(x, y, z, np_, nn, Fref, n, d)
    for a, k in zip(f, ("fx", "fy", "fz")):
        _eq(a, g[k])
    k = oracle.kick([get("in_kx"), get("in_ky"), get("in_kz")], [g["fx"], g["fy"], g["fz"]], float(g["dt"]))
    for a, key in zip(k, ("kx", "ky", "kz")):
        _eq(a, g[key])
    xs = oracle.drift([x, y, z], [g["kx"], g["ky"], g["kz"]], float(g["dt"]))
    for a, key in zip(xs, ("x", "y", "z")):
        _eq(a, g[key])
    _eq(oracle.excite(g["kz"], np_, nn, float(g["eta"]), int(g["zprot"]), int(g["nnuc"]), 0), g["kz_exc"])
    st = oracle.stats(g["x"], g["y"], g["z"], np_, nn, n, d)
    assert np.abs(st - g["stats"]).max() <= 2e-9 * max(1.0, np.abs(g["stats"]).max())


def test_kernels_vs_golden(oracle):
    g = dict(np.load(GOLDEN / "kernels_12x10x14.npz"))
    _kernel_checks(oracle, g, lambda k: np.ascontiguousarray(g[k]))


def test_loop_vs_golden(oracle):
    """The oracle step loop against the reference's cpu_simulate call
    sequence (8 steps, boost at t=1 fm/c, reference SOR settings)."""
    g = dict(np.load(GOLDEN / "loop_12x10x14_8steps.npz"))
    n = tuple(int(v) for v in g["n"])
    d = tuple(float(v) for v in g["d_max"])
    sk = [float(v) for v in g["skyrme"]]
    init = {k: np.ascontiguousarray(g["in_" + k]) for k in ("x", "y", "z", "kx", "ky", "kz")}
    steps = int(g["steps"])
    out, stats, its, _, warn = oracle.simulate(init, int(g["np_"]), int(g["nn"]), n, d, *sk, float(g["dt"]), steps,
                                               float(g["t_exc"]), float(g["eta"]), 0, int(g["zprot"]),
                                               int(g["nnuc"]), float(g["sigma_r"]), int(g["ppn"]), 1.5, 1e-4, 512)
    assert warn == 0
    # deposit quantisation (1e-12) and smearing order (1e-16) are the only
    # differences; over 8 steps they stay far below 1e-9 relative
    for k in ("x", "y", "z", "kx", "ky", "kz"):
        scale = np.abs(g["out_" + k]).max()
        assert np.abs(out[k] - g["out_" + k]).max() <= 1e-9 * scale, k
    for k in ("fx", "fy", "fz"):
        scale = np.abs(g["out_" + k]).max()
        assert np.abs(out[k] - g["out_" + k]).max() <= 1e-7 * scale, k
    assert np.abs(stats - g["stats"]).max() <= 1e-9 * np.abs(g["stats"]).max()


@pytest.mark.parametrize("n,d", [((16, 14, 18), (6.0, 5.5, 7.0)), ((24, 24, 24), (6.0, 6.0, 6.0))])
def test_kernels_vs_live_reference(oracle, ref, n, d):
    """Same checks against the live reference build on fresh inputs (multi-
    threaded reference runs: its order-dependent sums are compared with
    tolerance anyway)."""
    G = int(np.prod(n))
    np_, nn = 300, 420
    st = particles(5, np_, nn, 0.6 * min(d), spill=0.02)
    g = {"n": np.array(n), "d_max": np.array(d), "np_": np_, "nn": nn, "sigma_r": SIGMA_R, "ppn": 9,
         "skyrme": np.array([-356.0, 303.0, 32.0, 7.0 / 6.0]), "zprot": 82, "nnuc": 126, "dt": 0.5, "eta": 0.076}
    for k in ("x", "y", "z", "kx", "ky", "kz"):
        g["in_" + k] = st[k]
    g["counts"] = ref.deposit(st["x"], st["y"], st["z"], np_, nn, n, d)
    g["rho"] = ref.smear(g["counts"], n, d, SIGMA_R, 9)
    g["U"] = ref.skyrme(g["rho"], n, d, -356.0, 303.0, 32.0, 7.0 / 6.0)
    g["phi_bc"] = ref.coulomb_bc(np.zeros(G), n, d, 82)
    g["phi"] = ref.sor(g["phi_bc"], g["rho"][:G], n, d)
    g["U_merged"] = ref.merge(g["U"], g["phi"], n, d)
    g["Fx"], g["Fy"], g["Fz"] = ref.fdm(g["U_merged"], n, d)
    g["fx"], g["fy"], g["fz"] = ref.gather(st["x"], st["y"], st["z"], np_, nn, [g["Fx"], g["Fy"], g["Fz"]], n, d)
    g["kx"], g["ky"], g["kz"] = ref.kick([st["kx"], st["ky"], st["kz"]], [g["fx"], g["fy"], g["fz"]], 0.5)
    g["x"], g["y"], g["z"] = ref.drift([st["x"], st["y"], st["z"]], [g["kx"], g["ky"], g["kz"]], 0.5)
    g["kz_exc"] = ref.excite(g["kz"], np_, nn, 0.076, 82, 126)
    g["stats"] = np.array([ref.center_of_mass(g["z"], np_, nn, 0), ref.center_of_mass(g["z"], np_, nn, 1),
                           ref.mean_squared_radius(g["x"], g["y"], g["z"], np_, nn, n, d, 0),
                           ref.mean_squared_radius(g["x"], g["y"], g["z"], np_, nn, n, d, 1)])
    _kernel_checks(oracle, g, lambda k: g[k])


def test_smear_literal_vs_reference_full(oracle, ref):
    """The reference's O(G^2) smearing loop against the oracle's term-ordered
    stencil and its factorised form on a dense random count field."""
    n, d = (14, 12, 16), (5.0, 4.0, 6.0)
    rng = np.random.default_rng(3)
    counts = rng.uniform(0, 5, 2 * int(np.prod(n)))
    want = ref.smear(counts, n, d, SIGMA_R, 13)
    for literal in (False, True):
        got = oracle.smear(counts, n, d, SIGMA_R, 13, literal=literal)
        assert np.abs(got - want).max() <= 4e-15 * np.abs(want).max()


def test_sor_omega_param_matches_reference_default(oracle, ref):
    n, d = (20, 20, 20), (6.0, 6.0, 6.0)
    G = int(np.prod(n))
    rho = smooth_field(9, n, 0.07)
    phi0 = ref.coulomb_bc(np.zeros(G), n, d, 26)
    want = ref.sor(phi0, rho, n, d)
    got, it, md, _ = oracle.sor(phi0, rho, n, d, 1.5, 1e-4, 512)
    _eq(got, want)
