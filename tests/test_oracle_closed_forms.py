"""Closed-form and invariant pins of the CPU oracle, chosen so that a plausible
slip (a swapped weight, a wrong sign, a transposed index, a dropped term)
fails at least one of them.  No GPU."""
import math

import numpy as np
import pytest

from inputs import SIGMA_R, particles, smooth_field

N3 = (16, 12, 20)
D3 = (4.0, 3.0, 5.0)  # spacing 0.5 on every axis (dyadic, exact positions)


def node_pos(n, d):
    return [d[a] * (2.0 * np.arange(n[a]) / n[a] - 1.0) for a in range(3)]


# ---- deposit -------------------------------------------------------------------
def test_deposit_particle_on_node_is_unit(oracle):
    xs, ys, zs = node_pos(N3, D3)
    x = np.array([xs[3], xs[10]])
    y = np.array([ys[5], ys[2]])
    z = np.array([zs[7], zs[19]])
    counts, acc, skipped = oracle.deposit(x, y, z, 1, 1, N3, D3)
    G = int(np.prod(N3))
    assert skipped == 0
    idx_p = (3 * N3[1] + 5) * N3[2] + 7
    idx_n = G + (10 * N3[1] + 2) * N3[2] + 19
    assert counts[idx_p] == 1.0 and counts[idx_n] == 1.0
    assert np.count_nonzero(counts) == 2


def test_deposit_mass_and_first_moment(oracle):
    """Interior particles: weights sum to 1 and the weighted node positions
    reproduce the particle position (CIC preserves the first moment); a swapped
    (d, t) pair or a transposed axis breaks this."""
    rng = np.random.default_rng(0)
    h = 0.5
    n = 200
    x = rng.uniform(-D3[0] + h, D3[0] - 2 * h, n)
    y = rng.uniform(-D3[1] + h, D3[1] - 2 * h, n)
    z = rng.uniform(-D3[2] + h, D3[2] - 2 * h, n)
    counts, _, _ = oracle.deposit(x, y, z, n, 0, N3, D3)
    G = int(np.prod(N3))
    c = counts[:G].reshape(N3)
    xs, ys, zs = node_pos(N3, D3)
    assert abs(c.sum() - n) <= n * 8 * 2.0 ** -41
    assert abs((c.sum(axis=(1, 2)) * xs).sum() - x.sum()) < 1e-9
    assert abs((c.sum(axis=(0, 2)) * ys).sum() - y.sum()) < 1e-9
    assert abs((c.sum(axis=(0, 1)) * zs).sum() - z.sum()) < 1e-9
    assert counts[G:].sum() == 0.0


def test_deposit_upper_layer_drops_corners_and_outside(oracle):
    """Corners with index n are dropped (include/tools.hpp:67-94) and
    particles outside [0, n) or NaN are skipped."""
    h = 0.5
    x = np.array([D3[0] - 0.25 * h, 0.1, np.nan, 100.0, -D3[0] - 1e-9])
    y = np.array([0.0, 0.0, 0.0, 0.0, 0.0])
    z = np.array([0.0, D3[2] - 0.5 * h, 0.0, 0.0, 0.0])
    counts, _, skipped = oracle.deposit(x, y, z, 5, 0, N3, D3)
    assert skipped == 3
    # particle 0 keeps the t_x = 0.25 share, particle 1 the t_z = 0.5 share
    assert abs(counts.sum() - (0.25 + 0.5)) < 1e-12


def test_deposit_linear_in_particles(oracle):
    st = particles(1, 100, 130, 3.0)
    a, _, _ = oracle.deposit(st["x"], st["y"], st["z"], 100, 130, N3, D3)
    a1, acc1, _ = oracle.deposit(st["x"][:100], st["y"][:100], st["z"][:100], 100, 0, N3, D3)
    a2, acc2, _ = oracle.deposit(st["x"][100:], st["y"][100:], st["z"][100:], 0, 130, N3, D3)
    assert np.array_equal(a, a1 + a2)


def test_deposit_gather_transpose(oracle):
    """<deposit(p), F> = sum_i gather(F)(x_i): the two use one weight rule."""
    st = particles(2, 150, 170, 3.0)
    F = smooth_field(1, N3, 2.0, positive=False)
    F2 = np.concatenate([F, -0.5 * F])
    counts, _, _ = oracle.deposit(st["x"], st["y"], st["z"], 150, 170, N3, D3)
    f = oracle.gather(st["x"], st["y"], st["z"], 150, 170, [F2, F2, F2], N3, D3)
    lhs = float(counts @ F2)
    rhs = math.fsum(f[0])
    assert abs(lhs - rhs) <= 1e-9 * (np.abs(counts) @ np.abs(F2))


# ---- smearing ----------------------------------------------------------------------
@pytest.mark.parametrize("n,d", [((16, 16, 16), (4.0, 4.0, 4.0)), ((20, 18, 22), (4.0, 3.6, 4.4))])
def test_smear_unit_count_is_truncated_gaussian(oracle, n, d):
    G = int(np.prod(n))
    p = (n[0] // 2, n[1] // 2 - 1, n[2] // 2 + 1)
    counts = np.zeros(2 * G)
    counts[G + (p[0] * n[1] + p[1]) * n[2] + p[2]] = 1.0  # neutron species
    ppn = 5
    rho = oracle.smear(counts, n, d, SIGMA_R, ppn)
    h = [2.0 * d[a] / n[a] for a in range(3)]
    i, j, k = np.meshgrid(*[np.arange(a) for a in n], indexing="ij")
    d2 = ((i - p[0]) * h[0]) ** 2 + ((j - p[1]) * h[1]) ** 2 + ((k - p[2]) * h[2]) ** 2
    want = np.where(d2 <= 16 * SIGMA_R ** 2, np.exp(-d2 / (2 * SIGMA_R ** 2)), 0.0)
    want = want / ppn / (2 * np.pi * SIGMA_R ** 2) ** 1.5
    assert np.all(rho[:G] == 0.0)
    assert np.abs(rho[G:] - want.ravel()).max() <= 1e-14 * want.max()
    # truncated Gaussian keeps P(r <= 4 sigma) = erf(2 sqrt 2) - 4 sqrt(2/pi) e^-8 of the mass
    frac = math.erf(4 / math.sqrt(2)) - math.sqrt(2 / math.pi) * 4 * math.exp(-8)
    mass = rho[G:].sum() * h[0] * h[1] * h[2] * ppn
    assert abs(mass - frac) < 0.02


def test_smear_is_symmetric_and_clipped(oracle):
    n, d = (12, 14, 10), (3.0, 3.5, 2.5)
    rng = np.random.default_rng(5)
    a = rng.uniform(0, 1, 2 * int(np.prod(n)))
    b = rng.uniform(0, 1, 2 * int(np.prod(n)))
    Sa = oracle.smear(a, n, d, SIGMA_R, 3)
    Sb = oracle.smear(b, n, d, SIGMA_R, 3)
    assert abs(Sa @ b - a @ Sb) <= 1e-13 * abs(Sa @ b)


def test_smear_brute_force_tiny(oracle):
    """O(G^2) brute force over all node pairs on an 8^3 grid."""
    n, d = (8, 8, 8), (2.0, 2.0, 2.0)
    G = 512
    rng = np.random.default_rng(6)
    counts = rng.uniform(0, 2, 2 * G)
    pos = np.stack(np.meshgrid(*node_pos(n, d), indexing="ij"), -1).reshape(G, 3)
    d2 = ((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1)
    K = np.where(d2 <= 16 * SIGMA_R ** 2, np.exp(-d2 / (2 * SIGMA_R ** 2)), 0.0)
    term = 1.0 / 4 / (2 * np.pi * SIGMA_R ** 2) ** 1.5
    want = np.concatenate([K @ counts[:G], K @ counts[G:]]) * term
    got = oracle.smear(counts, n, d, SIGMA_R, 4)
    assert np.abs(got - want).max() <= 1e-14 * want.max()


# ---- Skyrme -----------------------------------------------------------------------
def test_skyrme_special_points(oracle):
    A, B, C, g = -356.0, 303.0, 32.0, 7.0 / 6.0
    n, d = (4, 1, 1), (1.0, 1.0, 1.0)
    rho_p = np.array([0.0, 0.08, 0.0, 0.16])
    rho_n = np.array([0.0, 0.08, 0.16, 0.0])
    U = oracle.skyrme(np.concatenate([rho_p, rho_n]), n, d, A, B, C, g)
    Up, Un = U[:4], U[4:]
    assert Up[0] == 0.0 and Un[0] == 0.0
    assert abs(Up[1] - (A + B)) < 1e-12 and abs(Un[1] - (A + B)) < 1e-12
    assert abs(Up[2] - (A + B - 2 * C)) < 1e-12 and abs(Un[2] - (A + B + 2 * C)) < 1e-12
    assert abs(Up[3] - (A + B + 2 * C)) < 1e-12 and abs(Un[3] - (A + B - 2 * C)) < 1e-12


def test_deterministic_pow_accuracy(oracle):
    """hv_det_pow (the Skyrme t^gamma) against numpy's pow: within 48 ulp over
    densities 1e-12 .. 4 rho0 (exp(y log x) carries log's rounding times
    |y log x| <= 32), within 12 ulp over the nuclear range 1e-3 .. 3 rho0."""
    t = np.concatenate([np.geomspace(1e-12, 4.0, 50000), np.random.default_rng(1).uniform(1e-3, 3.0, 50000)])
    G = t.size
    rho = np.concatenate([t * 0.16, np.zeros(G)])
    for gamma in (7.0 / 6.0, 2.0, 0.5):
        U = oracle.skyrme(rho, (G, 1, 1), (1.0, 1.0, 1.0), 0.0, 1.0, 0.0, gamma)
        tt = rho[:G] / 0.16
        want = np.power(tt, gamma)
        ulp = np.abs(U[:G] - want) / np.spacing(want)
        assert ulp.max() <= 48, gamma
        band = (tt > 1e-3) & (tt < 3.0)
        assert ulp[band].max() <= 12, gamma


# ---- SOR ------------------------------------------------------------------------------
def _laplacian(phi, n, d):
    h = [2.0 * d[a] / n[a] for a in range(3)]
    f = phi.reshape(n)
    L = np.zeros_like(f)
    L[1:-1, 1:-1, 1:-1] = ((f[2:, 1:-1, 1:-1] + f[:-2, 1:-1, 1:-1] - 2 * f[1:-1, 1:-1, 1:-1]) / h[0] ** 2
                           + (f[1:-1, 2:, 1:-1] + f[1:-1, :-2, 1:-1] - 2 * f[1:-1, 1:-1, 1:-1]) / h[1] ** 2
                           + (f[1:-1, 1:-1, 2:] + f[1:-1, 1:-1, :-2] - 2 * f[1:-1, 1:-1, 1:-1]) / h[2] ** 2)
    return L.ravel()


@pytest.mark.parametrize("omega", [1.0, 1.5, 1.8])
def test_sor_manufactured_solution(oracle, omega):
    n, d = (14, 12, 16), (3.5, 3.0, 4.0)
    target = smooth_field(3, n, 10.0, positive=False) + 3.0
    rho = -_laplacian(target, n, d) / (4 * np.pi * 1.44)
    phi0 = target.copy()
    phi0.reshape(n)[1:-1, 1:-1, 1:-1] = 0.0  # faces hold the solution, interior starts at 0
    phi, it, md, warn = oracle.sor(phi0, rho, n, d, omega, 1e-13, 5000)
    assert warn == 0 and md < 1e-13
    assert np.abs(phi - target).max() < 1e-9


def test_sor_uniform_sphere_matches_analytic(oracle):
    """A uniformly charged sphere against 1.44 Q/R (3/2 - r^2/2R^2) inside and
    1.44 Q / r outside (the form of include/physics_formulas.hpp:80-85)."""
    n, d = (48, 48, 48), (12.0, 12.0, 12.0)
    xs, ys, zs = node_pos(n, d)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    R, Q = 6.0, 50.0
    rho = np.where(r <= R, Q / (4.0 / 3.0 * np.pi * R ** 3), 0.0).ravel()
    Qgrid = rho.sum() * 0.5 ** 3
    phi0 = oracle.coulomb_bc(np.zeros(rho.size), n, d, 50)
    phi, it, md, warn = oracle.sor(phi0 * (Qgrid / 50.0), rho, n, d, 1.9, 1e-10, 5000)
    exact = np.where(r <= R, 1.44 * Qgrid / R * (1.5 - 0.5 * r ** 2 / R ** 2), 1.44 * Qgrid / np.maximum(r, 1e-9))
    inner = (r < 9.0).ravel()
    assert np.abs(phi - exact.ravel())[inner].max() < 0.02 * exact.max()


def test_coulomb_faces_closed_form(oracle):
    n, d = (6, 8, 10), (3.0, 4.0, 5.0)
    phi = oracle.coulomb_bc(np.full(480, 7.0), n, d, 20)
    xs, ys, zs = node_pos(n, d)
    f = phi.reshape(n)
    assert np.all(f[1:-1, 1:-1, 1:-1] == 7.0)
    assert f[0, 0, 0] == 1.44 * 20 / math.sqrt(xs[0] ** 2 + ys[0] ** 2 + zs[0] ** 2)
    assert f[-1, 3, 4] == 1.44 * 20 / math.sqrt(xs[-1] ** 2 + ys[3] ** 2 + zs[4] ** 2)


# ---- FDM and gather ---------------------------------------------------------------
def test_fdm_linear_and_quadratic(oracle):
    n, d = (16, 8, 32), (4.0, 2.0, 8.0)  # power-of-two grid: node positions exact
    xs, ys, zs = node_pos(n, d)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    U1 = (0.5 * X - 1.25 * Y + 2.0 * Z + 3.0).ravel()
    U = np.concatenate([U1, -U1])
    Fx, Fy, Fz = oracle.fdm(U, n, d)
    G = U1.size
    assert np.all(Fx[:G] == -0.5) and np.all(Fy[:G] == 1.25) and np.all(Fz[:G] == -2.0)
    assert np.all(Fx[G:] == 0.5) and np.all(Fz[G:] == 2.0)
    U2 = (X ** 2).ravel()
    Fx, _, _ = oracle.fdm(np.concatenate([U2, U2]), n, d)
    inner = Fx[:G].reshape(n)[1:-1]
    assert np.array_equal(inner, -2.0 * X[1:-1])
    # one-sided first order on the faces: -(x1^2 - x0^2)/h
    assert np.all(Fx[:G].reshape(n)[0] == -(xs[1] ** 2 - xs[0] ** 2) / 0.5)


def test_gather_reproduces_linear_fields(oracle):
    rng = np.random.default_rng(8)
    n = 100
    x = rng.uniform(-D3[0], D3[0] - 0.5, n)
    y = rng.uniform(-D3[1], D3[1] - 0.5, n)
    z = rng.uniform(-D3[2], D3[2] - 0.5, n)
    xs, ys, zs = node_pos(N3, D3)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    F = np.concatenate([(2 * X - Y + 0.5 * Z + 1).ravel()] * 2)
    const = np.full_like(F, 3.0)
    f = oracle.gather(x, y, z, n // 2, n - n // 2, [F, const, F], N3, D3)
    assert np.abs(f[0] - (2 * x - y + 0.5 * z + 1)).max() < 1e-12
    assert np.abs(f[1] - 3.0).max() < 1e-13
    # outside the grid -> zero force
    g = oracle.gather(np.array([50.0]), np.array([0.0]), np.array([0.0]), 1, 0, [F, F, F], N3, D3)
    assert g[0][0] == 0.0 and g[1][0] == 0.0


# ---- push, boost, diagnostics --------------------------------------------------------
def test_leapfrog_reversibility(oracle):
    """N steps forward, momenta flipped, N steps more: positions return (the
    half-kick / drift order and the gather sign are pinned by this)."""
    n, d = (16, 16, 16), (4.0, 4.0, 4.0)
    st = particles(4, 60, 60, 2.5)
    init = {k: st[k] for k in ("x", "y", "z", "kx", "ky", "kz")}
    args = dict(A=-356.0, B=303.0, C=32.0, gamma=7.0 / 6.0)
    out, _, _, phi, _ = oracle.simulate(init, 60, 60, n, d, *args.values(), 0.5, 6, 1e30, 0.0, 1, 20, 20,
                                        SIGMA_R, 3, 1.8, 1e-13, 5000)
    back = {k: out[k] for k in ("x", "y", "z")}
    back.update({k: -out[k] for k in ("kx", "ky", "kz")})
    out2, _, _, _, _ = oracle.simulate(back, 60, 60, n, d, *args.values(), 0.5, 6, 1e30, 0.0, 1, 20, 20,
                                       SIGMA_R, 3, 1.8, 1e-13, 5000)
    for k in ("x", "y", "z"):
        assert np.abs(out2[k] - init[k]).max() < 1e-7
    moved = max(np.abs(out[k] - init[k]).max() for k in ("x", "y", "z"))
    assert moved > 0.1


def test_kick_drift_constant_force(oracle):
    k = [np.array([0.1, -0.2]), np.array([0.0, 0.3]), np.array([1.0, 0.5])]
    f = [np.array([2.0, 4.0]), np.array([-1.0, 0.0]), np.array([0.5, 8.0])]
    dt = 0.5
    k1 = oracle.kick(k, f, dt)
    for a in range(3):
        assert np.allclose(k1[a], k[a] + f[a] * dt / (2 * 197.33), rtol=0, atol=1e-16)
    x = [np.zeros(2)] * 3
    x1 = oracle.drift(x, k1, dt)
    for a in range(3):
        assert np.allclose(x1[a], k1[a] * dt * 197.33 / 935.0, rtol=0, atol=1e-16)


def test_boost_momentum_bookkeeping(oracle):
    z, n, ppn = 50, 82, 10
    np_, nn = z * ppn, n * ppn
    kz = np.zeros(np_ + nn)
    kw = oracle.excite(kz, np_, nn, 0.05, z, n, 0)
    assert abs(kw.sum()) < 1e-12  # protons +eta N/A, neutrons -eta Z/A: total momentum kept
    assert np.all(kw[:np_] == 0.05 * n / (z + n)) and np.all(kw[np_:] == -0.05 * z / (z + n))
    ks = oracle.excite(kz, np_, nn, 0.05, z, n, 1)
    assert np.all(ks[:np_] == 0.05) and np.all(ks[np_:] == -0.05)


def test_stats_exact_on_lattice(oracle):
    rng = np.random.default_rng(9)
    np_, nn = 300, 500
    # values on a 2^-10 lattice are exact in the 2^-30 fixed point
    x, y, z = [rng.integers(-3000, 3000, np_ + nn) / 1024.0 for _ in range(3)]
    n, d = (8, 8, 8), (2.0, 2.0, 2.0)
    st = oracle.stats(x, y, z, np_, nn, n, d)
    assert st[0] == math.fsum(z[:np_]) / np_ and st[1] == math.fsum(z[np_:]) / nn
    inb = (np.abs(x) <= 2.0) & (np.abs(y) <= 2.0) & (np.abs(z) <= 2.0)
    r2 = x * x + y * y + z * z
    assert st[2] == math.fsum(r2[:np_][inb[:np_]]) / inb[:np_].sum()
    assert st[3] == math.fsum(r2[np_:][inb[np_:]]) / inb[np_:].sum()


@pytest.mark.parametrize("b", [0.5 ** 2, 0.4 ** 2, 0.25 ** 2, (2 * 12.0 / 96) ** 2, (2 * 7.5 / 24) ** 2,
                               (2 * 5.0 / 16) ** 2, (24.0 / 96) ** 2, 0.3 ** 2, 1.0 / 3.0])
def test_fixed_divisor_division_is_ieee(oracle, b):
    """The SOR kernels divide by dx^2, dy^2, dz^2 through hv_div_fixed
    (q = a y, r = fma(-q, b, a), q + r y): it must give a / b bit for bit."""
    import ctypes
    fn = oracle.L.hvo_div_fixed_mismatches
    fn.restype = ctypes.c_longlong
    fn.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_longlong, ctypes.c_ulonglong]
    for lo, hi in ((0.0, 1.0), (1.0, 200.0), (-50.0, 50.0), (1e-6, 1e-3)):
        assert fn(b, lo, hi, 2_000_000, 12345) == 0
