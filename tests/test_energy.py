"""Total-energy diagnostic: pins of the oracle's hvo_total_energy (no reference
counterpart; the kinetic prefactor is kinetic_energy(),
include/physics_formulas.hpp:45-49, and the potential energy density is the
functional whose derivative is skyrme_potential, :71-77, plus phi).

Pins: the kinetic part in closed form; the functional derivative of the grid
part, by central finite differences of the energy in one cell's density,
equals dV times the potential the loop computes (Skyrme U_p, U_n from
hvo_skyrme_potentials; phi from an SOR solve with zero Dirichlet faces, which
makes phi linear and symmetric in rho_p)."""
import numpy as np
import pytest

from inputs import smooth_field

A, B, C, GAMMA = -356.0, 303.0, 32.0, 7.0 / 6.0
HBARC, MC2 = 197.33, 935.0


def _dv(n, d):
    return float(np.prod([2.0 * dm / nn for nn, dm in zip(n, d)]))


def test_kinetic_energy_closed_form(oracle):
    n, d = (6, 6, 6), (3.0, 3.0, 3.0)
    G = int(np.prod(n))
    N = 1000
    kx = np.full(N, 0.75)
    ky = np.full(N, -0.5)
    kz = np.zeros(N)
    e = oracle.total_energy((kx, ky, kz), np.zeros(2 * G), np.zeros(G), n, d, A, B, C, GAMMA, 10)
    want = HBARC ** 2 / (2 * MC2) * N * (0.75 ** 2 + 0.5 ** 2) / 10
    assert e[1] == pytest.approx(want, rel=1e-12)
    assert e[2] == 0.0 and e[3] == 0.0 and e[0] == e[1]


def test_skyrme_energy_derivative_is_potential(oracle):
    n, d = (10, 9, 8), (4.0, 3.6, 3.2)
    G = int(np.prod(n))
    rho_p = smooth_field(3, n, 0.07)
    rho_n = smooth_field(4, n, 0.09)
    rho = np.concatenate([rho_p, rho_n])
    U = oracle.skyrme(rho, n, d, A, B, C, GAMMA)
    zero_k = (np.zeros(1), np.zeros(1), np.zeros(1))
    phi = np.zeros(G)
    dV = _dv(n, d)
    h = 1e-4
    cells = [q for q in np.argsort(rho_p + rho_n)[-40:]][::8]  # dense cells
    for q in cells:
        for sp in (0, 1):
            idx = q + sp * G
            r1 = rho.copy()
            r1[idx] += h
            r2 = rho.copy()
            r2[idx] -= h
            e1 = oracle.total_energy(zero_k, r1, phi, n, d, A, B, C, GAMMA, 1)[2]
            e2 = oracle.total_energy(zero_k, r2, phi, n, d, A, B, C, GAMMA, 1)[2]
            fd = (e1 - e2) / (2 * h)
            assert fd == pytest.approx(dV * U[idx], rel=2e-5, abs=1e-7), (q, sp)


def test_coulomb_energy_derivative_is_phi(oracle):
    n, d = (10, 10, 10), (4.0, 4.0, 4.0)
    G = int(np.prod(n))
    rho_p = smooth_field(7, n, 0.08)
    zero_k = (np.zeros(1), np.zeros(1), np.zeros(1))
    dV = _dv(n, d)

    def solve(rp):
        phi, it, md, warn = oracle.sor(np.zeros(G), rp, n, d, 1.5, 1e-12, 20000)
        assert warn == 0
        return phi

    def e_coul(rp):
        rho = np.concatenate([rp, np.zeros(G)])
        return oracle.total_energy(zero_k, rho, solve(rp), n, d, A, B, C, GAMMA, 1)[3]

    phi0 = solve(rho_p)
    # E_coul is quadratic in rho_p, so the central difference is exact for any
    # step; a large one keeps the 2^-30 quantum of every cell's term (all
    # cells change with phi) far below the difference
    h = 5e-2
    for q in [int(np.argmax(rho_p)), G // 2 + 3, 5 * 100 + 5 * 10 + 2]:
        r1 = rho_p.copy()
        r1[q] += h
        r2 = rho_p.copy()
        r2[q] -= h
        fd = (e_coul(r1) - e_coul(r2)) / (2 * h)
        assert fd == pytest.approx(dV * phi0[q], rel=1e-6, abs=1e-8), q


def test_total_is_sum_of_parts(oracle):
    n, d = (8, 8, 8), (3.2, 3.2, 3.2)
    G = int(np.prod(n))
    rng = np.random.default_rng(5)
    k = tuple(rng.normal(0, 1, 300) for _ in range(3))
    rho = np.concatenate([smooth_field(1, n, 0.08), smooth_field(2, n, 0.08)])
    phi = smooth_field(9, n, 20.0)
    e = oracle.total_energy(k, rho, phi, n, d, A, B, C, GAMMA, 3)
    assert e[0] == e[1] + e[2] + e[3]
    # order-free: permuting the particles changes no bit
    perm = rng.permutation(300)
    e2 = oracle.total_energy(tuple(a[perm].copy() for a in k), rho, phi, n, d, A, B, C, GAMMA, 3)
    assert np.array_equal(e, e2)
