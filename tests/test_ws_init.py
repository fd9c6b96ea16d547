"""The reference's Woods-Saxon / Fermi-level initial state
(initialize_particles, include/physics.hpp:35-97) restated in
hv_init_woods_saxon and its pieces.

Pins: a sampling round (random balls, energies, counts, Fermi-level move,
checking particles) equals the reference's own templates bit for bit from the
same rand() seed; the particle densities equal the reference's
compute_particle_densities (GPU exp against libm exp: 1e-12); chi squared
equals what the reference prints; the well fit recovers an exact Woods-Saxon
profile and agrees with MINPACK's Levenberg-Marquardt (scipy) on noisy data.
GSL's trust-region fit itself is absent here: where the reference's fit would
stop (1e-4 relative steps) is parity unpinned."""
import ctypes

import numpy as np
import pytest

from hvamd import initial
from hvamd.sim import Skyrme, sigma_from_fwhm
from oracle_lib import REF_PATH, Reference

LIBC = ctypes.CDLL(None)
SK = Skyrme(-356.8, 303.9, 32.0, 7.0 / 6.0)
KEYS = ("x", "y", "z", "kx", "ky", "kz")


def _ref():
    if not Reference.available:
        pytest.skip("oracle/_ref/libhvref.so not built (needs the reference checkout)")
    L = ctypes.CDLL(str(REF_PATH))
    P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    L.ref_ws_sample_round.argtypes = [I, I, I, D, P, P, P, P] + [P] * 6
    L.ref_ws_max_particles.argtypes = [I, I, I]
    L.ref_ws_particle_densities.argtypes = [P, P, P, I, I, D, I, P, P]
    L.ref_ws_chi_squared.argtypes = [P] * 5 + [I, I, D, D, D, D, P, I, P]
    return L


def _wells(z, n):
    r_max = initial.nuclear_radius(z + n)
    return [initial.WoodsSaxon(-50.0, 0.8 * r_max, 0.66), initial.WoodsSaxon(-50.0, 0.8 * r_max, 0.66)]


@pytest.mark.parametrize("z,n,ppn", [(8, 8, 20), (20, 28, 12)])
def test_sample_round_equals_reference(z, n, ppn):
    L = _ref()
    ws = _wells(z, n)
    sigma_k = sigma_from_fwhm(0.346)
    assert initial.max_particles(initial.nuclear_radius(z + n), 1.5, ppn) == L.ref_ws_max_particles(z, n, ppn)
    LIBC.srand(11)
    got, f, d, c = initial.sample_round(z, n, ppn, sigma_k, ws, [-8.0, -12.0])
    LIBC.srand(11)
    tot = (z + n) * ppn
    want = {k: np.zeros(tot) for k in KEYS}
    wf = np.array([-8.0, -12.0])
    wd = np.zeros(2)
    wc = np.zeros(6, dtype=np.int32)
    wsa = np.array([ws[0].V0, ws[0].R12, ws[0].a, ws[1].V0, ws[1].R12, ws[1].a])
    L.ref_ws_sample_round(z, n, ppn, sigma_k, wsa.ctypes.data, wf.ctypes.data, wd.ctypes.data, wc.ctypes.data,
                          *[want[k].ctypes.data for k in KEYS])
    assert np.array_equal(c, wc), (c, wc)
    assert np.array_equal(d, wd) and np.array_equal(f, wf)
    for k in KEYS:
        assert np.array_equal(got[k], want[k]), k
    # every kept particle is followed by its mirror image
    assert np.array_equal(got["x"][1::2], -got["x"][0::2]) and np.array_equal(got["kz"][1::2], -got["kz"][0::2])


def _ws_profile(r, w):
    return w.V0 / (1.0 + np.exp((r - w.R12) / w.a))


def test_fit_recovers_an_exact_woods_saxon_well():
    """With B = C = 0 the Skyrme potential is A (rho_p + rho_n) / rho0, so
    densities chosen as V_ws(r) rho0 / (2 A) make the data an exact well: the
    fit returns it from a distant start (a closed form, no GPU)."""
    rng = np.random.default_rng(3)
    true = [initial.WoodsSaxon(-47.0, 5.2, 0.61), initial.WoodsSaxon(-53.0, 5.6, 0.7)]
    np_, nn = 3000, 4000
    pos = rng.uniform(-9, 9, (np_ + nn, 3))
    r = np.sqrt((pos ** 2).sum(1))
    sk = Skyrme(-356.8, 0.0, 0.0, 7.0 / 6.0)
    v = np.where(np.arange(np_ + nn) < np_, _ws_profile(r, true[0]), _ws_profile(r, true[1]))
    rho = v * 0.16 / (2.0 * sk.A)
    start = [initial.WoodsSaxon(-30.0, 7.0, 1.0), initial.WoodsSaxon(-30.0, 7.0, 1.0)]
    ws, its = initial.fit(pos[:, 0], pos[:, 1], pos[:, 2], rho, rho, np_, nn, sk, start)
    for got, want in zip(ws, true):
        assert abs(got.V0 - want.V0) < 1e-8 and abs(got.R12 - want.R12) < 1e-8 and abs(got.a - want.a) < 1e-8
    assert (its > 0).all()
    chi = initial.chi_squared(pos[:, 0], pos[:, 1], pos[:, 2], rho, rho, np_, nn, sk, ws, 10)
    assert chi.max() < 1e-12


def test_fit_agrees_with_minpack_levenberg_marquardt():
    """Noisy data from the full Skyrme potential: the fitted wells equal
    scipy's MINPACK Levenberg-Marquardt minimiser of the same residual."""
    from scipy.optimize import least_squares
    rng = np.random.default_rng(5)
    np_, nn = 2500, 3500
    pos = rng.normal(0, 3.5, (np_ + nn, 3))
    r = np.sqrt((pos ** 2).sum(1))
    prof = 0.16 / (1.0 + np.exp((r - 5.5) / 0.55))
    rp = 0.45 * prof * (1 + 0.05 * rng.standard_normal(r.size))
    rn = 0.55 * prof * (1 + 0.05 * rng.standard_normal(r.size))
    start = [initial.WoodsSaxon(-50.0, 6.0, 0.66), initial.WoodsSaxon(-50.0, 6.0, 0.66)]
    ws, _ = initial.fit(pos[:, 0], pos[:, 1], pos[:, 2], rp, rn, np_, nn, SK, start)
    for t, (lo, hi) in enumerate(((0, np_), (np_, np_ + nn))):
        tau = -1.0 if t == 0 else 1.0
        rr, p_, n_ = r[lo:hi], rp[lo:hi], rn[lo:hi]
        tt = (p_ + n_) / 0.16
        v = SK.A * tt + SK.B * tt ** SK.gamma + 2 * tau * SK.C * (n_ - p_) / 0.16
        res = least_squares(lambda q: v - q[0] / (1 + np.exp((rr - q[1]) / q[2])), [-50.0, 6.0, 0.66], method="lm",
                            xtol=1e-14, ftol=1e-14, gtol=1e-14)
        got = np.array([ws[t].V0, ws[t].R12, ws[t].a])
        assert np.allclose(got, res.x, rtol=1e-7, atol=0), (got, res.x)


def test_chi_squared_equals_the_reference_print():
    L = _ref()
    rng = np.random.default_rng(9)
    np_, nn = 700, 900
    pos = np.ascontiguousarray(rng.normal(0, 3.0, (3, np_ + nn)))
    r = np.sqrt((pos ** 2).sum(0))
    rp = np.ascontiguousarray(0.07 / (1 + np.exp((r - 5) / 0.6)))
    rn = np.ascontiguousarray(0.09 / (1 + np.exp((r - 5.2) / 0.6)))
    ws = [initial.WoodsSaxon(-48.0, 5.1, 0.62), initial.WoodsSaxon(-52.0, 5.3, 0.66)]
    got = initial.chi_squared(pos[0], pos[1], pos[2], rp, rn, np_, nn, SK, ws, 7)
    wsa = np.array([ws[0].V0, ws[0].R12, ws[0].a, ws[1].V0, ws[1].R12, ws[1].a])
    out = np.zeros(2)
    x, y, z = (np.ascontiguousarray(pos[i]) for i in range(3))
    assert L.ref_ws_chi_squared(x.ctypes.data, y.ctypes.data, z.ctypes.data, rp.ctypes.data, rn.ctypes.data, np_, nn,
                                SK.A, SK.B, SK.C, SK.gamma, wsa.ctypes.data, 7, out.ctypes.data) == 0
    assert np.all(np.abs(got - out) <= 0.005 + 1e-12 * np.abs(out)), (got, out)
    assert got.min() > 1.0  # the figures carry information at the printed precision


@pytest.mark.gpu
def test_particle_densities_equal_the_reference():
    L = _ref()
    LIBC.srand(21)
    z, n, ppn = 20, 28, 30
    st, _, _, _ = initial.sample_round(z, n, ppn, sigma_from_fwhm(0.346), _wells(z, n), [-8.0, -12.0])
    np_, nn = z * ppn, n * ppn
    sigma_r = sigma_from_fwhm(1.444)
    print("particle densities", np_ + nn, flush=True)
    rp, rn = initial.particle_densities(st["x"], st["y"], st["z"], np_, nn, sigma_r, ppn)
    wp, wn = np.zeros(np_ + nn), np.zeros(np_ + nn)
    L.ref_ws_particle_densities(st["x"].ctypes.data, st["y"].ctypes.data, st["z"].ctypes.data, np_, nn, sigma_r, ppn,
                                wp.ctypes.data, wn.ctypes.data)
    assert np.abs(rp - wp).max() <= 1e-12 * np.abs(wp).max()
    assert np.abs(rn - wn).max() <= 1e-12 * np.abs(wn).max()
    assert wp.max() > 0.01 and wn.max() > 0.01  # nuclear densities, fm^-3


@pytest.mark.gpu
def test_woods_saxon_nucleus_settles_and_repeats():
    """The reference's default nucleus (Z = 50, N = 82, its input defaults,
    100 test particles per nucleon, seed 128): the Fermi levels settle within
    the 32 rounds, the wells and the nucleus are physical, chi squared is what
    the pieces give, and the same seed gives the same bits."""
    print("woods-saxon init Z=50 N=82 ppn=100", flush=True)
    a = initial.woods_saxon_nucleus(50, 82, 100)
    b = initial.woods_saxon_nucleus(50, 82, 100)
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    print("rounds", a["iterations"], "fermi", a["fermi"], "wells", a["ws"], "chi2", a["chi2"])
    assert a["converged"] and 1 <= a["iterations"] < 32
    for w in a["ws"]:
        assert -90.0 < w.V0 < -20.0 and 4.0 < w.R12 < 8.5 and 0.2 < w.a < 2.0
    assert -25.0 < a["fermi"][0] < 0.0 and -25.0 < a["fermi"][1] < 0.0
    r = np.sqrt(a["x"] ** 2 + a["y"] ** 2 + a["z"] ** 2)
    rms = np.sqrt((r ** 2).mean())
    assert 3.5 < rms < 6.5  # fm; r_rms = 0.94 A^(1/3) = 4.8 fm for a sharp sphere of 1.2 A^(1/3)
    assert np.isfinite(a["chi2"]).all() and (a["chi2"] > 0).all()
