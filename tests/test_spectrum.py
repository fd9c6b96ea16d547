"""Strength function and centroid (restating scripts/programs.py:82-124),
pinned by closed forms.  No GPU."""
import numpy as np

from hvamd import spectrum


def damped(E0, gamma, t, D0=1.0):
    hbarc = spectrum.HBARC
    return D0 * np.exp(-gamma * t / (2 * hbarc)) * np.sin(E0 * t / hbarc)


def test_dipole_moment_formula():
    d = spectrum.dipole_moment(np.array([0.3]), np.array([-0.2]), 82, 126)
    assert abs(d[0] - 82 * 126 / 208 * 0.5) < 1e-15


def test_peak_at_resonance_energy():
    t = 0.5 * np.arange(8000)
    for E0 in (12.0, 15.5, 21.0):
        e, s = spectrum.strength_function(t, damped(E0, 2.0, t), t[-1] + 0.5, 0.076)
        de = e[1] - e[0]
        assert abs(e[np.argmax(s)] - E0) <= de
        # a pure sine driven by a positive kick has positive strength at the peak
        assert s.max() > 0


def test_centroid_of_symmetric_line():
    t = 0.5 * np.arange(8000)
    e, s = spectrum.strength_function(t, damped(15.0, 3.0, t), t[-1] + 0.5, 0.076)
    c = spectrum.centroid(e, s, 7.0, 23.0)
    assert abs(c - 15.0) < 0.1


def test_two_lines_centroid_is_weighted_mean():
    t = 0.5 * np.arange(16000)
    d = damped(12.0, 1.0, t, 1.0) + damped(20.0, 1.0, t, 1.0)
    e, s = spectrum.strength_function(t, d, t[-1] + 0.5, 0.05)
    c = spectrum.centroid(e, s, 5.0, 30.0)
    # each line's integrated strength is proportional to E0 x amplitude^2 ... for
    # a sine of fixed amplitude the FFT weight is the same, so the centroid is
    # the plain mean of the two energies
    assert abs(c - 16.0) < 0.15


def test_eta_scales_strength():
    t = 0.5 * np.arange(2000)
    d = damped(15.0, 2.0, t)
    _, s1 = spectrum.strength_function(t, d, t[-1] + 0.5, 0.05)
    _, s2 = spectrum.strength_function(t, d, t[-1] + 0.5, 0.10)
    assert np.allclose(s1, 2 * s2)


def test_stats_start_index_is_the_excitation_line():
    # t_exc = 100 fm/c of t_f = 800 over 800 steps written every 4: line 25
    assert spectrum.stats_start_index(100.0, 800.0, 800, 4) == 25
    assert spectrum.stats_start_index(0.0, 800.0, 800, 4) == 0
    # truncation, not rounding (int() of 24.999...)
    assert spectrum.stats_start_index(99.99, 800.0, 800, 4) == 24


def test_energy_spectrum_file_from_a_stats_file(tmp_path):
    """A stats.txt whose lines before t_exc hold a different signal: the
    spectrum comes from the lines from t_exc on only (a damped line at E0 is
    found at E0), and energy_spectrum.txt holds the same two columns."""
    z, n, t_f, steps, substeps, t_exc = 50, 82, 800.0, 3200, 2, 100.0
    dt = t_f / steps * substeps
    t = dt * np.arange(steps // substeps)
    start = spectrum.stats_start_index(t_exc, t_f, steps, substeps)
    fac = n * z / (n + z)
    dip = np.where(t >= t[start], damped(14.0, 2.0, t - t[start], 0.05), 0.3 * np.sin(t * 40.0 / spectrum.HBARC))
    cm_p, cm_n = dip / (2 * fac), -dip / (2 * fac)
    rows = np.column_stack((t, np.full_like(t, 5.0), np.full_like(t, 5.1), cm_p, cm_n))
    path = tmp_path / "stats.txt"
    with open(path, "w") as f:
        for r in rows:
            f.write("%e %e %e %e %e\n" % tuple(r))
    out = tmp_path / "energy_spectrum.txt"
    e, s = spectrum.energy_spectrum_file(path, z, n, t_f, steps, t_exc, substeps, 0.05, out)
    assert abs(e[np.argmax(s)] - 14.0) <= e[1] - e[0]
    back = np.loadtxt(out)
    assert back.shape == (e.size, 2)
    assert np.allclose(back[:, 0], e, rtol=1e-6) and np.allclose(back[:, 1], s, rtol=1e-6, atol=1e-12 * abs(s).max())
    # the spacing of the energy grid is 2 pi hbar c / (samples x dt)
    m = steps // substeps - start
    assert abs((e[1] - e[0]) - 2 * np.pi * spectrum.HBARC / (m * dt)) < 1e-9
