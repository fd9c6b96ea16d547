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
