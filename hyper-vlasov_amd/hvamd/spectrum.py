"""Dipole strength function and its centroid from a D(t) series.

Restates the reference's post-processing (scripts/programs.py:82-124):
D(t) = N Z / A (<z>_p - <z>_n) (:99), linear detrend (:100), cos^2 taper
from the first sample to t_f (:104), S(omega) = Im(conj(FFT(D)) dt) / (pi eta
hbar c) (:109-113), kept for omega > 0 and reported against E = hbar omega
(:115-117).  The centroid E_c = sum E S / sum S over an energy window has no
counterpart in the reference; it is the scalar the precision comparison uses.
energy_spectrum_file() is the reference's whole post-processing step: it reads
the stats.txt a run wrote (hv_vlasov, or the reference binary), keeps the
lines from t_exc on and writes energy_spectrum.txt.
"""
from __future__ import annotations

import numpy as np
from scipy import signal
from scipy.fft import fft, fftfreq

HBARC = 197.33


def dipole_moment(com_p: np.ndarray, com_n: np.ndarray, z: int, n: int) -> np.ndarray:
    """D = N Z / A (<z>_p - <z>_n)  (scripts/programs.py:99)."""
    return (n * z) / (n + z) * (np.asarray(com_p) - np.asarray(com_n))


def strength_function(time: np.ndarray, dipole: np.ndarray, t_f: float, eta: float):
    """(E [MeV], S(E)) from D(t) sampled at uniform `time` [fm/c] up to t_f.

    time[0] is the first sample after the excitation (the reference slices the
    stats from t_exc)."""
    time = np.asarray(time, dtype=np.float64)
    return _strength(time, dipole, t_f, eta, float(time[1] - time[0]))


def _strength(time: np.ndarray, dipole: np.ndarray, t_f: float, eta: float, dt: float):
    d = signal.detrend(np.asarray(dipole, dtype=np.float64))
    t0 = time[0]
    d = d * np.cos(np.pi * (time - t0) / (2.0 * (t_f - t0))) ** 2
    freq = fftfreq(d.size, d=dt)
    omega = 2.0 * np.pi * freq
    raw = np.conjugate(fft(d)) * dt
    if eta > 1e-5:
        s = np.imag(raw) / (np.pi * eta * HBARC)
    else:
        s = np.imag(raw) / (np.pi * HBARC)
    pos = omega > 0
    return omega[pos] * HBARC, s[pos]


def centroid(energy: np.ndarray, strength: np.ndarray, e_lo: float = 5.0, e_hi: float = 40.0) -> float:
    """E_c = sum E S / sum S over e_lo <= E <= e_hi (uniform energy bins)."""
    m = (energy >= e_lo) & (energy <= e_hi)
    s = strength[m]
    return float((energy[m] * s).sum() / s.sum())


def dipole_response(stats: np.ndarray, dt: float, z: int, n: int, eta: float, e_lo: float = 5.0,
                    e_hi: float = 40.0):
    """From per-step stats rows (com_p, com_n, msr_p, msr_n) of a run boosted
    at t = 0: returns (time, D, E, S, centroid)."""
    stats = np.asarray(stats)
    time = dt * np.arange(stats.shape[0])
    d = dipole_moment(stats[:, 0], stats[:, 1], z, n)
    t_f = dt * stats.shape[0]
    e, s = strength_function(time, d, t_f, eta)
    return time, d, e, s, centroid(e, s, e_lo, e_hi)


def stats_start_index(t_exc: float, t_f: float, steps: int, substeps: int) -> int:
    """First stats.txt line used for the spectrum: int(t_exc / t_f * steps /
    substeps) (scripts/programs.py:95), the line of the excitation step."""
    return int(t_exc / t_f * steps / substeps)


def energy_spectrum_file(stats_path, z: int, n: int, t_f: float, steps: int, t_exc: float, substeps: int,
                         eta_exc: float, out_path=None):
    """(E, S) from a stats.txt written every `substeps` steps (columns: time,
    sqrt<r^2>_p, sqrt<r^2>_n, <z>_p, <z>_n; src/hyper_vlasov.cpp:69-70), as
    scripts/programs.py:82-124 computes it: the lines from the excitation on
    (:95-98), D = N Z / A (<z>_p - <z>_n) (:99), sample spacing t_f / steps x
    substeps (:103); written to out_path (energy_spectrum.txt) as two columns
    in %e format (:119-120) when given."""
    data = np.loadtxt(stats_path, dtype=np.float64, ndmin=2)
    start = stats_start_index(t_exc, t_f, steps, substeps)
    time = data[start:, 0]
    d = dipole_moment(data[start:, 3], data[start:, 4], z, n)
    dt = t_f / steps * substeps
    e, s = _strength(time, d, t_f, eta_exc, dt)
    if out_path is not None:
        np.savetxt(out_path, np.column_stack((e, s)), fmt="%e")
    return e, s
