"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and, where it was
built, of the reference harness (oracle/_ref/libhvref.so).  Test
infrastructure: used only by tests/ as the checker."""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_PATH = ROOT / "oracle" / "liboracle.so"
REF_PATH = ROOT / "oracle" / "_ref" / "libhvref.so"

_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int


class OWorld(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("d_max", ctypes.c_double * 3)]


class OSkyrme(ctypes.Structure):
    _fields_ = [("A", _D), ("B", _D), ("C", _D), ("gamma", _D)]


class OParams(ctypes.Structure):
    _fields_ = [("dt", _D), ("steps", _I), ("t_exc", _D), ("eta", _D), ("boost_mode", _I), ("z", _I),
                ("n", _I), ("sigma_r", _D), ("part_per_nucleon", _I), ("omega", _D), ("sor_tol", _D),
                ("sor_max_it", _I)]


def p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def oworld(n, d_max):
    return OWorld((ctypes.c_int * 3)(*n), (ctypes.c_double * 3)(*d_max))


class Oracle:
    def __init__(self):
        if not ORACLE_PATH.exists():
            raise RuntimeError(f"{ORACLE_PATH} not built (make -C oracle)")
        L = ctypes.CDLL(str(ORACLE_PATH))
        L.hvo_deposit.restype = _I
        L.hvo_sor.restype = _I
        L.hvo_simulate.restype = _I
        L.hvo_center_of_mass.restype = _D
        L.hvo_mean_squared_radius.restype = _D
        for name in ("hvo_smear", "hvo_smear_literal"):
            getattr(L, name).argtypes = [_P, _P, _P, _D, _I]
        L.hvo_kick.argtypes = [_P] * 6 + [_I, _D]
        L.hvo_drift.argtypes = [_P] * 6 + [_I, _D]
        L.hvo_excite.argtypes = [_P, _I, _I, _D, _I, _I, _I]
        L.hvo_sor.argtypes = [_P, _P, _P, _D, _D, _I, _P, _P]
        L.hvo_center_of_mass.argtypes = [_P, _I, _I, _I]
        L.hvo_mean_squared_radius.argtypes = [_P, _P, _P, _I, _I, _P, _I]
        L.hvo_total_energy.argtypes = [_P, _P, _P, _I, _P, _P, _P, _P, _I, _P]
        self.L = L

    # each returns new arrays; inputs are float64 numpy arrays
    def deposit(self, x, y, z, np_, nn, n, d_max):
        G = int(np.prod(n))
        acc = np.zeros(2 * G, dtype=np.int64)
        counts = np.zeros(2 * G)
        w = oworld(n, d_max)
        skipped = self.L.hvo_deposit(p(x), p(y), p(z), np_, nn, ctypes.byref(w), p(acc), p(counts))
        return counts, acc, skipped

    def smear(self, counts, n, d_max, sigma_r, ppn, literal=False):
        rho = np.zeros_like(counts)
        w = oworld(n, d_max)
        fn = self.L.hvo_smear_literal if literal else self.L.hvo_smear
        fn(p(counts), p(rho), ctypes.byref(w), sigma_r, ppn)
        return rho

    def skyrme(self, rho, n, d_max, A, B, C, gamma):
        U = np.zeros_like(rho)
        w = oworld(n, d_max)
        s = OSkyrme(A, B, C, gamma)
        self.L.hvo_skyrme_potentials(p(rho), p(U), ctypes.byref(w), ctypes.byref(s))
        return U

    def coulomb_bc(self, phi, n, d_max, z):
        phi = phi.copy()
        w = oworld(n, d_max)
        self.L.hvo_coulomb_bc(p(phi), ctypes.byref(w), z)
        return phi

    def sor(self, phi, rho_p, n, d_max, omega, tol, max_it):
        phi = phi.copy()
        w = oworld(n, d_max)
        it = ctypes.c_int()
        md = ctypes.c_double()
        warn = self.L.hvo_sor(p(phi), p(rho_p), ctypes.byref(w), omega, tol, max_it, ctypes.byref(it),
                              ctypes.byref(md))
        return phi, it.value, md.value, warn

    def merge(self, U, phi, n, d_max):
        U = U.copy()
        w = oworld(n, d_max)
        self.L.hvo_merge(p(U), p(phi), ctypes.byref(w))
        return U

    def fdm(self, U, n, d_max):
        F = [np.zeros_like(U) for _ in range(3)]
        w = oworld(n, d_max)
        self.L.hvo_fdm(p(U), p(F[0]), p(F[1]), p(F[2]), ctypes.byref(w))
        return F

    def gather(self, x, y, z, np_, nn, F, n, d_max):
        f = [np.zeros(np_ + nn) for _ in range(3)]
        w = oworld(n, d_max)
        self.L.hvo_gather(p(x), p(y), p(z), np_, nn, p(F[0]), p(F[1]), p(F[2]), ctypes.byref(w), p(f[0]),
                          p(f[1]), p(f[2]))
        return f

    def kick(self, k, f, dt):
        k = [a.copy() for a in k]
        self.L.hvo_kick(p(k[0]), p(k[1]), p(k[2]), p(f[0]), p(f[1]), p(f[2]), len(k[0]), dt)
        return k

    def drift(self, x, k, dt):
        x = [a.copy() for a in x]
        self.L.hvo_drift(p(x[0]), p(x[1]), p(x[2]), p(k[0]), p(k[1]), p(k[2]), len(x[0]), dt)
        return x

    def excite(self, kz, np_, nn, eta, z, n, mode):
        kz = kz.copy()
        self.L.hvo_excite(p(kz), np_, nn, eta, z, n, mode)
        return kz

    def total_energy(self, k, rho, phi, n, d_max, A, B, C, gamma, ppn):
        """[E, E_kin, E_sky, E_coul] (hvo_total_energy)."""
        out = np.zeros(4)
        w = oworld(n, d_max)
        s = OSkyrme(A, B, C, gamma)
        self.L.hvo_total_energy(p(k[0]), p(k[1]), p(k[2]), len(k[0]), p(rho), p(phi), ctypes.byref(w),
                                ctypes.byref(s), ppn, p(out))
        return out

    def stats(self, x, y, z, np_, nn, n, d_max):
        st = np.zeros(4)
        w = oworld(n, d_max)
        self.L.hvo_stats(p(x), p(y), p(z), np_, nn, ctypes.byref(w), p(st))
        return st

    def simulate(self, state, np_, nn, n, d_max, A, B, C, gamma, dt, steps, t_exc, eta, boost_mode, z, nnuc,
                 sigma_r, ppn, omega, tol, max_it, phi=None):
        """state: dict of x..fz (copied).  Returns (state, stats[steps,4], iters[steps+1], phi, warn)."""
        st = {k: np.ascontiguousarray(v, dtype=np.float64).copy() for k, v in state.items()}
        for k in ("fx", "fy", "fz"):
            if k not in st:
                st[k] = np.zeros(np_ + nn)
        w = oworld(n, d_max)
        s = OSkyrme(A, B, C, gamma)
        prm = OParams(dt, steps, t_exc, eta, boost_mode, z, nnuc, sigma_r, ppn, omega, tol, max_it)
        G = int(np.prod(n))
        phi_io = np.zeros(G) if phi is None else phi.copy()
        stats = np.zeros(4 * steps)
        its = np.zeros(steps + 1, dtype=np.int32)
        warn = self.L.hvo_simulate(*[p(st[k]) for k in ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz")],
                                   np_, nn, ctypes.byref(w), ctypes.byref(s), ctypes.byref(prm),
                                   None if phi is None else p(phi_io), p(stats), p(its))
        return st, stats.reshape(steps, 4), its, (phi_io if phi is not None else None), warn


class Reference:
    """The reference's own kernels (compiled from its headers by oracle/Makefile)."""

    available = REF_PATH.exists()

    def __init__(self):
        if not REF_PATH.exists():
            raise RuntimeError("oracle/_ref/libhvref.so not built")
        L = ctypes.CDLL(str(REF_PATH))
        L.ref_center_of_mass.restype = _D
        L.ref_mean_squared_radius.restype = _D
        L.ref_center_of_mass.argtypes = [_P, _I, _I, _I]
        L.ref_mean_squared_radius.argtypes = [_P, _P, _P, _I, _I, _P, _P, _I]
        L.ref_smear.argtypes = [_P, _P, _P, _P, _D, _I]
        L.ref_skyrme.argtypes = [_P, _P, _P, _P, _D, _D, _D, _D]
        L.ref_kick.argtypes = [_P] * 6 + [_I, _I, _D]
        L.ref_drift.argtypes = [_P] * 6 + [_I, _I, _D]
        L.ref_excite.argtypes = [_P, _I, _I, _D, _I, _I]
        L.ref_coulomb_bc.argtypes = [_P, _P, _P, _I]
        L.ref_write_vtk.restype = _I
        L.ref_write_vtk.argtypes = [ctypes.c_char_p, _P, _P] + [_P] * 5 + [_I]
        L.ref_simulate.argtypes = ([_P] * 9 + [_I, _I, _P, _P] + [_D] * 4 + [_D, _I, _D, _D, _I, _I, _D, _I]
                                   + [_P, _P])
        self.L = L

    @staticmethod
    def _nd(n, d_max):
        return (ctypes.c_int * 3)(*n), (ctypes.c_double * 3)(*d_max)

    def write_vtk(self, path, n, d_max, fx, fy, fz, density, potentials, single=False):
        nn_, dd = self._nd(n, d_max)
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (fx, fy, fz, density, potentials)]
        rc = self.L.ref_write_vtk(str(path).encode(), nn_, dd, *[p(v) for v in a], 1 if single else 0)
        if rc:
            raise OSError(f"cannot write {path}")

    def deposit(self, x, y, z, np_, nn, n, d_max):
        counts = np.zeros(2 * int(np.prod(n)))
        nn_, dd = self._nd(n, d_max)
        self.L.ref_deposit(p(x), p(y), p(z), np_, nn, nn_, dd, p(counts))
        return counts

    def smear(self, counts, n, d_max, sigma_r, ppn):
        d = counts.copy()
        t = np.zeros_like(d)
        nn_, dd = self._nd(n, d_max)
        self.L.ref_smear(p(d), p(t), nn_, dd, sigma_r, ppn)
        return d

    def skyrme(self, rho, n, d_max, A, B, C, gamma):
        U = np.zeros_like(rho)
        nn_, dd = self._nd(n, d_max)
        self.L.ref_skyrme(p(rho), p(U), nn_, dd, A, B, C, gamma)
        return U

    def coulomb_bc(self, phi, n, d_max, z):
        phi = phi.copy()
        nn_, dd = self._nd(n, d_max)
        self.L.ref_coulomb_bc(p(phi), nn_, dd, z)
        return phi

    def sor(self, phi, rho_p, n, d_max):
        phi = phi.copy()
        nn_, dd = self._nd(n, d_max)
        self.L.ref_sor(p(phi), p(rho_p), nn_, dd)
        return phi

    def merge(self, U, phi, n, d_max):
        U = U.copy()
        nn_, dd = self._nd(n, d_max)
        self.L.ref_merge(p(U), p(phi), nn_, dd)
        return U

    def fdm(self, U, n, d_max):
        F = [np.zeros_like(U) for _ in range(3)]
        nn_, dd = self._nd(n, d_max)
        self.L.ref_fdm(p(U), p(F[0]), p(F[1]), p(F[2]), nn_, dd)
        return F

    def gather(self, x, y, z, np_, nn, F, n, d_max):
        f = [np.zeros(np_ + nn) for _ in range(3)]
        nn_, dd = self._nd(n, d_max)
        self.L.ref_gather(p(x), p(y), p(z), np_, nn, p(F[0]), p(F[1]), p(F[2]), nn_, dd, p(f[0]), p(f[1]),
                          p(f[2]))
        return f

    def kick(self, k, f, dt):
        k = [a.copy() for a in k]
        f = [a.copy() for a in f]
        self.L.ref_kick(p(k[0]), p(k[1]), p(k[2]), p(f[0]), p(f[1]), p(f[2]), len(k[0]), 0, dt)
        return k

    def drift(self, x, k, dt):
        x = [a.copy() for a in x]
        k = [a.copy() for a in k]
        self.L.ref_drift(p(x[0]), p(x[1]), p(x[2]), p(k[0]), p(k[1]), p(k[2]), len(x[0]), 0, dt)
        return x

    def excite(self, kz, np_, nn, eta, z, n):
        kz = kz.copy()
        self.L.ref_excite(p(kz), np_, nn, eta, z, n)
        return kz

    def center_of_mass(self, z, np_, nn, type_):
        zc = np.ascontiguousarray(z, dtype=np.float64).copy()
        return self.L.ref_center_of_mass(p(zc), np_, nn, type_)

    def mean_squared_radius(self, x, y, z, np_, nn, n, d_max, type_):
        nn_, dd = self._nd(n, d_max)
        xc, yc, zc = [np.ascontiguousarray(a, dtype=np.float64).copy() for a in (x, y, z)]
        return self.L.ref_mean_squared_radius(p(xc), p(yc), p(zc), np_, nn, nn_, dd, type_)

    def simulate(self, state, np_, nn, n, d_max, A, B, C, gamma, t_f, steps, t_exc, eta, z, nnuc, sigma_r, ppn):
        st = {k: np.ascontiguousarray(v, dtype=np.float64).copy() for k, v in state.items()}
        for k in ("fx", "fy", "fz"):
            if k not in st:
                st[k] = np.zeros(np_ + nn)
        nn_, dd = self._nd(n, d_max)
        stats = np.zeros(4 * steps)
        phi = np.zeros(int(np.prod(n)))
        self.L.ref_simulate(*[p(st[k]) for k in ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz")], np_, nn, nn_,
                            dd, A, B, C, gamma, t_f, steps, t_exc, eta, z, nnuc, sigma_r, ppn, p(stats), p(phi))
        return st, stats.reshape(steps, 4), phi
