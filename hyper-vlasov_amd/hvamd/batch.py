"""Ensembles of nuclei stepped together (hv_batch_* in include/hv_abi.h).

Each instance has its own grid, particle counts, Skyrme set and loop
parameters; one launch per loop stage advances all of them.  Instance i of a
batch gives the bits of a solo Simulation of it with the same options and the
per-sweep SOR form (HV_OPT_SOR_RESIDENT_F32 = 0 in the mixed form)."""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import HV_PREC_F64, HvError, HvParams, HvSkyrme, HvWorld
from .sim import Params, Simulation, Skyrme, World


class Batch:
    def __init__(self, worlds: Sequence[World], n_protons: Sequence[int], n_neutrons: Sequence[int],
                 precision: int = HV_PREC_F64):
        self.lib = _lib.load()
        self.n = len(worlds)
        if not (len(n_protons) == len(n_neutrons) == self.n) or self.n == 0:
            raise ValueError("one world and one particle count pair per instance")
        self.worlds = list(worlds)
        self.np = [int(v) for v in n_protons]
        self.nn = [int(v) for v in n_neutrons]
        self.precision = precision
        wa = (HvWorld * self.n)(*[w.c() for w in worlds])
        pa = (ctypes.c_int * self.n)(*self.np)
        na = (ctypes.c_int * self.n)(*self.nn)
        h = ctypes.c_void_p()
        rc = self.lib.hv_batch_create(self.n, wa, pa, na, precision, ctypes.byref(h))
        if rc < 0:
            raise HvError("hv_batch_create", rc, "bad instance or device allocation failure")
        self.b = h
        self.last_warnings = 0

    def close(self):
        if getattr(self, "b", None):
            self.lib.hv_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _rc(self, fn: str, rc: int) -> int:
        if rc < 0:
            raise HvError(fn, rc, self.lib.hv_batch_last_error(self.b).decode())
        self.last_warnings = rc
        return rc

    def instance(self, i: int) -> Simulation:
        """Non-owning Simulation view of instance i (upload/download, energy...)."""
        ctx = self.lib.hv_batch_instance(self.b, i)
        if not ctx:
            raise IndexError(i)
        return Simulation.view(ctx, self.worlds[i], self.np[i], self.nn[i], self.precision)

    def set_option(self, option: int, value: int) -> int:
        return self._rc("hv_batch_set_option", self.lib.hv_batch_set_option(self.b, option, value))

    def get_option(self, option: int) -> int:
        v = ctypes.c_int()
        self._rc("hv_batch_get_option", self.lib.hv_batch_get_option(self.b, option, ctypes.byref(v)))
        return v.value

    def _arrays(self, skyrmes: Sequence[Skyrme], params: Sequence[Params]):
        if len(skyrmes) != self.n or len(params) != self.n:
            raise ValueError("one Skyrme set and one Params per instance")
        sk = (HvSkyrme * self.n)(*[s.c() for s in skyrmes])
        pr = (HvParams * self.n)(*[p.c() for p in params])
        return sk, pr

    def run_init(self, skyrmes, params) -> np.ndarray:
        sk, pr = self._arrays(skyrmes, params)
        its = np.zeros(self.n, dtype=np.int32)
        self._rc("hv_batch_run_init", self.lib.hv_batch_run_init(self.b, sk, pr, its.ctypes.data_as(_lib._IP)))
        return its

    def run_steps(self, skyrmes, params, first_step: int, nsteps: int):
        """(stats[n, nsteps, 4], sor_iters[n, nsteps])."""
        sk, pr = self._arrays(skyrmes, params)
        stats = np.zeros((self.n, nsteps, 4))
        its = np.zeros((self.n, nsteps), dtype=np.int32)
        self._rc("hv_batch_run_steps", self.lib.hv_batch_run_steps(self.b, sk, pr, first_step, nsteps,
                                                                   stats.ctypes.data_as(ctypes.c_void_p),
                                                                   its.ctypes.data_as(ctypes.c_void_p)))
        return stats, its

    def profile_enable(self, on: bool = True):
        self._rc("hv_batch_profile_enable", self.lib.hv_batch_profile_enable(self.b, 1 if on else 0))

    def sor_counts(self) -> dict:
        """Sweep launches and interior-node iterations (summed over instances)
        of the FP32 correction and of FP64 since profile_enable."""
        v = [ctypes.c_longlong() for _ in range(4)]
        self._rc("hv_batch_sor_counts", self.lib.hv_batch_sor_counts(self.b, *[ctypes.byref(x) for x in v]))
        return {"sweeps_f32": v[0].value, "sweeps_f64": v[1].value, "node_iters_f32": v[2].value,
                "node_iters_f64": v[3].value}

    def graph_info(self) -> dict:
        """Steps replayed from a captured step graph, and graphs captured."""
        v = [ctypes.c_longlong() for _ in range(2)]
        self._rc("hv_batch_graph_info", self.lib.hv_batch_graph_info(self.b, *[ctypes.byref(x) for x in v]))
        return {"graph_steps": v[0].value, "captures": v[1].value}

    def profile_read(self) -> dict:
        out = {}
        for k, name in enumerate(_lib.KCLASS_NAMES):
            ms = ctypes.c_double()
            n = ctypes.c_longlong()
            self._rc("hv_batch_profile_read", self.lib.hv_batch_profile_read(self.b, k, ctypes.byref(ms),
                                                                             ctypes.byref(n)))
            out[name] = (ms.value, n.value)
        return out
