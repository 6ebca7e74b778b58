"""Batched engine: many independent periodic cases of uneven size in one run.

``Batch`` wraps the plbm_batch_* entry points of libplbm.so.  ``sweep_cases`` builds the
case list of the uneven-size sweep (grid sizes drawn from a fixed set, viscosities
log-uniform, half Taylor-Green vortices and half doubly periodic shear layers) with the
engine's counter-based generator, so the same seed gives the same cases everywhere.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import check, dptr, lib
from .lattice import ShearLayer, TaylorGreen, csqr, noise, padded_ld

SWEEP_SIZES = (32, 48, 64, 96, 128, 192, 256, 384, 512)


@dataclass
class Case:
    n: int
    nu: float
    kind: str          # "taylor_green" or "shear_layer"
    u0: float

    @property
    def omega(self) -> float:
        # set_properties with dt = 1 (src/fvm_bardow.F90:254-257): omega = 1/(3 nu + 1/2)
        return 1.0 / (3.0 * self.nu + 0.5)

    def fields(self):
        """Initial (rho, ux, uy) on the (n, n) grid, Fortran order."""
        n = self.n
        if self.kind == "taylor_green":
            k = 2 * math.pi / n
            p, ux, uy = TaylorGreen(n, n, k, k, self.u0, self.nu).eval(0.0)
            return np.asfortranarray(p / csqr + 1.0), ux, uy
        # tanh profile of width n/80 (k = 80 in units of the period), 5% sinusoidal v
        return ShearLayer(float(n), self.u0, 80.0, 0.05).eval(n, n)


def sweep_cases(ncases: int = 256, seed: int = 4, sizes=SWEEP_SIZES, nu_min=1e-3,
                nu_max=5e-2) -> list[Case]:
    """Case c: n = sizes[floor(U_2c * len(sizes))], nu = exp(log nu_min + U_2c+1 *
    log(nu_max/nu_min)); the first half Taylor-Green (u0 = 0.05), the rest shear layers
    (u0 = 0.04).  U_i are the engine's deviates of `seed`."""
    out = []
    for c in range(ncases):
        n = sizes[min(len(sizes) - 1, int(noise(seed, 2 * c) * len(sizes)))]
        nu = math.exp(math.log(nu_min) + noise(seed, 2 * c + 1) * math.log(nu_max / nu_min))
        tg = c < ncases // 2
        out.append(Case(n, nu, "taylor_green" if tg else "shear_layer", 0.05 if tg else 0.04))
    return out


class Batch:
    """Owner of one plbm_batch."""

    def __init__(self, sizes, storage: str | int = "f32s", model: str | int = "bgk"):
        sizes = [(int(a), int(b)) for a, b in sizes]
        nx = (ctypes.c_int * len(sizes))(*[s[0] for s in sizes])
        ny = (ctypes.c_int * len(sizes))(*[s[1] for s in sizes])
        h = ctypes.c_void_p()
        check(lib().plbm_batch_create(len(sizes), nx, ny, _abi.STORAGE_IDS.get(storage, storage),
                                      _abi.MODEL_IDS.get(model, model), ctypes.byref(h)))
        self.h = h
        self.sizes = sizes

    @classmethod
    def from_cases(cls, cases, storage="f32s", model="bgk", magic=0.25):
        b = cls([(c.n, c.n) for c in cases], storage, model)
        b.set_params([c.omega for c in cases], [magic] * len(cases))
        for i, c in enumerate(cases):
            b.equilibrium_init(i, *c.fields())
        return b

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            check(lib().plbm_batch_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        n, cells, nb = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_size_t()
        check(lib().plbm_batch_info(self.h, ctypes.byref(n), ctypes.byref(cells), ctypes.byref(nb)))
        return {"ncases": n.value, "cells": cells.value, "arena_bytes": nb.value}

    def set_params(self, omega, trt_magic=None):
        om = np.ascontiguousarray(omega, dtype=np.float64)
        mg = None if trt_magic is None else np.ascontiguousarray(trt_magic, dtype=np.float64)
        check(lib().plbm_batch_set_params(self.h, dptr(om), None if mg is None else dptr(mg)))

    def equilibrium_init(self, c, rho, ux, uy):
        a = [np.asfortranarray(v, dtype=np.float64) for v in (rho, ux, uy)]
        check(lib().plbm_batch_equilibrium_init(self.h, c, *(dptr(v) for v in a)))

    def upload_f(self, c, f):
        f = np.asfortranarray(f, dtype=np.float64)
        check(lib().plbm_batch_upload_f(self.h, c, dptr(f), f.shape[0]))

    def download_f(self, c, ld=None):
        nx, ny = self.sizes[c]
        ld = padded_ld(ny) if ld is None else ld
        f = np.zeros((ld, nx, 9), order="F")
        check(lib().plbm_batch_download_f(self.h, c, dptr(f), ld))
        return f

    def macros(self, c):
        nx, ny = self.sizes[c]
        r, u, v = (np.zeros((ny, nx), order="F") for _ in range(3))
        check(lib().plbm_batch_macros(self.h, c, dptr(r), dptr(u), dptr(v)))
        return r, u, v

    def reduce(self):
        out = np.zeros((len(self.sizes), _abi.NRED))
        check(lib().plbm_batch_reduce(self.h, dptr(out.reshape(-1))))
        return out

    def run(self, nsteps, graph=True):
        fn = lib().plbm_batch_run if graph else lib().plbm_batch_run_direct
        check(fn(self.h, nsteps))

    def sync(self):
        check(lib().plbm_batch_sync(self.h))

    def set_launch(self, ty=128, tx=1, resident_cells=0, resident_block=256):
        check(lib().plbm_batch_set_launch(self.h, ty, tx, resident_cells, resident_block))

    def get_launch(self):
        v = [ctypes.c_int() for _ in range(4)]
        check(lib().plbm_batch_get_launch(self.h, *(ctypes.byref(x) for x in v)))
        return tuple(x.value for x in v)

    def set_steps_per_pass(self, steps: int):
        """Steps per launch of the tiled cases (1: one step per launch; 2..8: temporally
        blocked passes)."""
        check(lib().plbm_batch_set_steps_per_pass(self.h, steps))

    def get_steps_per_pass(self) -> int:
        v = ctypes.c_int()
        check(lib().plbm_batch_get_steps_per_pass(self.h, ctypes.byref(v)))
        return v.value

    def set_pass_columns(self, columns: int):
        """Columns each wavefront of a pass walks (segments of a case); results do not
        depend on it."""
        check(lib().plbm_batch_set_pass_columns(self.h, columns))

    def get_pass_columns(self) -> int:
        v = ctypes.c_int()
        check(lib().plbm_batch_get_pass_columns(self.h, ctypes.byref(v)))
        return v.value

    def autotune(self, steps=10) -> float:
        ms = ctypes.c_double()
        check(lib().plbm_batch_autotune(self.h, steps, ctypes.byref(ms)))
        return ms.value
