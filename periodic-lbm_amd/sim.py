"""Plugin driver for the reference's sim/ protocol (c_<name>_{init,step,vars,free,norm}).

Behaviour of ivan-pi/periodic-lbm sim/cases.py:13-233,388-425 (SimPlugin, TGVStudy, run,
perfsweep), written for this engine: ``SimPlugin(path)`` loads any library that exports
the protocol — this package's libplbm.so exports it as c_plbm_* (include/plbm_sim.h) —
and arrays are Fortran-ordered (nx, ny) / (nx, ny, 2) with x fastest.
"""
from __future__ import annotations

import ctypes
import time
from pathlib import Path

import numpy as np

from ._abi import LIBPATH


class SimPlugin:
    def __init__(self, path: str | Path = LIBPATH):
        self.ptr = None
        name = Path(path).stem
        name = name[3:] if name.startswith("lib") else name
        self.lib = ctypes.CDLL(str(path))
        get = lambda s: self.lib["c_" + name + "_" + s]
        self.c_init, self.c_vars, self.c_step = get("init"), get("vars"), get("step")
        self.c_free, self.c_norm = get("free"), get("norm")
        f2 = np.ctypeslib.ndpointer(np.float64, ndim=2, flags="F_CONTIGUOUS")
        f3 = np.ctypeslib.ndpointer(np.float64, ndim=3, flags="F_CONTIGUOUS")
        self.c_norm.restype = ctypes.c_double
        self.c_norm.argtypes = (ctypes.c_int, ctypes.c_int, f2, f2)
        self.c_init.restype = ctypes.c_void_p
        self.c_init.argtypes = (ctypes.c_int, ctypes.c_int, ctypes.c_double, f2, f3,
                                ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)
        self.c_step.restype = None
        self.c_step.argtypes = (ctypes.c_void_p, ctypes.c_double)
        w2 = np.ctypeslib.ndpointer(np.float64, ndim=2, flags="F_CONTIGUOUS,WRITEABLE")
        w3 = np.ctypeslib.ndpointer(np.float64, ndim=3, flags="F_CONTIGUOUS,WRITEABLE")
        self.c_vars.restype = None
        self.c_vars.argtypes = (ctypes.c_void_p, w2, w3)
        self.c_free.restype = None
        self.c_free.argtypes = (ctypes.c_void_p,)
        self._params = None

    def my_norm(self, u, ua):
        u = np.require(u, np.float64, ["F_CONTIGUOUS"])
        ua = np.require(ua, np.float64, ["F_CONTIGUOUS"])
        nx, ny = u.shape
        return self.c_norm(nx, ny, u, ua)

    def init(self, grid_size, dt, p, u, sigma=None, storage: int | None = None):
        """p: pressure (nx, ny); u: (nx, ny, 2).  `storage` selects the engine's
        population storage (passed through the protocol's params pointer)."""
        if sigma is not None:
            raise NotImplementedError("non-equilibrium initialisation")
        self.grid_size = tuple(grid_size)
        self._dt = dt
        nx, ny = grid_size
        p = np.require(p, np.float64, ["F_CONTIGUOUS"])
        u = np.require(u, np.float64, ["F_CONTIGUOUS"])
        self._params = None if storage is None else ctypes.c_int(storage)
        params = None if self._params is None else ctypes.cast(ctypes.byref(self._params),
                                                               ctypes.c_void_p)
        self.ptr = self.c_init(nx, ny, ctypes.c_double(dt), p, u, None, params)
        if not self.ptr:
            raise RuntimeError("plugin init failed (see stderr)")

    @property
    def dt(self):
        return self._dt

    def vars(self):
        if self.ptr is None:
            raise RuntimeError("The plugin has not been initialised. Call .init() first")
        rho = np.empty(self.grid_size, dtype=np.float64, order="F")
        u = np.empty(self.grid_size + (2,), dtype=np.float64, order="F")
        self.c_vars(self.ptr, rho, u)
        return rho, u

    def step(self, omega):
        if self.ptr is not None:
            self.c_step(self.ptr, ctypes.c_double(omega))

    def free(self):
        if self.ptr is not None:
            self.c_free(self.ptr)
            self.ptr = None


class TGVStudy:
    """Analytic Taylor-Green vortex on an (nx, ny) grid of cell centres (x fastest)."""

    def __init__(self, grid_size, umax, k, nu):
        self.umax, self.nu, self.k = umax, nu, k
        nx, ny = grid_size
        x = np.arange(nx, dtype=np.float64) + 0.5
        y = np.arange(ny, dtype=np.float64) + 0.5
        self._xv, self._yv = np.meshgrid(x, y, indexing="ij")

    @property
    def decay_time(self):
        kx, ky = self.k
        return 1.0 / (self.nu * (kx ** 2 + ky ** 2))

    @property
    def grid(self):
        return self._xv, self._yv

    def solution(self, t, ret_stress_tensor=False):
        td = self.decay_time
        kx, ky = self.k
        U = self.umax
        x, y = self._xv, self._yv
        e1, e2 = np.exp(-t / td), np.exp(-2.0 * t / td)
        u = np.empty(x.shape + (2,), order="F")
        u[:, :, 0] = -U * np.sqrt(ky / kx) * np.cos(kx * x) * np.sin(ky * y) * e1
        u[:, :, 1] = U * np.sqrt(kx / ky) * np.sin(kx * x) * np.cos(ky * y) * e1
        p = np.asfortranarray(-0.25 * U ** 2 * ((ky / kx) * np.cos(2 * kx * x)
                                                + (kx / ky) * np.cos(2 * ky * y)) * e2)
        if ret_stress_tensor:
            S = np.empty(x.shape + (3,), order="F")
            S[:, :, 0] = U * np.sqrt(kx * ky) * np.sin(kx * x) * np.sin(ky * y) * e1
            S[:, :, 1] = 0.5 * U * (np.sqrt(kx ** 3 / ky) - np.sqrt(ky ** 3 / kx)) \
                * np.cos(kx * x) * np.cos(ky * y) * e1
            S[:, :, 2] = -S[:, :, 0]
            return p, u, S
        return p, u

    def norm1(self, t, u):
        """max |u - u_ref| (velocity-magnitude error, infinity norm)."""
        _, ur = self.solution(t)
        return float(np.max(np.hypot(u[:, :, 0] - ur[:, :, 0], u[:, :, 1] - ur[:, :, 1])))

    def norm2(self, t, u):
        """|| |u_ref| - |u| ||_2 / || |u_ref| ||_2."""
        _, ur = self.solution(t)
        mag, magref = np.hypot(u[:, :, 0], u[:, :, 1]), np.hypot(ur[:, :, 0], ur[:, :, 1])
        return float(np.linalg.norm(magref - mag) / np.linalg.norm(magref))

    def norm2x(self, t, ux):
        _, ur = self.solution(t)
        return float(np.linalg.norm(ux - ur[:, :, 0]) / np.linalg.norm(ur[:, :, 0]))

    def norm2y(self, t, uy):
        _, ur = self.solution(t)
        return float(np.linalg.norm(uy - ur[:, :, 1]) / np.linalg.norm(ur[:, :, 1]))


def run(sim: SimPlugin, omega: float, nsteps: int):
    """nsteps plugin steps; returns (simulated time, wall seconds)."""
    start = time.time()
    tm = 0.0
    for _ in range(nsteps):
        sim.step(omega)
        tm += sim.dt
    vars_ = sim.vars()  # forces completion of the queued device work
    del vars_
    return tm, time.time() - start


def perfsweep(sim: SimPlugin, lo: int, hi: int, step: int = 20, tsteps: int = 1000,
              dt: float | None = None):
    """MLUPS of TGV runs for grids lo..hi (Re = 100, umax = 0.01 lo/n); returns
    (grids, mlups) arrays."""
    grids = list(range(lo, hi, step))
    mlups = np.empty(len(grids))
    cs = np.sqrt(1.0 / 3.0)
    for i, n in enumerate(grids):
        umax = 0.01 / (n / lo)
        nu = umax * n / 100.0
        tau = nu / cs ** 2
        dt_ = 1.0 if dt is None else dt
        omega = dt_ / (tau + 0.5 * dt_)
        case = TGVStudy((n, n), umax=umax, k=(2 * np.pi / n, 2 * np.pi / n), nu=nu)
        p0, u0 = case.solution(0.0)
        sim.init((n, n), dt_, p0, u0)
        _, el = run(sim, omega, tsteps)
        mlups[i] = n * n / 1e6 * tsteps / el
        sim.free()
    return np.array(grids), mlups
