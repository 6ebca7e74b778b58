"""Python mirror of the reference's Fortran module API on the GPU engine.

Names, argument meaning and buffer bookkeeping follow ivan-pi/periodic-lbm:

  alloc_grid / dealloc_grid / set_properties / set_pdf_to_equilibrium / update_macros /
  perform_step / perform_triple_step / equilibrium / cx / cy / csqr   src/fvm_bardow.F90
  perform_lbm_step / lbm_stream                                       src/periodic_lbm.f90
  collide_bgk                                                         src/collision_bgk.F90
  collide_trt / magic_number / lambda_d                               src/collision_trt.F90
  collide_rr                                                          src/collision_regularized.F90
  TaylorGreen (taylor_green_t) / pi                                   src/benchmarks/taylor_green.f90
  ShearLayer (shear_layer_t)                                          src/benchmarks/shear_layer.f90
  stream_fvm_bardow / stream_fdm_bardow / stream_fdm_sofonea          src/fvm_bardow.F90
  perform_dugks_step / dugks_collide / dugks_stream / tau_over_dt     src/periodic_dugks.F90

Host arrays use the reference's layouts: grid.f is f(ld, nx, 0:8, nf) and grid.mf is
mf(ny, nx, 3), both Fortran-ordered numpy arrays; grid.rho / ux / uy are views of mf.
Populations live on the device; grid.f is a host mirror refreshed by sync_host().
Buffer ids iold / inew / imid are 1-based as in the reference.
"""
from __future__ import annotations

import ctypes
import math
from typing import Callable, Optional

import numpy as np

from . import _abi
from ._abi import check, dptr, lib

cx = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=np.float64)
cy = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=np.float64)
w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, dtype=np.float64)
csqr = 1.0 / 3.0
pi = 4.0 * math.atan(1.0)


def padded_ld(ny: int) -> int:
    """src/fvm_bardow.F90:144-147: leading dimension rounded up to a multiple of 16."""
    r = ny % 16
    return ny + (16 - r) if r else ny


class Lattice:
    """Owner of one device lattice (plbm_create ... plbm_destroy)."""

    def __init__(self, nx: int, ny: int, nf: int = 2, storage: str | int = "f64"):
        st = _abi.STORAGE_IDS.get(storage, storage)
        h = ctypes.c_void_p()
        check(lib().plbm_create(nx, ny, nf, st, ctypes.byref(h)))
        self.h = h
        self.nx, self.ny, self.nf, self.storage = nx, ny, nf, st

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            check(lib().plbm_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- parameters / info
    def info(self) -> _abi.LatticeInfo:
        out = _abi.LatticeInfo()
        check(lib().plbm_get_info(self.h, ctypes.byref(out)))
        return out

    def set_params(self, omega: float, trt_magic: float = 0.25):
        check(lib().plbm_set_params(self.h, omega, trt_magic))

    def set_f16_scale(self, scale: float):
        check(lib().plbm_set_f16_scale(self.h, scale))

    def set_stream(self, stream_ptr: int | None):
        check(lib().plbm_set_stream(self.h, stream_ptr))

    def buffer_ptr(self, buf: int) -> int:
        p = ctypes.c_void_p()
        check(lib().plbm_buffer_ptr(self.h, buf, ctypes.byref(p)))
        return p.value

    # --- launch shape
    def set_launch(self, model: str | int, block: int, xcd_remap: int = 1):
        check(lib().plbm_set_launch(self.h, _abi.MODEL_IDS.get(model, model), block, xcd_remap))

    def get_launch(self, model: str | int) -> tuple[int, int]:
        b, r = ctypes.c_int(), ctypes.c_int()
        check(lib().plbm_get_launch(self.h, _abi.MODEL_IDS.get(model, model), ctypes.byref(b),
                                    ctypes.byref(r)))
        return b.value, r.value

    def set_steps_per_launch(self, model: str | int, steps: int):
        """Temporal blocking depth of run(): 1 single steps, 2..8 fused steps per pass,
        0 automatic.  FP64 results are bit-identical for every value."""
        check(lib().plbm_set_steps_per_launch(self.h, _abi.MODEL_IDS.get(model, model), steps))

    def run_pass(self, model: str | int, steps: int, src: int, dst: int):
        """One pass of `steps` fused steps src -> dst (no intermediate states stored)."""
        check(lib().plbm_pass(self.h, _abi.MODEL_IDS.get(model, model), steps, src, dst))

    def get_steps_per_launch(self, model: str | int) -> int:
        v = ctypes.c_int()
        check(lib().plbm_get_steps_per_launch(self.h, _abi.MODEL_IDS.get(model, model),
                                              ctypes.byref(v)))
        return v.value

    def set_pass_segments(self, segments: int):
        """Column segments per row band of a pass (one wavefront each); 0 automatic.
        Results do not depend on it."""
        check(lib().plbm_set_pass_segments(self.h, segments))

    def get_pass_segments(self) -> int:
        v = ctypes.c_int()
        check(lib().plbm_get_pass_segments(self.h, ctypes.byref(v)))
        return v.value

    def set_run_timing(self, enable: bool):
        """Record HIP events around every launch of later run() calls."""
        check(lib().plbm_set_run_timing(self.h, int(bool(enable))))

    def run_launch_times(self) -> list[tuple[int, float]]:
        """(steps advanced, ms) of each launch of the last run() (waits for it)."""
        n = ctypes.c_int()
        check(lib().plbm_run_launch_times(self.h, 0, None, None, ctypes.byref(n)))
        d = (ctypes.c_int * max(1, n.value))()
        ms = (ctypes.c_float * max(1, n.value))()
        check(lib().plbm_run_launch_times(self.h, n.value, d, ms, ctypes.byref(n)))
        return [(d[k], ms[k]) for k in range(n.value)]

    def set_pass_lag_buffer(self, enable: bool):
        """Third buffer of runs in passes (default on): the last pass also writes the
        state one step before the latest; off, runs end on a single step instead.  Results
        are identical either way."""
        check(lib().plbm_set_pass_lag_buffer(self.h, int(bool(enable))))

    def autotune(self, model: str | int, buf: int = 0, reps: int = 10) -> float:
        """Time every launch shape of the fused step and the temporal-blocking depths; keep
        the fastest.  Returns ms per lattice-update step."""
        ms = ctypes.c_double()
        check(lib().plbm_autotune(self.h, _abi.MODEL_IDS.get(model, model), buf, reps,
                                  ctypes.byref(ms)))
        return ms.value

    # --- populations (0-based buffer ids)
    def upload_f(self, buf: int, f: np.ndarray):
        f = np.asfortranarray(f, dtype=np.float64)
        if f.shape[1:] != (self.nx, 9) or f.shape[0] < self.ny:
            raise ValueError(f"f must be (ld>=ny, nx, 9), got {f.shape}")
        check(lib().plbm_upload_f(self.h, buf, dptr(f), f.shape[0]))

    def download_f(self, buf: int, ld: int | None = None, out: np.ndarray | None = None):
        ld = padded_ld(self.ny) if ld is None else ld
        if out is None:
            out = np.zeros((ld, self.nx, 9), dtype=np.float64, order="F")
        check(lib().plbm_download_f(self.h, buf, dptr(out), out.shape[0]))
        return out

    def equilibrium_init(self, buf: int, rho, ux, uy):
        a = [np.asfortranarray(v, dtype=np.float64) for v in (rho, ux, uy)]
        for v in a:
            if v.shape != (self.ny, self.nx):
                raise ValueError(f"fields must be (ny, nx) = {(self.ny, self.nx)}")
        check(lib().plbm_equilibrium_init(self.h, buf, *(dptr(v) for v in a)))

    def equilibrium_init_dev(self, buf: int, rho_ptr: int, ux_ptr: int, uy_ptr: int):
        check(lib().plbm_equilibrium_init_dev(self.h, buf, rho_ptr, ux_ptr, uy_ptr))

    # --- stepping
    def step(self, model: str | int, src: int, dst: int):
        check(lib().plbm_step(self.h, _abi.MODEL_IDS.get(model, model), src, dst))

    def run(self, model: str | int, nsteps: int, iold: int, inew: int) -> tuple[int, int]:
        a, b = ctypes.c_int(iold), ctypes.c_int(inew)
        check(lib().plbm_run(self.h, _abi.MODEL_IDS.get(model, model), nsteps, ctypes.byref(a),
                             ctypes.byref(b)))
        return a.value, b.value

    def stream(self, src: int, dst: int):
        check(lib().plbm_stream(self.h, src, dst))

    def stream_op(self, op: str | int, src: int, dst: int, dt: float = 1.0):
        """op: "exact" | "fvm_bardow" | "fdm_bardow" | "fdm_sofonea"."""
        check(lib().plbm_stream_op(self.h, _abi.STREAM_OPS.get(op, op), src, dst, dt))

    def step_op(self, op: str | int, model: str | int, src: int, dst: int, dt: float = 1.0):
        """Streaming operator `op` then collision `model`, fused: src -> dst."""
        check(lib().plbm_step_op(self.h, _abi.STREAM_OPS.get(op, op),
                                 _abi.MODEL_IDS.get(model, model), src, dst, dt))

    def run_op(self, op: str | int, model: str | int, nsteps: int, iold: int, inew: int,
               dt: float = 1.0) -> tuple[int, int]:
        """nsteps fused (streaming operator, collision) steps with the reference's slot
        bookkeeping; FP64 storage runs them in temporally blocked passes."""
        a, b = ctypes.c_int(iold), ctypes.c_int(inew)
        check(lib().plbm_run_op(self.h, _abi.STREAM_OPS.get(op, op),
                                _abi.MODEL_IDS.get(model, model), nsteps, ctypes.byref(a),
                                ctypes.byref(b), dt))
        return a.value, b.value

    def set_op_steps_per_pass(self, steps: int):
        check(lib().plbm_set_op_steps_per_pass(self.h, steps))

    def get_op_steps_per_pass(self, op: str | int = "fvm_bardow", model: str | int = "bgk") -> int:
        v = ctypes.c_int()
        check(lib().plbm_get_op_steps_per_pass(self.h, _abi.STREAM_OPS.get(op, op),
                                               _abi.MODEL_IDS.get(model, model), ctypes.byref(v)))
        return v.value

    def collide(self, model: str | int, buf: int):
        check(lib().plbm_collide(self.h, _abi.MODEL_IDS.get(model, model), buf))

    def dugks_collide(self, src: int, dst: int, omega_full: float, omega_half: float):
        """dugks_collide: dst = BGK(src, omega_full), src = BGK(src, omega_half)."""
        check(lib().plbm_dugks_collide(self.h, src, dst, omega_full, omega_half))

    def dugks_stream(self, src: int, dst: int, dt: float):
        """dugks_stream: dst -= dt c_q . face flux balance of src."""
        check(lib().plbm_dugks_stream(self.h, src, dst, dt))

    def dugks_stream_faces(self, src: int, dst: int, dt: float, omega_face: float):
        """dugks_stream of the reference's -DDUGKS build: face collisions (rate omega_face)
        before the flux balance."""
        check(lib().plbm_dugks_stream_faces(self.h, src, dst, dt, omega_face))

    def dugks_step_faces(self, iold: int, inew: int, nsteps: int, dt: float,
                         omega_full: float, omega_half: float, omega_face: float):
        """nsteps x perform_dugks_step of the -DDUGKS build (slot rules of dugks_step)."""
        check(lib().plbm_dugks_step_faces(self.h, iold, inew, nsteps, dt, omega_full,
                                          omega_half, omega_face))

    def dugks_set_steps_per_pass(self, steps: int):
        """Passes of plbm_dugks_step: 1 = collide + flux every step, 2..3 = temporally
        blocked passes, 0 = automatic (bit-identical for every value)."""
        check(lib().plbm_dugks_set_steps_per_pass(self.h, steps))

    def dugks_get_steps_per_pass(self) -> int:
        v = ctypes.c_int()
        check(lib().plbm_dugks_get_steps_per_pass(self.h, ctypes.byref(v)))
        return v.value

    def dugks_step(self, iold: int, inew: int, nsteps: int, dt: float, omega_full: float,
                   omega_half: float):
        """nsteps x perform_dugks_step on slots (iold, inew); slot contents afterwards
        are the reference's buffers of the same index (ids swap once per step)."""
        check(lib().plbm_dugks_step(self.h, iold, inew, nsteps, dt, omega_full, omega_half))

    def copy(self, src: int, dst: int):
        check(lib().plbm_copy(self.h, src, dst))

    def sync(self):
        check(lib().plbm_sync(self.h))

    # --- read-back
    def macros(self, buf: int):
        rho, ux, uy = (np.zeros((self.ny, self.nx), order="F") for _ in range(3))
        check(lib().plbm_macros(self.h, buf, dptr(rho), dptr(ux), dptr(uy)))
        return rho, ux, uy

    def macros_dev(self, buf: int, rho_ptr: int, ux_ptr: int, uy_ptr: int):
        check(lib().plbm_macros_dev(self.h, buf, rho_ptr, ux_ptr, uy_ptr))

    def reduce(self, buf: int, uxa=None, uya=None) -> np.ndarray:
        out = np.zeros(_abi.NRED)
        if uxa is None:
            check(lib().plbm_reduce(self.h, buf, None, None, dptr(out)))
        else:
            a = np.asfortranarray(uxa, dtype=np.float64)
            b = np.asfortranarray(uya, dtype=np.float64)
            check(lib().plbm_reduce(self.h, buf, dptr(a), dptr(b), dptr(out)))
        return out

    def reduce_dev(self, buf: int, uxa_ptr: int | None = None, uya_ptr: int | None = None):
        out = np.zeros(_abi.NRED)
        check(lib().plbm_reduce_dev(self.h, buf, uxa_ptr, uya_ptr, dptr(out)))
        return out


# ================================================================= reference-named API

class LatticeGrid:
    """lattice_grid (src/fvm_bardow.F90:37-69) with device-resident populations."""

    def __init__(self):
        self.nx = self.ny = 0
        self.nu = self.dt = self.tau = self.omega = self.trt_magic = math.nan
        self.csqr = csqr
        self.iold = self.inew = self.imid = 0
        self.collision: Optional[Callable[["LatticeGrid"], None]] = None
        self.streaming: Optional[Callable[["LatticeGrid"], None]] = None
        self.logger: Optional[Callable[["LatticeGrid", int], None]] = None
        self.filename = "results"
        self.foldername: Optional[str] = None
        self.logfile = "lattice_grid_log.txt"
        self.logunit = None
        self.mf = None
        self.lattice: Optional[Lattice] = None
        self._f = None
        self._host_current = False  # host mirror holds the current state (edited by user ops)
        self._host_valid = False    # host mirror equals the device state

    # host views of the macroscopic fields (pointer aliases of mf, src/fvm_bardow.F90:164-166)
    @property
    def rho(self):
        return self.mf[:, :, 0]

    @property
    def ux(self):
        return self.mf[:, :, 1]

    @property
    def uy(self):
        return self.mf[:, :, 2]

    @property
    def f(self) -> np.ndarray:
        """Host mirror f(ld, nx, 0:8, nf), synchronised from the device on access."""
        sync_host(self)
        return self._f

    def set_output_folder(self, foldername: str, verbose: bool = False):
        import os
        os.makedirs(foldername, exist_ok=True)
        if verbose:
            print(f"mkdir: created directory '{foldername}'")
        self.foldername = foldername


def alloc_grid(nx: int, ny: int, nf: int = 2, log: bool = True,
               storage: str | int = "f64") -> LatticeGrid:
    """alloc_grid (src/fvm_bardow.F90:129-202).  `storage` selects the device population
    storage: "f64" (reference arithmetic), "f64s", "f32s", "f16s" (shifted)."""
    g = LatticeGrid()
    g.nx, g.ny = nx, ny
    g._f = np.zeros((padded_ld(ny), nx, 9, nf), dtype=np.float64, order="F")
    g.mf = np.zeros((ny, nx, 3), dtype=np.float64, order="F")
    g.inew, g.iold = 1, 2
    g.imid = 3 if nf > 2 else -1
    g.lattice = Lattice(nx, ny, nf, storage)
    if log:
        g.logunit = open(g.logfile, "w")
    return g


def dealloc_grid(grid: LatticeGrid) -> None:
    """dealloc_grid (src/fvm_bardow.F90:204-240)."""
    if grid.logunit is not None:
        grid.logunit.close()
        grid.logunit = None
    if grid.lattice is not None:
        grid.lattice.close()
        grid.lattice = None
    grid._f = None
    grid.mf = None


def set_properties(grid: LatticeGrid, nu: float, dt: float, magic: float | None = None):
    """set_properties (src/fvm_bardow.F90:242-269)."""
    grid.nu, grid.dt = nu, dt
    grid.csqr = csqr
    tau = (1.0 / csqr) * nu
    grid.tau = tau
    grid.omega = dt / (tau + 0.5 * dt)
    grid.trt_magic = magic if magic is not None else (tau / dt) ** 2
    print(" trt magic = ", grid.trt_magic)
    grid.lattice.set_params(grid.omega, grid.trt_magic)


def sync_host(grid: LatticeGrid) -> None:
    """Refresh the host mirror grid.f from the device."""
    if grid._host_current or grid._host_valid:
        return
    grid.lattice.sync()
    for b in range(grid._f.shape[3]):
        grid.lattice.download_f(b, out=grid._f[:, :, :, b])
    grid._host_valid = True


def sync_device(grid: LatticeGrid) -> None:
    """Upload the host mirror after it was written directly."""
    for b in range(grid._f.shape[3]):
        grid.lattice.upload_f(b, grid._f[:, :, :, b])
    grid._host_current = False
    grid._host_valid = True


def _require_device(grid: LatticeGrid):
    if grid._host_current:
        sync_device(grid)
    grid.lattice.set_params(grid.omega, grid.trt_magic)


def _device_written(grid: LatticeGrid):
    grid._host_current = False
    grid._host_valid = False


def _require_host(grid: LatticeGrid):
    sync_host(grid)
    grid._host_current = True


def lbm_stream(grid: LatticeGrid) -> None:
    """lbm_stream (src/periodic_lbm.f90:32-129): f(iold) -> f(inew) on the device."""
    _require_device(grid)
    grid.lattice.stream(grid.iold - 1, grid.inew - 1)
    _device_written(grid)


def _stream_op(grid: LatticeGrid, op: str):
    _require_device(grid)
    grid.lattice.stream_op(op, grid.iold - 1, grid.inew - 1, grid.dt)
    _device_written(grid)


def stream_fvm_bardow(grid: LatticeGrid) -> None:
    """stream_fvm_bardow (src/fvm_bardow.F90:393-509) on the device, time step grid.dt."""
    _stream_op(grid, "fvm_bardow")


# stream_fdm_bardow's stencil.  The reference chooses it when it is compiled
# (src/fvm_bardow.F90:591-669: -DFDM_WLS, -DFDM_WLS_GAUSS_V1, -DFDM_WLS_GAUSS_V2,
# -DFDM_ISO, otherwise central differences); set_fdm_bardow_stencil chooses it at run time.
FDM_BARDOW_STENCILS = {"default": "fdm_bardow", "wls": "fdm_bardow_wls",
                       "wls_gauss_v1": "fdm_bardow_wls_gauss_v1",
                       "wls_gauss_v2": "fdm_bardow_wls_gauss_v2", "iso": "fdm_bardow_iso"}
_fdm_bardow_op = "fdm_bardow"


def set_fdm_bardow_stencil(name: str) -> None:
    """Stencil of stream_fdm_bardow from now on: "default", "wls", "wls_gauss_v1",
    "wls_gauss_v2" or "iso" (the reference's FDM_* build options)."""
    global _fdm_bardow_op
    if name not in FDM_BARDOW_STENCILS:
        raise ValueError(f"unknown stencil {name!r}; one of {sorted(FDM_BARDOW_STENCILS)}")
    _fdm_bardow_op = FDM_BARDOW_STENCILS[name]


def stream_fdm_bardow(grid: LatticeGrid) -> None:
    """stream_fdm_bardow (src/fvm_bardow.F90:511-685) on the device, with the stencil of
    set_fdm_bardow_stencil (default: central differences)."""
    _stream_op(grid, _fdm_bardow_op)


def stream_fdm_sofonea(grid: LatticeGrid) -> None:
    """stream_fdm_sofonea (src/fvm_bardow.F90:688-893) on the device."""
    _stream_op(grid, "fdm_sofonea")


def _collide(grid: LatticeGrid, model: int):
    _require_device(grid)
    grid.lattice.collide(model, grid.inew - 1)
    _device_written(grid)


def collide_bgk(grid: LatticeGrid) -> None:
    """collide_bgk (src/collision_bgk.F90:17-82), in place on f(inew)."""
    _collide(grid, _abi.BGK)


def collide_bgk_split(grid: LatticeGrid) -> None:
    """collide_bgk as the reference's -DSPLIT build runs it (bgk_kernel_cache,
    src/collision_bgk.F90:84-176), in place on f(inew)."""
    _collide(grid, _abi.BGK_SPLIT)


def collide_trt(grid: LatticeGrid) -> None:
    """collide_trt (src/collision_trt.F90:41-160), in place on f(inew)."""
    _collide(grid, _abi.TRT)


def collide_rr(grid: LatticeGrid) -> None:
    """collide_rr (src/collision_regularized.F90:21-202), in place on f(inew)."""
    _collide(grid, _abi.RR)


def tau_over_dt(grid: LatticeGrid) -> float:
    """tau_over_dt (src/periodic_dugks.F90:40-44)."""
    return grid.tau / grid.dt


# DUGKS with face collisions: the reference builds it with -DDUGKS
# (src/periodic_dugks.F90:68,199,267,308); set_dugks_face_collisions chooses it at run time.
_dugks_faces = False


def set_dugks_face_collisions(enable: bool) -> None:
    """DUGKS as the reference's -DDUGKS build from now on (True): the half collision at
    0.75 of its rate and face collisions of rate 1/(4 tau/dt + 1) in dugks_stream."""
    global _dugks_faces
    _dugks_faces = bool(enable)


def _dugks_omega_half(grid: LatticeGrid) -> float:
    om = 1.0 / (tau_over_dt(grid) + 0.5)  # src/periodic_dugks.F90:56
    return 0.75 * om if _dugks_faces else om  # :68-71 with -DDUGKS


def _dugks_omega_face(grid: LatticeGrid) -> float:
    return 1.0 / (4.0 * tau_over_dt(grid) + 1.0)  # src/periodic_dugks.F90:179


def _dugks_steps(grid: LatticeGrid, nsteps: int) -> None:
    if _dugks_faces:
        grid.lattice.dugks_step_faces(grid.iold - 1, grid.inew - 1, nsteps, grid.dt,
                                      grid.omega, _dugks_omega_half(grid),
                                      _dugks_omega_face(grid))
    else:
        grid.lattice.dugks_step(grid.iold - 1, grid.inew - 1, nsteps, grid.dt, grid.omega,
                                _dugks_omega_half(grid))


def dugks_collide(grid: LatticeGrid) -> None:
    """dugks_collide (src/periodic_dugks.F90:46-77) on the device: f(inew) = BGK(f(iold),
    grid.omega), f(iold) = BGK(f(iold), 1/(tau/dt + 1/2)) in place."""
    _require_device(grid)
    grid.lattice.dugks_collide(grid.iold - 1, grid.inew - 1, grid.omega, _dugks_omega_half(grid))
    _device_written(grid)


def dugks_stream(grid: LatticeGrid) -> None:
    """dugks_stream (src/periodic_dugks.F90:168-299) on the device: f(inew) minus the
    face-flux balance of f(iold), time step grid.dt."""
    _require_device(grid)
    if _dugks_faces:
        grid.lattice.dugks_stream_faces(grid.iold - 1, grid.inew - 1, grid.dt,
                                        _dugks_omega_face(grid))
    else:
        grid.lattice.dugks_stream(grid.iold - 1, grid.inew - 1, grid.dt)
    _device_written(grid)


def perform_dugks_step(grid: LatticeGrid) -> None:
    """perform_dugks_step (src/periodic_dugks.F90:25-38): grid.collision, grid.streaming,
    swap.  With dugks_collide + dugks_stream in the slots the step runs fused."""
    if grid.collision is dugks_collide and grid.streaming is dugks_stream:
        _require_device(grid)
        _dugks_steps(grid, 1)
        _device_written(grid)
    else:
        _dispatch(grid, grid.collision, _ENGINE_COLLISIONS)
        _dispatch(grid, grid.streaming, _ENGINE_STREAMING)
    grid.iold, grid.inew = grid.inew, grid.iold


def perform_dugks_steps(grid: LatticeGrid, nsteps: int) -> None:
    """Engine extension: the effect of `nsteps` perform_dugks_step calls in one call; with
    dugks_collide + dugks_stream in the slots all steps but the last run in temporally
    blocked passes (bit-identical, same slot contents)."""
    if grid.collision is dugks_collide and grid.streaming is dugks_stream:
        _require_device(grid)
        _dugks_steps(grid, nsteps)
        _device_written(grid)
        if nsteps % 2:
            grid.iold, grid.inew = grid.inew, grid.iold
    else:
        for _ in range(nsteps):
            perform_dugks_step(grid)


_ENGINE_COLLISIONS = {collide_bgk: _abi.BGK, collide_trt: _abi.TRT, collide_rr: _abi.RR,
                      collide_bgk_split: _abi.BGK_SPLIT, dugks_collide: None}
_ENGINE_STREAMING = (lbm_stream, stream_fvm_bardow, stream_fdm_bardow, stream_fdm_sofonea,
                     dugks_stream)


def _dispatch(grid: LatticeGrid, op, engine_ops):
    if op not in engine_ops:
        _require_host(grid)
    op(grid)


_FUSABLE_STREAMING = {lbm_stream: "exact", stream_fvm_bardow: "fvm_bardow",
                      stream_fdm_bardow: "fdm_bardow", stream_fdm_sofonea: "fdm_sofonea"}


def _fusable_op(streaming):
    op = _FUSABLE_STREAMING.get(streaming)
    return _fdm_bardow_op if op == "fdm_bardow" else op


def perform_step(grid: LatticeGrid) -> None:
    """perform_step / perform_lbm_step (src/fvm_bardow.F90:307-320,
    src/periodic_lbm.f90:15-29): a pair of engine operators (lbm_stream or a FV/FD
    streaming operator, and collide_bgk/trt/rr) runs as one fused kernel."""
    model = _ENGINE_COLLISIONS.get(grid.collision)
    op = _fusable_op(grid.streaming)
    if model is not None and op is not None:
        _require_device(grid)
        grid.lattice.step_op(op, model, grid.iold - 1, grid.inew - 1, grid.dt)
        _device_written(grid)
    else:
        _dispatch(grid, grid.streaming, _ENGINE_STREAMING)
        _dispatch(grid, grid.collision, _ENGINE_COLLISIONS)
    grid.iold, grid.inew = grid.inew, grid.iold


perform_lbm_step = perform_step


def perform_steps(grid: LatticeGrid, nsteps: int) -> None:
    """Engine extension: the effect of `nsteps` perform_step calls in one call.  With an
    engine streaming operator (lbm_stream or a FV/FD operator) and an engine collision the
    steps run in temporally blocked passes (Lattice.run / run_op; bit-identical to single
    steps in FP64 storage, same slot contents); any other pair is stepped one
    perform_step at a time."""
    model = _ENGINE_COLLISIONS.get(grid.collision)
    op = _fusable_op(grid.streaming)
    if model is not None and op is not None:
        _require_device(grid)
        iold, inew = grid.lattice.run_op(op, model, nsteps, grid.iold - 1, grid.inew - 1,
                                         grid.dt)
        _device_written(grid)
        grid.iold, grid.inew = iold + 1, inew + 1
    else:
        for _ in range(nsteps):
            perform_step(grid)


def perform_triple_step(grid: LatticeGrid) -> None:
    """perform_triple_step (src/fvm_bardow.F90:322-340)."""
    _dispatch(grid, grid.streaming, _ENGINE_STREAMING)
    if grid._host_current:
        grid._f[:, :, :, grid.iold - 1] = grid._f[:, :, :, grid.inew - 1]
    else:
        grid.lattice.copy(grid.inew - 1, grid.iold - 1)
        _device_written(grid)
    _dispatch(grid, grid.collision, _ENGINE_COLLISIONS)
    grid.iold, grid.inew, grid.imid = grid.inew, grid.imid, grid.iold


def set_pdf_to_equilibrium(grid: LatticeGrid) -> None:
    """set_pdf_to_equilibrium (src/fvm_bardow.F90:272-305): f(iold) = feq(rho, u)."""
    if grid._host_current:
        sync_device(grid)
    grid.lattice.equilibrium_init(grid.iold - 1, grid.rho, grid.ux, grid.uy)
    _device_written(grid)


def update_macros(grid: LatticeGrid) -> None:
    """update_macros (src/fvm_bardow.F90:343-390): rho, u of f(inew) into grid.mf."""
    if grid._host_current:
        sync_device(grid)
    rho, ux, uy = grid.lattice.macros(grid.inew - 1)
    grid.mf[:, :, 0], grid.mf[:, :, 1], grid.mf[:, :, 2] = rho, ux, uy


def equilibrium(rho: float, ux: float, uy: float) -> np.ndarray:
    """equilibrium (src/fvm_bardow.F90:99-126), host side, same evaluation order."""
    uxx, uyy = ux * ux, uy * uy
    indp = 1.0 - 1.5 * (uxx + uyy)
    w0, ws, wd = 4.0 / 9.0, 1.0 / 9.0, 1.0 / 36.0
    uxpy, uxmy = ux + uy, ux - uy
    return np.array([
        (w0 * rho) * indp,
        (ws * rho) * ((indp + 3.0 * ux) + 4.5 * uxx),
        (ws * rho) * ((indp + 3.0 * uy) + 4.5 * uyy),
        (ws * rho) * ((indp - 3.0 * ux) + 4.5 * uxx),
        (ws * rho) * ((indp - 3.0 * uy) + 4.5 * uyy),
        (wd * rho) * ((indp + 3.0 * uxpy) + (4.5 * uxpy) * uxpy),
        (wd * rho) * ((indp - 3.0 * uxmy) + (4.5 * uxmy) * uxmy),
        (wd * rho) * ((indp - 3.0 * uxpy) + (4.5 * uxpy) * uxpy),
        (wd * rho) * ((indp + 3.0 * uxmy) + (4.5 * uxmy) * uxmy),
    ])


def magic_number(le: float, ld: float) -> float:
    """src/collision_trt.F90:24-28."""
    return lib().plbm_magic_number(le, ld)


def lambda_d(omega: float, x: float) -> float:
    """src/collision_trt.F90:30-34."""
    return lib().plbm_lambda_d(omega, x)


class TaylorGreen:
    """taylor_green_t (src/benchmarks/taylor_green.f90:12-84)."""

    def __init__(self, nx: int, ny: int, kx: float, ky: float, umax: float, nu: float):
        self.nx, self.ny, self.kx, self.ky, self.umax, self.nu = nx, ny, kx, ky, umax, nu
        self.td = 1.0 / (nu * (kx * kx + ky * ky))

    def decay_time(self) -> float:
        return self.td

    def eval(self, t: float):
        """Returns (p, ux, uy) as (ny, nx) Fortran-ordered arrays."""
        p, ux, uy = (np.zeros((self.ny, self.nx), order="F") for _ in range(3))
        check(lib().plbm_field_taylor_green(self.nx, self.ny, self.kx, self.ky, self.umax,
                                            self.nu, t, dptr(p), dptr(ux), dptr(uy)))
        return p, ux, uy


class VortexCase:
    """vortex_case_t (src/benchmarks/barotropic_vortex_case.F90:20-83)."""

    def __init__(self, U0, xc, yc, Rc, eps, rho0=1.0, csqr_=1.0 / 3.0):
        self.U0, self.xc, self.yc, self.Rc, self.eps = U0, xc, yc, Rc, eps
        self.rho0, self.csqr = rho0, csqr_

    def eval(self, nx: int, ny: int):
        rho, ux, uy = (np.zeros((ny, nx), order="F") for _ in range(3))
        check(lib().plbm_field_barotropic_vortex(nx, ny, self.U0, self.xc, self.yc, self.Rc,
                                                 self.eps, self.rho0, self.csqr, dptr(rho),
                                                 dptr(ux), dptr(uy)))
        return rho, ux, uy


class ShearLayer:
    """shear_layer_t (src/benchmarks/shear_layer.f90:10-52)."""

    def __init__(self, length: float, U0: float, k: float, delta: float, rho0: float = 1.0):
        self.length, self.U0, self.k, self.delta, self.rho0 = length, U0, k, delta, rho0

    def eval(self, nx: int, ny: int):
        rho, ux, uy = (np.zeros((ny, nx), order="F") for _ in range(3))
        check(lib().plbm_field_shear_layer(nx, ny, self.length, self.U0, self.k, self.delta,
                                           dptr(rho), dptr(ux), dptr(uy)))
        rho[:] = self.rho0
        return rho, ux, uy


def add_noise(ux: np.ndarray, uy: np.ndarray, seed: int, amp: float) -> None:
    """In place: u += amp*(2U-1) with the engine's counter-based generator."""
    if not (ux.flags.f_contiguous and uy.flags.f_contiguous):
        raise ValueError("fields must be Fortran-contiguous")
    check(lib().plbm_field_add_noise(ux.size, seed, amp, dptr(ux), dptr(uy)))


def noise(seed: int, counter: int) -> float:
    return lib().plbm_noise(seed, counter)


def copy_bandwidth(nbytes: int, iters: int = 20) -> float:
    """Measured HBM copy bandwidth in GB/s (read + write bytes / kernel time)."""
    ms = ctypes.c_double()
    check(lib().plbm_copy_bandwidth(nbytes, iters, ctypes.byref(ms)))
    return 2.0 * nbytes / (ms.value * 1e-3) / 1e9
