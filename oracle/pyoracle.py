"""Test infrastructure only — ctypes bindings for the oracle.

Two checkers live under oracle/:

* ``Oracle``   — liboracle.so, the C restatement of the reference hot path (oracle.c).
* ``RefLib``   — oracle/_ref/libplbm_ref.so, the reference Fortran itself (flang-built
                 from ivan-pi/periodic-lbm src/ by oracle/Makefile), driven through
                 ref_driver.f90.  Present only where the reference tree was available
                 at build time.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may import this
module.  All arrays are numpy float64 in the reference's Fortran layout: populations
``f(ld, nx, 0:8)`` and macroscopic fields ``(ny, nx)``, both column-major.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBORACLE = HERE / "liboracle.so"
LIBREF = HERE / "_ref" / "libplbm_ref.so"
LIBREF_SPLIT = HERE / "_ref" / "split" / "libplbm_ref.so"  # built with -DSPLIT
REF_APP = HERE / "_ref" / "main_taylor_green_ref"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i = ctypes.c_int
_d = ctypes.c_double

MODELS = {"bgk": 0, "trt": 1, "rr": 2, "bgk_split": 3}


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim <= 1)
    return a.ctypes.data_as(_dp)


def padded_ld(ny: int) -> int:
    """src/fvm_bardow.F90:144-147."""
    r = ny % 16
    return ny + (16 - r) if r else ny


def new_f(nx: int, ny: int, ld: int | None = None) -> np.ndarray:
    ld = padded_ld(ny) if ld is None else ld
    return np.zeros((ld, nx, 9), dtype=np.float64, order="F")


def new_field(nx: int, ny: int) -> np.ndarray:
    return np.zeros((ny, nx), dtype=np.float64, order="F")


class Oracle:
    """liboracle.so (C restatement)."""

    def __init__(self, path: Path | str = LIBORACLE):
        if not Path(path).exists():
            raise FileNotFoundError(f"{path} not built: run `make -C oracle liboracle`")
        L = ctypes.CDLL(str(path))
        self.lib = L
        L.orc_equilibrium.argtypes = [_d, _d, _d, _dp]
        L.orc_equilibrium_shifted.argtypes = [_d, _d, _d, _dp]
        L.orc_set_properties.argtypes = [_d, _d, _d, _i, _dp]
        L.orc_magic_number.argtypes = [_d, _d]
        L.orc_magic_number.restype = _d
        L.orc_lambda_d.argtypes = [_d, _d]
        L.orc_lambda_d.restype = _d
        L.orc_padded_ld.argtypes = [_i]
        L.orc_padded_ld.restype = _i
        L.orc_stream.argtypes = [_i, _i, _i, _dp, _dp]
        for name in ("orc_stream_fvm_bardow", "orc_stream_fdm_bardow", "orc_stream_fdm_sofonea"):
            getattr(L, name).argtypes = [_i, _i, _i, _dp, _dp, _d]
        L.orc_stream_fdm_bardow_v.argtypes = [_i, _i, _i, _i, _dp, _dp, _d]
        L.orc_collide_bgk.argtypes = [_i, _i, _i, _dp, _d]
        L.orc_dugks_collide.argtypes = [_i, _i, _i, _dp, _dp, _d, _d]
        L.orc_dugks_stream.argtypes = [_i, _i, _i, _dp, _dp, _d]
        L.orc_dugks_stream_face.argtypes = [_i, _i, _i, _dp, _dp, _d, _d]
        L.orc_collide_trt.argtypes = [_i, _i, _i, _dp, _d, _d]
        L.orc_collide_rr.argtypes = [_i, _i, _i, _dp, _d]
        L.orc_collide.argtypes = [_i, _i, _i, _i, _dp, _d, _d]
        L.orc_collide_shifted.argtypes = [_i, _i, _i, _i, _dp, _d, _d]
        L.orc_update_macros.argtypes = [_i, _i, _i, _dp, _dp, _dp, _dp]
        L.orc_update_macros_shifted.argtypes = [_i, _i, _i, _dp, _dp, _dp, _dp]
        L.orc_set_pdf_to_equilibrium.argtypes = [_i, _i, _i, _dp, _dp, _dp, _dp]
        L.orc_run.argtypes = [_i, _i, _i, _i, _dp, _dp, _i, _d, _d, _ip]
        L.orc_tgv_eval.argtypes = [_i, _i, _d, _d, _d, _d, _d, _dp, _dp, _dp]
        L.orc_shear_eval.argtypes = [_i, _i, _d, _d, _d, _d, _dp, _dp, _dp]
        L.orc_noise.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.orc_noise.restype = _d
        L.orc_l2_norm.argtypes = [_i, _dp, _dp, _dp, _dp]
        L.orc_l2_norm.restype = _d

    # scalar helpers
    def equilibrium(self, rho, ux, uy, shifted=False):
        out = np.zeros(9)
        fn = self.lib.orc_equilibrium_shifted if shifted else self.lib.orc_equilibrium
        fn(rho, ux, uy, _ptr(out))
        return out

    def set_properties(self, nu, dt, magic=None):
        out = np.zeros(6)
        self.lib.orc_set_properties(nu, dt, 0.0 if magic is None else magic,
                                    0 if magic is None else 1, _ptr(out))
        return dict(zip(["nu", "dt", "tau", "omega", "trt_magic", "csqr"], out))

    def magic_number(self, le, ld):
        return self.lib.orc_magic_number(le, ld)

    def lambda_d(self, omega, x):
        return self.lib.orc_lambda_d(omega, x)

    def noise(self, seed, counter):
        return self.lib.orc_noise(seed, counter)

    # array kernels (in place where the reference is in place)
    def stream(self, nx, ny, fsrc, fdst):
        self.lib.orc_stream(nx, ny, fsrc.shape[0], _ptr(fsrc), _ptr(fdst))

    # stream_fdm_bardow's compile-time stencils (src/fvm_bardow.F90:591-669)
    FDM_VARIANTS = {"fdm_bardow_wls": 1, "fdm_bardow_wls_gauss_v1": 2,
                    "fdm_bardow_wls_gauss_v2": 3, "fdm_bardow_iso": 4}

    def stream_op(self, op, nx, ny, fsrc, fdst, dt=1.0):
        """op: "exact" | "fvm_bardow" | "fdm_bardow" | "fdm_sofonea" | one of
        FDM_VARIANTS."""
        if op == "exact":
            return self.stream(nx, ny, fsrc, fdst)
        if op in self.FDM_VARIANTS:
            return self.lib.orc_stream_fdm_bardow_v(self.FDM_VARIANTS[op], nx, ny, fsrc.shape[0],
                                                    _ptr(fsrc), _ptr(fdst), dt)
        getattr(self.lib, "orc_stream_" + op)(nx, ny, fsrc.shape[0], _ptr(fsrc), _ptr(fdst), dt)

    def collide(self, model, nx, ny, f, omega, trt_magic=0.25):
        self.lib.orc_collide(MODELS.get(model, model), nx, ny, f.shape[0], _ptr(f), omega,
                             trt_magic)

    def collide_trt(self, nx, ny, f, lambda_e, lambda_d):
        self.lib.orc_collide_trt(nx, ny, f.shape[0], _ptr(f), lambda_e, lambda_d)

    def dugks_collide(self, nx, ny, fold, fnew, omega_full, omega_half):
        """dugks_collide: fnew = BGK(fold, omega_full); fold = BGK(fold, omega_half)."""
        self.lib.orc_dugks_collide(nx, ny, fold.shape[0], _ptr(fold), _ptr(fnew), omega_full,
                                   omega_half)

    def dugks_stream(self, nx, ny, ft, fp, dt):
        """dugks_stream: fp -= flux balance of the face values of ft."""
        self.lib.orc_dugks_stream(nx, ny, ft.shape[0], _ptr(ft), _ptr(fp), dt)

    def dugks_stream_face(self, nx, ny, ft, fp, dt, omega_face):
        """dugks_stream of the -DDUGKS build: face collisions (rate omega_face) before the
        flux balance."""
        self.lib.orc_dugks_stream_face(nx, ny, ft.shape[0], _ptr(ft), _ptr(fp), dt, omega_face)

    def dugks_run(self, nx, ny, fa, fb, nsteps, dt, omega_full, omega_half):
        """perform_dugks_step x nsteps on (fa = iold, fb = inew), with the reference's
        buffer swap; returns (iold, inew) after the last step."""
        for _ in range(nsteps):
            self.dugks_collide(nx, ny, fa, fb, omega_full, omega_half)
            self.dugks_stream(nx, ny, fa, fb, dt)
            fa, fb = fb, fa
        return fa, fb

    def collide_shifted(self, model, nx, ny, g, omega, lambda_odd=0.0):
        self.lib.orc_collide_shifted(MODELS.get(model, model), nx, ny, g.shape[0], _ptr(g),
                                     omega, lambda_odd)

    def update_macros(self, nx, ny, f, shifted=False):
        rho, ux, uy = new_field(nx, ny), new_field(nx, ny), new_field(nx, ny)
        fn = self.lib.orc_update_macros_shifted if shifted else self.lib.orc_update_macros
        fn(nx, ny, f.shape[0], _ptr(f), _ptr(rho), _ptr(ux), _ptr(uy))
        return rho, ux, uy

    def set_pdf_to_equilibrium(self, nx, ny, rho, ux, uy, f):
        self.lib.orc_set_pdf_to_equilibrium(nx, ny, f.shape[0], _ptr(rho), _ptr(ux), _ptr(uy),
                                            _ptr(f))

    def run(self, model, nx, ny, f_iold, f_inew, nsteps, omega, trt_magic=0.25):
        """perform_lbm_step x nsteps.  Returns (latest, previous): the buffer holding the
        last post-collision state and the one holding the state a step earlier."""
        flag = ctypes.c_int(0)
        self.lib.orc_run(MODELS.get(model, model), nx, ny, f_iold.shape[0], _ptr(f_iold),
                         _ptr(f_inew), nsteps, omega, trt_magic, ctypes.byref(flag))
        return (f_inew, f_iold) if flag.value else (f_iold, f_inew)

    def tgv(self, nx, ny, kx, ky, umax, nu, t):
        p, ux, uy = new_field(nx, ny), new_field(nx, ny), new_field(nx, ny)
        self.lib.orc_tgv_eval(nx, ny, kx, ky, umax, nu, t, _ptr(p), _ptr(ux), _ptr(uy))
        return p, ux, uy

    def shear(self, nx, ny, length, u0, k, delta):
        rho, ux, uy = new_field(nx, ny), new_field(nx, ny), new_field(nx, ny)
        self.lib.orc_shear_eval(nx, ny, length, u0, k, delta, _ptr(rho), _ptr(ux), _ptr(uy))
        return rho, ux, uy

    def l2_norm(self, ux, uy, uxa, uya):
        n = ux.size
        return self.lib.orc_l2_norm(n, _ptr(np.ravel(ux, "F")), _ptr(np.ravel(uy, "F")),
                                    _ptr(np.ravel(uxa, "F")), _ptr(np.ravel(uya, "F")))


class RefLib:
    """oracle/_ref/libplbm_ref.so: the reference Fortran through ref_driver.f90."""

    def __init__(self, path: Path | str = LIBREF):
        if not Path(path).exists():
            raise FileNotFoundError(f"{path} not built (needs the reference tree)")
        L = ctypes.CDLL(str(path))
        self.lib = L
        L.ref_create.argtypes = [_i, _i, _i, _i]
        L.ref_create.restype = ctypes.c_void_p
        for name, args in {
            "ref_destroy": [], "ref_set_pdf_to_equilibrium": [], "ref_stream": [],
            "ref_collide": [], "ref_update_macros": [],
        }.items():
            getattr(L, name).argtypes = [ctypes.c_void_p] + args
        L.ref_set_model.argtypes = [ctypes.c_void_p, _i]
        L.ref_set_properties.argtypes = [ctypes.c_void_p, _d, _d, _d, _i]
        L.ref_get_params.argtypes = [ctypes.c_void_p, _dp]
        L.ref_set_omega.argtypes = [ctypes.c_void_p, _d, _d]
        L.ref_set_macros.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.ref_get_macros.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.ref_perform_lbm_step.argtypes = [ctypes.c_void_p, _i]
        L.ref_get_indices.argtypes = [ctypes.c_void_p, _ip]
        L.ref_get_f.argtypes = [ctypes.c_void_p, _i, _dp]
        L.ref_set_f.argtypes = [ctypes.c_void_p, _i, _dp]
        L.ref_equilibrium.argtypes = [_d, _d, _d, _dp]
        L.ref_magic_number.argtypes = [_d, _d]
        L.ref_magic_number.restype = _d
        L.ref_lambda_d.argtypes = [_d, _d]
        L.ref_lambda_d.restype = _d
        L.ref_tgv_eval.argtypes = [_i, _i, _d, _d, _d, _d, _d, _dp, _dp, _dp]
        L.ref_shear_eval.argtypes = [_i, _i, _d, _d, _d, _d, _dp, _dp, _dp]
        L.ref_vortex_eval.argtypes = [_i, _i, _d, _d, _d, _d, _d, _dp, _dp, _dp]
        L.ref_set_streaming.argtypes = [ctypes.c_void_p, _i]
        L.ref_streaming.argtypes = [ctypes.c_void_p]
        L.ref_perform_step.argtypes = [ctypes.c_void_p, _i]
        L.ref_perform_triple_step.argtypes = [ctypes.c_void_p, _i]
        L.ref_set_dugks.argtypes = [ctypes.c_void_p]
        L.ref_perform_dugks_step.argtypes = [ctypes.c_void_p, _i]

    def equilibrium(self, rho, ux, uy):
        out = np.zeros(9)
        self.lib.ref_equilibrium(rho, ux, uy, _ptr(out))
        return out

    def magic_number(self, le, ld):
        return self.lib.ref_magic_number(le, ld)

    def lambda_d(self, omega, x):
        return self.lib.ref_lambda_d(omega, x)

    def tgv(self, nx, ny, kx, ky, umax, nu, t):
        p, ux, uy = new_field(nx, ny), new_field(nx, ny), new_field(nx, ny)
        self.lib.ref_tgv_eval(nx, ny, kx, ky, umax, nu, t, _ptr(p), _ptr(ux), _ptr(uy))
        return p, ux, uy

    def shear(self, nx, ny, length, u0, k, delta):
        rho, ux, uy = new_field(nx, ny), new_field(nx, ny), new_field(nx, ny)
        self.lib.ref_shear_eval(nx, ny, length, u0, k, delta, _ptr(rho), _ptr(ux), _ptr(uy))
        return rho, ux, uy

    def vortex(self, nx, ny, u0, xc, yc, rc, eps):
        rho, ux, uy = new_field(nx, ny), new_field(nx, ny), new_field(nx, ny)
        self.lib.ref_vortex_eval(nx, ny, u0, xc, yc, rc, eps, _ptr(rho), _ptr(ux), _ptr(uy))
        return rho, ux, uy

    def grid(self, nx, ny, model, nf=2):
        return RefGrid(self, nx, ny, model, nf)


class RefGrid:
    """One reference ``lattice_grid`` (alloc_grid ... dealloc_grid)."""

    def __init__(self, ref: RefLib, nx, ny, model, nf=2):
        self.ref, self.nx, self.ny = ref, nx, ny
        self.h = ref.lib.ref_create(nx, ny, nf, MODELS.get(model, model))
        self.ld = self.indices()[3]

    def close(self):
        if self.h:
            self.ref.lib.ref_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_model(self, model):
        self.ref.lib.ref_set_model(self.h, MODELS.get(model, model))

    def set_properties(self, nu, dt, magic=None):
        self.ref.lib.ref_set_properties(self.h, nu, dt, 0.0 if magic is None else magic,
                                        0 if magic is None else 1)

    def set_omega(self, omega, magic):
        self.ref.lib.ref_set_omega(self.h, omega, magic)

    def params(self):
        out = np.zeros(6)
        self.ref.lib.ref_get_params(self.h, _ptr(out))
        return dict(zip(["nu", "dt", "tau", "omega", "trt_magic", "csqr"], out))

    def set_macros(self, rho, ux, uy):
        f = lambda a: np.asfortranarray(a, dtype=np.float64)
        self.ref.lib.ref_set_macros(self.h, _ptr(f(rho)), _ptr(f(ux)), _ptr(f(uy)))

    def macros(self):
        rho, ux, uy = (new_field(self.nx, self.ny) for _ in range(3))
        self.ref.lib.ref_get_macros(self.h, _ptr(rho), _ptr(ux), _ptr(uy))
        return rho, ux, uy

    def set_pdf_to_equilibrium(self):
        self.ref.lib.ref_set_pdf_to_equilibrium(self.h)

    def step(self, n=1):
        self.ref.lib.ref_perform_lbm_step(self.h, n)

    def stream(self):
        self.ref.lib.ref_stream(self.h)

    STREAMING = {"exact": 0, "fvm_bardow": 1, "fdm_bardow": 2, "fdm_sofonea": 3}

    def set_streaming(self, op):
        self.ref.lib.ref_set_streaming(self.h, self.STREAMING.get(op, op))

    def streaming(self):
        self.ref.lib.ref_streaming(self.h)

    def perform_step(self, n=1):
        self.ref.lib.ref_perform_step(self.h, n)

    def triple_step(self, n=1):
        """perform_triple_step (src/fvm_bardow.F90:322-340) n times (nf = 3 grids)."""
        self.ref.lib.ref_perform_triple_step(self.h, n)

    def collide(self):
        self.ref.lib.ref_collide(self.h)

    def set_dugks(self):
        """grid%collision => dugks_collide, grid%streaming => dugks_stream."""
        self.ref.lib.ref_set_dugks(self.h)

    def dugks_step(self, n=1):
        self.ref.lib.ref_perform_dugks_step(self.h, n)

    def update_macros(self):
        self.ref.lib.ref_update_macros(self.h)

    def indices(self):
        out = (ctypes.c_int * 4)()
        self.ref.lib.ref_get_indices(self.h, out)
        return list(out)  # iold, inew, imid, ld (1-based buffer ids)

    def get_f(self, buf):
        f = new_f(self.nx, self.ny, self.ld)
        self.ref.lib.ref_get_f(self.h, buf, _ptr(f))
        return f

    def set_f(self, buf, f):
        self.ref.lib.ref_set_f(self.h, buf, _ptr(np.asfortranarray(f)))


def have_ref() -> bool:
    return LIBREF.exists()
