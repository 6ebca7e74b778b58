"""periodic-lbm_amd — D2Q9 periodic lattice Boltzmann engine for AMD Instinct MI355X.

Host-side Python mirror of the reference's module API (see lattice.py) over the HIP
engine in libplbm.so (C ABI: include/plbm.h).  The directory name contains a hyphen, so
load it with ``load_package()`` from ``__graft_entry__`` (or importlib) under the module
name ``periodic_lbm_amd``.
"""
from . import _abi
from ._abi import PlbmError
from .lattice import (
    Lattice, LatticeGrid, ShearLayer, TaylorGreen, add_noise, alloc_grid, collide_bgk,
    collide_rr, collide_trt, copy_bandwidth, csqr, cx, cy, dealloc_grid, equilibrium,
    lambda_d, lbm_stream, magic_number, noise, padded_ld, perform_lbm_step, perform_step,
    perform_triple_step, pi, set_pdf_to_equilibrium, set_properties, sync_device, sync_host,
    update_macros, w,
)

__all__ = [
    "Lattice", "LatticeGrid", "ShearLayer", "TaylorGreen", "PlbmError", "add_noise",
    "alloc_grid", "collide_bgk", "collide_rr", "collide_trt", "copy_bandwidth", "csqr", "cx",
    "cy", "dealloc_grid", "equilibrium", "lambda_d", "lbm_stream", "magic_number", "noise",
    "padded_ld", "perform_lbm_step", "perform_step", "perform_triple_step", "pi",
    "set_pdf_to_equilibrium", "set_properties", "sync_device", "sync_host", "update_macros",
    "w",
]
from .lattice import (  # noqa: E402
    VortexCase, dugks_collide, dugks_stream, perform_dugks_step, stream_fdm_bardow,
    stream_fdm_sofonea, stream_fvm_bardow, tau_over_dt,
)
from .batch import Batch, Case, sweep_cases  # noqa: E402

__all__ += ["VortexCase", "dugks_collide", "dugks_stream", "perform_dugks_step",
            "stream_fdm_bardow", "stream_fdm_sofonea", "stream_fvm_bardow", "tau_over_dt"]

__all__ += ["Batch", "Case", "sweep_cases"]
from .lattice import (  # noqa: E402
    FDM_BARDOW_STENCILS, collide_bgk_split, perform_dugks_steps, perform_steps,
    set_dugks_face_collisions, set_fdm_bardow_stencil,
)

__all__ += ["FDM_BARDOW_STENCILS", "collide_bgk_split", "perform_dugks_steps", "perform_steps",
            "set_dugks_face_collisions", "set_fdm_bardow_stencil"]
from .adjoint import Adjoint, band_limited_field, velocity_misfit  # noqa: E402

__all__ += ["Adjoint", "band_limited_field", "velocity_misfit"]
from . import sim  # noqa: E402

__all__ += ["sim"]
