"""The finite-volume / finite-difference streaming operators on the GPU against the
oracle (bit-exact in FP64), through the C ABI and the reference-named Python API, and
the reference's unchanged app/main_vortex.f90 on the engine."""
from __future__ import annotations

import json
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

import pyoracle as po

pytestmark = pytest.mark.gpu
OPS = ["fvm_bardow", "fdm_bardow", "fdm_sofonea"]
# stream_fdm_bardow's compile-time stencils (src/fvm_bardow.F90:591-669)
FDM_VARIANTS = ["fdm_bardow_wls", "fdm_bardow_wls_gauss_v1", "fdm_bardow_wls_gauss_v2",
                "fdm_bardow_iso"]


def fields(nx, ny, seed):
    rng = np.random.default_rng(seed)
    return (np.asfortranarray(1 + 0.01 * rng.standard_normal((ny, nx))),
            np.asfortranarray(0.05 * rng.standard_normal((ny, nx))),
            np.asfortranarray(0.05 * rng.standard_normal((ny, nx))))


# (1100, 1000) and (3300, 2500) run the stencil kernels with 8- and 32-column strips whose
# last strip is ragged; the small grids use single-column strips.
SIZES = [(40, 40, 1.0), (33, 70, 0.45), (130, 5, 0.8), (1100, 1000, 0.7), (3300, 2500, 1.0)]


@pytest.mark.parametrize("op", OPS + FDM_VARIANTS)
@pytest.mark.parametrize("nx,ny,dt", SIZES)
def test_streaming_operator_bit_exact(P, gpu, oracle, op, nx, ny, dt):
    print(f"case {op} {nx}x{ny} dt={dt}", flush=True)
    rho, ux, uy = fields(nx, ny, nx + ny)
    omega = 1.0 / (3 * 0.01 / dt + 0.5)
    lat = P.Lattice(nx, ny, 2, "f64")
    lat.set_params(omega, 0.25)
    lat.equilibrium_init(0, rho, ux, uy)
    a, b = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, a)
    for s in range(6):
        src, dst = s % 2, 1 - s % 2
        lat.stream_op(op, src, dst, dt)
        lat.collide("bgk", dst)
        oracle.stream_op(op, nx, ny, a, b, dt)
        oracle.collide("bgk", nx, ny, b, omega)
        a, b = b, a
    np.testing.assert_array_equal(lat.download_f(0)[:ny], a[:ny])


@pytest.mark.parametrize("model", ["bgk", "trt", "rr"])
@pytest.mark.parametrize("op", OPS + FDM_VARIANTS)
@pytest.mark.parametrize("nx,ny,dt", [SIZES[0], SIZES[2], SIZES[3], SIZES[4]])
def test_fused_operator_step_bit_exact(P, gpu, oracle, op, model, nx, ny, dt):
    """plbm_step_op (stream operator + collision in one pass) against the oracle's
    stream_op then collide, 3 steps, FP64."""
    print(f"case fused {op}+{model} {nx}x{ny} dt={dt}", flush=True)
    rho, ux, uy = fields(nx, ny, nx + 2 * ny)
    omega = 1.0 / (3 * 0.01 / dt + 0.5)
    lat = P.Lattice(nx, ny, 2, "f64")
    lat.set_params(omega, 0.25)
    lat.equilibrium_init(0, rho, ux, uy)
    a, b = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, a)
    for s in range(3):
        lat.step_op(op, model, s % 2, 1 - s % 2, dt)
        oracle.stream_op(op, nx, ny, a, b, dt)
        oracle.collide(model, nx, ny, b, omega, 0.25)
        a, b = b, a
    np.testing.assert_array_equal(lat.download_f(1)[:ny], a[:ny])


@pytest.mark.parametrize("op", OPS)
def test_fused_operator_step_shifted_storage(P, gpu, op):
    """FP32-shifted storage: the fused pass skips the intermediate FP32 rounding of the
    unfused pair, so both agree to FP32 roundoff."""
    nx, ny = 300, 200
    rho, ux, uy = fields(nx, ny, 5)
    out = []
    for fused in (True, False):
        lat = P.Lattice(nx, ny, 2, "f32s")
        lat.set_params(1.7, 0.25)
        lat.equilibrium_init(0, rho, ux, uy)
        for s in range(10):
            if fused:
                lat.step_op(op, "trt", s % 2, 1 - s % 2, 0.9)
            else:
                lat.stream_op(op, s % 2, 1 - s % 2, 0.9)
                lat.collide("trt", 1 - s % 2)
        out.append(lat.macros(0))
    for a, b in zip(out[0], out[1]):
        np.testing.assert_allclose(a, b, rtol=0, atol=2e-6)


@pytest.mark.parametrize("op", OPS)
def test_streaming_operator_shifted_storage(P, gpu, oracle, op):
    nx = ny = 48
    rho, ux, uy = fields(nx, ny, 3)
    ref = None
    for storage, tol in (("f64", 0.0), ("f64s", 1e-14), ("f32s", 1e-6)):
        lat = P.Lattice(nx, ny, 2, storage)
        lat.set_params(1.6, 0.25)
        lat.equilibrium_init(0, rho, ux, uy)
        for s in range(10):
            lat.stream_op(op, s % 2, 1 - s % 2, 0.7)
            lat.collide("trt", 1 - s % 2)
        _, u, v = lat.macros(0)
        if ref is None:
            ref = (u, v)
        err = np.sqrt(np.sum((u - ref[0]) ** 2 + (v - ref[1]) ** 2) / np.sum(ref[0] ** 2 + ref[1] ** 2))
        assert err <= tol, (storage, err)


@pytest.mark.parametrize("fused", [True, False])
def test_reference_named_api_with_fvm_streaming(P, gpu, oracle, fused):
    """perform_step with grid.streaming = stream_fvm_bardow (the fused device pass), and
    the same pair called one by one through the slots."""
    n, steps = 32, 12
    vc = P.VortexCase(0.05, n / 2, n / 2, n / 10, 0.1)
    rho, ux, uy = vc.eval(n, n)
    g = P.alloc_grid(n, n, 2, log=False)
    g.mf[:, :, 0], g.mf[:, :, 1], g.mf[:, :, 2] = rho, ux, uy
    P.set_properties(g, 1e-3, 0.8, 0.25)
    g.streaming, g.collision = P.stream_fvm_bardow, P.collide_bgk
    P.set_pdf_to_equilibrium(g)
    for _ in range(steps):
        if fused:
            P.perform_step(g)
        else:
            g.streaming(g)
            g.collision(g)
            g.iold, g.inew = g.inew, g.iold
    a, b = po.new_f(n, n), po.new_f(n, n)
    oracle.set_pdf_to_equilibrium(n, n, rho, ux, uy, a)
    for _ in range(steps):
        oracle.stream_op("fvm_bardow", n, n, a, b, 0.8)
        oracle.collide("bgk", n, n, b, g.omega)
        a, b = b, a
    np.testing.assert_array_equal(g.f[:n, :, :, g.iold - 1], a[:n])


def test_unchanged_vortex_app(gpu, tmp_path):
    """app/main_vortex.f90 (stream_fvm_bardow + collide_bgk through perform_step) compiled
    unmodified against the engine's modules: same printout (except the timing line) and
    byte-identical VTK / npy output as the reference build."""
    import hashlib
    exe = ROOT / "periodic-lbm_amd" / "_build" / "app" / "main_vortex"
    if not exe.exists():
        pytest.skip("app not built (the reference tree was absent at build time)")
    ref = json.loads((GOLDEN / "ref_vortex_app.json").read_text())
    r = subprocess.run([str(exe), *ref["args"]], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    strip = lambda s: [l.strip() for l in s.splitlines() if l.strip() and "MLUPS" not in l]
    assert strip(r.stdout) == strip(ref["stdout"])
    for name in ("results000000000.vtk", "results000000000.npy"):
        got = hashlib.sha256((tmp_path / "vortex" / name).read_bytes()).hexdigest()
        assert got == ref["files"][name], name


@pytest.mark.parametrize("per_pass", [2, 3])
@pytest.mark.parametrize("model", ["bgk", "trt", "rr"])
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("nx,ny,nsteps", [(200, 130, 7), (64, 64, 6), (5, 3, 5), (130, 1, 4)])
def test_operator_passes_equal_single_steps(P, gpu, op, model, per_pass, nx, ny, nsteps):
    """plbm_run_op in passes of 2 / 3 steps == single fused steps, both slots, FP64."""
    print(f"case {op}+{model} passes of {per_pass} {nx}x{ny} n={nsteps}", flush=True)
    rho, ux, uy = fields(nx, ny, nx + ny + nsteps)
    out = []
    for pp in (1, per_pass):
        lat = P.Lattice(nx, ny, 2, "f64")
        lat.set_params(1.0 / (3 * 0.01 / 0.8 + 0.5), 0.25)
        lat.set_op_steps_per_pass(pp)
        lat.equilibrium_init(0, rho, ux, uy)
        i, n = lat.run_op(op, model, nsteps, 0, 1, 0.8)
        out.append((i, n, lat.download_f(i), lat.download_f(n)))
    (i1, n1, a1, b1), (i2, n2, a2, b2) = out
    assert (i1, n1) == (i2, n2) == ((1, 0) if nsteps % 2 else (0, 1))
    np.testing.assert_array_equal(a1, a2)
    np.testing.assert_array_equal(b1, b2)


@pytest.mark.parametrize("lag_buffer", [True, False])
@pytest.mark.parametrize("per_pass", [2, 3])
def test_operator_pass_step_counts(P, gpu, per_pass, lag_buffer):
    """plbm_run_op for every step count 1..12, passes of 2 / 3 with the lagged state from
    the last pass (or a final single step), against single fused steps: both slots."""
    nx, ny = 96, 70
    print(f"case fd step counts per_pass={per_pass} lag_buffer={lag_buffer}", flush=True)
    for nsteps in range(1, 13):
        rho, ux, uy = fields(nx, ny, 3)
        out = []
        for pp in (1, per_pass):
            lat = P.Lattice(nx, ny, 2, "f64")
            lat.set_params(1.0 / (3 * 0.01 + 0.5), 0.25)
            lat.set_op_steps_per_pass(pp)
            lat.set_pass_lag_buffer(lag_buffer)
            lat.equilibrium_init(0, rho, ux, uy)
            i, n = lat.run_op("fdm_sofonea", "trt", nsteps, 0, 1, 1.0)
            out.append((i, n, lat.download_f(i), lat.download_f(n)))
            lat.close()
        (i1, n1, a1, b1), (i2, n2, a2, b2) = out
        assert (i1, n1) == (i2, n2), nsteps
        np.testing.assert_array_equal(a1, a2, err_msg=str(nsteps))
        np.testing.assert_array_equal(b1, b2, err_msg=str(nsteps))


def test_operator_passes_full_size(P, gpu, oracle):
    """4096^2, stream_fvm_bardow + BGK, 5 steps in passes (automatic depth) vs the oracle."""
    n, nsteps, dt = 4096, 5, 1.0
    print(f"case fvm passes {n}x{n}", flush=True)
    rho, ux, uy = fields(n, n, 77)
    omega = 1.0 / (3 * 0.01 + 0.5)
    lat = P.Lattice(n, n, 2, "f64")
    lat.set_params(omega, 0.25)
    assert lat.get_op_steps_per_pass() > 1
    lat.equilibrium_init(0, rho, ux, uy)
    i, _ = lat.run_op("fvm_bardow", "bgk", nsteps, 0, 1, dt)
    a, b = po.new_f(n, n), po.new_f(n, n)
    oracle.set_pdf_to_equilibrium(n, n, rho, ux, uy, a)
    for _ in range(nsteps):
        oracle.stream_op("fvm_bardow", n, n, a, b, dt)
        oracle.collide("bgk", n, n, b, omega)
        a, b = b, a
    np.testing.assert_array_equal(lat.download_f(i)[:n], a[:n])


@pytest.mark.parametrize("op", FDM_VARIANTS)
def test_fdm_stencil_variants_in_passes(P, gpu, op):
    """The -DFDM_* stencils through plbm_run_op in passes of 2 and 3 steps (with the
    lagged slot from the last pass) against single fused steps, both slots, FP64; and the
    reference-named Python API with the stencil chosen by set_fdm_bardow_stencil."""
    nx, ny, nsteps, dt = 200, 130, 7, 0.8
    print(f"case {op} passes {nx}x{ny}", flush=True)
    rho, ux, uy = fields(nx, ny, 41)
    out = []
    for pp in (1, 2, 3):
        lat = P.Lattice(nx, ny, 2, "f64")
        lat.set_params(1.0 / (3 * 0.01 / dt + 0.5), 0.25)
        lat.set_op_steps_per_pass(pp)
        lat.equilibrium_init(0, rho, ux, uy)
        i, n = lat.run_op(op, "trt", nsteps, 0, 1, dt)
        out.append((i, n, lat.download_f(i), lat.download_f(n)))
        lat.close()
    for i, n, a, b in out[1:]:
        assert (i, n) == out[0][:2]
        np.testing.assert_array_equal(a, out[0][2])
        np.testing.assert_array_equal(b, out[0][3])
    name = [k for k, v in P.FDM_BARDOW_STENCILS.items() if v == op][0]
    g = P.alloc_grid(nx, ny, 2, log=False)
    g.mf[:, :, 0], g.mf[:, :, 1], g.mf[:, :, 2] = rho, ux, uy
    P.set_properties(g, 0.01, dt, 0.25)
    g.streaming, g.collision = P.stream_fdm_bardow, P.collide_trt
    P.set_pdf_to_equilibrium(g)
    P.set_fdm_bardow_stencil(name)
    try:
        P.perform_steps(g, nsteps)
    finally:
        P.set_fdm_bardow_stencil("default")
    # the grid starts from the reference's ids (iold = 2, inew = 1): after an odd count
    # its latest state is in buffer 1 (0-based 0)
    assert (g.iold, g.inew) == (1, 2)
    P.sync_host(g)
    np.testing.assert_array_equal(g.f[:, :, :, g.iold - 1], out[0][2])
    np.testing.assert_array_equal(g.f[:, :, :, g.inew - 1], out[0][3])
    P.dealloc_grid(g)
