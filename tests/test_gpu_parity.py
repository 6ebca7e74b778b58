"""GPU parity: the HIP engine (through the C ABI) against the oracle and the golden
fixtures.  FP64 storage must be bit-identical to the reference; shifted storages are
checked against the FP64 oracle with the tolerances stated in each test."""
from __future__ import annotations

import hashlib
import json

import numpy as np
import pytest

from conftest import GOLDEN, tgv_fields

import pyoracle as po

pytestmark = pytest.mark.gpu
MODELS = ["bgk", "trt", "rr"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(np.asfortranarray(a)).tobytes(order="F")).hexdigest()


def random_fields(nx, ny, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    rho = np.asfortranarray(1 + 0.01 * rng.standard_normal((ny, nx)))
    ux = np.asfortranarray(amp * rng.standard_normal((ny, nx)))
    uy = np.asfortranarray(amp * rng.standard_normal((ny, nx)))
    return rho, ux, uy


def oracle_run(oracle, model, nx, ny, rho, ux, uy, steps, omega, magic=0.25):
    fa, fb = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, fa)
    return oracle.run(model, nx, ny, fa, fb, steps, omega, magic)


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("nx,ny,steps", [(64, 64, 50), (37, 53, 20), (1, 5, 7), (130, 3, 9),
                                         (96, 128, 11)])
def test_fused_step_bit_exact(P, gpu, oracle, model, nx, ny, steps):
    print(f"case fused {model} {nx}x{ny} steps={steps}", flush=True)
    rho, ux, uy = random_fields(nx, ny, nx * 1000 + ny)
    omega = 1.0 / (3 * 0.02 + 0.5)
    lat = P.Lattice(nx, ny, 2, "f64")
    lat.set_params(omega, 0.25)
    lat.equilibrium_init(0, rho, ux, uy)
    iold, inew = lat.run(model, steps, 0, 1)
    lat.sync()
    latest, prev = oracle_run(oracle, model, nx, ny, rho, ux, uy, steps, omega)
    np.testing.assert_array_equal(lat.download_f(iold)[:ny], latest[:ny])
    np.testing.assert_array_equal(lat.download_f(inew)[:ny], prev[:ny])
    r, u, v = lat.macros(inew)
    for a, b in zip((r, u, v), oracle.update_macros(nx, ny, prev)):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("model", MODELS)
def test_config1_taylor_green_matches_reference_fingerprint(P, gpu, model):
    """64x64 TGV, 1000 steps: the GPU's buffers hash to the reference Fortran's."""
    C = json.loads((GOLDEN / "ref_config1.json").read_text())
    M = C["models"][model]
    n, steps = C["n"], C["steps"]
    rho, ux, uy = tgv_fields(P, n, C["u0"], C["nu"])
    g = P.alloc_grid(n, n, 2, log=False)
    g.mf[:, :, 0], g.mf[:, :, 1], g.mf[:, :, 2] = rho, ux, uy
    P.set_properties(g, C["nu"], 1.0, 0.25)
    assert g.omega == M["omega"]
    g.streaming = P.lbm_stream
    g.collision = {"bgk": P.collide_bgk, "trt": P.collide_trt, "rr": P.collide_rr}[model]
    P.set_pdf_to_equilibrium(g)
    for _ in range(steps):
        P.perform_lbm_step(g)
    f = g.f
    assert sha(f[:n, :, :, g.iold - 1]) == M["sha256_latest"]
    assert sha(f[:n, :, :, g.inew - 1]) == M["sha256_prev"]
    P.update_macros(g)
    assert sha(g.mf) == M["sha256_macros_prev"]
    k = 2 * np.pi / n
    _, uxa, uya = P.TaylorGreen(n, n, k, k, C["u0"], C["nu"]).eval(float(steps))
    l2 = np.sqrt(np.sum((g.ux - uxa) ** 2 + (g.uy - uya) ** 2) / np.sum(uxa ** 2 + uya ** 2))
    assert l2 == M["l2_lag"]
    P.dealloc_grid(g)


@pytest.mark.parametrize("model", MODELS)
def test_unfused_operators_bit_exact(P, gpu, oracle, model):
    nx, ny = 48, 40
    rho, ux, uy = random_fields(nx, ny, 5)
    omega = 1.7
    lat = P.Lattice(nx, ny, 2, "f64")
    lat.set_params(omega, 0.25)
    lat.equilibrium_init(1, rho, ux, uy)
    fa, fb = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, fa)
    for _ in range(3):
        lat.stream(1, 0)
        lat.collide(model, 0)
        lat.stream(0, 1)
        lat.collide(model, 1)
        for (s, d) in ((fa, fb), (fb, fa)):
            oracle.stream(nx, ny, s, d)
            oracle.collide(model, nx, ny, d, omega, 0.25)
    lat.sync()
    np.testing.assert_array_equal(lat.download_f(1)[:ny], fa[:ny])


def test_upload_download_roundtrip(P, gpu):
    nx, ny = 33, 70
    rng = np.random.default_rng(1)
    f = np.asfortranarray(rng.random((P.padded_ld(ny), nx, 9)))
    w = P.w
    for storage, tol in (("f64", 0.0), ("f64s", 2.3e-16), ("f32s", 1.2e-7 * 1.0),
                         ("f16s", 1.0 / 1024 / 1024 * 2)):
        lat = P.Lattice(nx, ny, 2, storage)
        g = np.asfortranarray(w + 1e-3 * (f - 0.5)) if storage != "f64" else f
        lat.upload_f(1, g)
        back = lat.download_f(1, ld=g.shape[0])
        err = np.abs(back[:ny] - g[:ny]).max()
        assert err <= tol, (storage, err)
        lat.close()


@pytest.mark.parametrize("nx,ny,ld", [(33, 70, 70), (33, 70, 80), (8, 3, 16), (1, 5, 9)])
def test_pitched_transfers_keep_padding(P, gpu, nx, ny, ld):
    """Host arrays with pitch ld >= ny (the reference's padded f(ld, nx, 0:8)): upload
    reads rows 0..ny-1 of each column, download writes them and leaves the padding rows
    ny..ld-1 as the caller had them, for ld == ny (one copy per plane) and ld > ny."""
    rng = np.random.default_rng(3)
    f = np.asfortranarray(rng.random((ld, nx, 9)))
    lat = P.Lattice(nx, ny, 2, "f64")
    lat.upload_f(0, f)
    out = np.full((ld, nx, 9), -7.0, order="F")
    lat.download_f(0, ld=ld, out=out)
    lat.close()
    np.testing.assert_array_equal(out[:ny], f[:ny])
    assert np.all(out[ny:] == -7.0)


# ragged shapes, a single row or column, and a grid larger than one sweep of the 1024
# reduction workgroups (several cells per thread)
@pytest.mark.parametrize("nx,ny", [(70, 45), (5, 3), (1, 7), (300, 1), (1000, 530)])
def test_reductions(P, gpu, oracle, nx, ny):
    print(f"case reductions {nx}x{ny}", flush=True)
    rho, ux, uy = random_fields(nx, ny, 11)
    lat = P.Lattice(nx, ny, 2, "f64")
    lat.set_params(1.5, 0.25)
    lat.equilibrium_init(0, rho, ux, uy)
    _, uxa, uya = random_fields(nx, ny, 12)
    out = lat.reduce(0, uxa, uya)
    f = po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, f)
    r, u, v = oracle.update_macros(nx, ny, f)
    R = P._abi
    assert out[R.RED_MASS] == pytest.approx(r.sum(), rel=1e-14)
    assert out[R.RED_KE] == pytest.approx(0.5 * np.sum(r * (u * u + v * v)), rel=1e-13)
    assert out[R.RED_U2] == pytest.approx(np.sum(u * u + v * v), rel=1e-13)
    assert out[R.RED_ERR2] == pytest.approx(np.sum((u - uxa) ** 2 + (v - uya) ** 2), rel=1e-13)
    assert out[R.RED_REF2] == pytest.approx(np.sum(uxa ** 2 + uya ** 2), rel=1e-13)
    assert out[R.RED_UMAX] == pytest.approx(np.max(np.hypot(u, v)), rel=1e-15)
    assert out[R.RED_NONFINITE] == 0
    again = lat.reduce(0, uxa, uya)
    np.testing.assert_array_equal(out, again)  # deterministic


def test_nonfinite_cells_are_counted(P, gpu):
    nx, ny = 16, 16
    lat = P.Lattice(nx, ny, 2, "f64")
    f = np.asfortranarray(np.tile(P.w, (16, nx, 1)).astype(np.float64))
    f[3, 4, 2] = np.nan
    f[7, 1, 0] = np.inf
    lat.upload_f(0, f)
    out = lat.reduce(0)
    assert out[P._abi.RED_NONFINITE] == 2


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("storage,tol", [("f64s", 1e-12), ("f32s", 1e-5), ("f16s", 2e-2)])
def test_shifted_storage_accuracy(P, gpu, oracle, model, storage, tol):
    """Velocity relative L2 error vs the FP64 oracle after 1000 steps of the 64x64 TGV.
    Tolerances: F64S 1e-12, F32S 1e-5, F16S 2e-2 (FP16 has an 11-bit significand)."""
    n, steps, nu = 64, 1000, 0.01
    rho, ux, uy = tgv_fields(P, n, 0.05, nu)
    omega = 1.0 / (3 * nu + 0.5)
    lat = P.Lattice(n, n, 2, storage)
    lat.set_params(omega, 0.25)
    lat.equilibrium_init(0, rho, ux, uy)
    iold, _ = lat.run(model, steps, 0, 1)
    r, u, v = lat.macros(iold)
    latest, _ = oracle_run(oracle, model, n, n, rho, ux, uy, steps, omega)
    r0, u0, v0 = oracle.update_macros(n, n, latest)
    err = np.sqrt(np.sum((u - u0) ** 2 + (v - v0) ** 2) / np.sum(u0 ** 2 + v0 ** 2))
    mass = abs(r.sum() - r0.sum()) / r0.sum()
    print(f"{storage} {model}: velocity rel L2 {err:.3e}, mass drift {mass:.3e}")
    assert err < tol
    assert mass < tol


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("per_pass", [1, 8])
def test_fp32_storage_conserves_mass_and_momentum(P, gpu, model, per_pass):
    """FP32-shifted storage, 128^2 random flow, 4000 steps: total mass and momentum stay
    within 1e-7 of their initial values (relative to the mass).  The FP32 weights are not
    representable; an operator that ignored that would drift by ~1e-8 per step (4e-5
    here), while the conserving FP32 collisions leave only unbiased rounding, whose random
    walk over 4000 steps is ~1e-9."""
    n, steps = 128, 4000
    print(f"case fp32 conservation {model} per_pass={per_pass}", flush=True)
    rho, ux, uy = random_fields(n, n, 21, amp=0.02)
    lat = P.Lattice(n, n, 2, "f32s")
    lat.set_params(1.0 / (3 * 0.05 + 0.5), 0.25)
    lat.set_steps_per_launch(model, per_pass)
    lat.equilibrium_init(0, rho, ux, uy)
    r0, u0, v0 = lat.macros(0)
    m0, px0, py0 = r0.sum(), (r0 * u0).sum(), (r0 * v0).sum()
    iold, _ = lat.run(model, steps, 0, 1)
    r, u, v = lat.macros(iold)
    dm = abs(r.sum() - m0) / m0
    dpx = abs((r * u).sum() - px0) / m0
    dpy = abs((r * v).sum() - py0) / m0
    print(f"f32s {model} per_pass={per_pass}: mass {dm:.2e} momentum {dpx:.2e} {dpy:.2e}")
    assert dm < 1e-7 and dpx < 1e-7 and dpy < 1e-7
    lat.close()


def test_fused_equals_unfused_at_4096(P, gpu):
    """Full-size property: one fused step equals lbm_stream + collide_trt bit for bit,
    and mass is conserved over 50 steps (4096^2, TRT, FP64)."""
    n = 4096
    rho, ux, uy = tgv_fields(P, n, 0.04, 0.005)
    P.add_noise(ux, uy, 1, 1e-3)
    omega = 1.0 / (3 * 0.005 + 0.5)
    lat = P.Lattice(n, n, 3, "f64")
    lat.set_params(omega, 0.25)
    lat.equilibrium_init(0, rho, ux, uy)
    m0 = lat.reduce(0)[P._abi.RED_MASS]
    lat.step("trt", 0, 1)
    lat.stream(0, 2)
    lat.collide("trt", 2)
    a = lat.reduce(1)
    b = lat.reduce(2)
    np.testing.assert_array_equal(a, b)
    iold, _ = lat.run("trt", 49, 1, 0)
    m1 = lat.reduce(iold)[P._abi.RED_MASS]
    assert abs(m1 - m0) / m0 < 1e-13


def test_errors_are_reported(P, gpu):
    lat = P.Lattice(8, 8, 2, "f64")
    with pytest.raises(P.PlbmError, match="parameters not set"):
        lat.step("bgk", 0, 1)
    lat.set_params(1.2, 0.25)
    with pytest.raises(P.PlbmError, match="out of range"):
        lat.step("bgk", 0, 2)
    with pytest.raises(P.PlbmError, match="distinct"):
        lat.step("bgk", 1, 1)
    with pytest.raises(P.PlbmError, match="unknown collision"):
        lat.step(9, 0, 1)


@pytest.mark.parametrize("block,remap", [(64, 0), (128, 1), (512, 0), (1024, 1)])
def test_launch_shapes_bit_exact(P, gpu, oracle, block, remap):
    """Every launch shape the autotuner may pick computes the same cells (ragged grid)."""
    nx, ny, steps = 45, 200, 6
    print(f"case launch block={block} remap={remap}", flush=True)
    rho, ux, uy = random_fields(nx, ny, 21)
    lat = P.Lattice(nx, ny, 2, "f64")
    lat.set_params(1.6, 0.25)
    lat.set_launch("rr", block, remap)
    assert lat.get_launch("rr") == (block, remap)
    lat.equilibrium_init(0, rho, ux, uy)
    iold, _ = lat.run("rr", steps, 0, 1)
    latest, _ = oracle_run(oracle, "rr", nx, ny, rho, ux, uy, steps, 1.6)
    np.testing.assert_array_equal(lat.download_f(iold)[:ny], latest[:ny])


def test_autotune_keeps_state_and_results(P, gpu, oracle):
    nx, ny = 300, 260
    rho, ux, uy = random_fields(nx, ny, 22)
    lat = P.Lattice(nx, ny, 2, "f64")
    lat.set_params(1.9, 0.25)
    lat.equilibrium_init(0, rho, ux, uy)
    before = lat.download_f(0)
    ms = lat.autotune("trt", 0, reps=2)
    assert ms > 0
    block, remap = lat.get_launch("trt")
    assert block % 64 == 0 and 64 <= block <= 1024 and remap in (0, 1)
    np.testing.assert_array_equal(lat.download_f(0), before)  # tuner wrote only scratch
    iold, _ = lat.run("trt", 5, 0, 1)
    latest, _ = oracle_run(oracle, "trt", nx, ny, rho, ux, uy, 5, 1.9)
    np.testing.assert_array_equal(lat.download_f(iold)[:ny], latest[:ny])
    with pytest.raises(P.PlbmError):
        lat.set_launch("trt", 96, 0)


@pytest.mark.parametrize("n,model,steps", [(4096, "trt", 20), (8192, "rr", 3), (8192, "trt", 3)])
def test_full_size_bit_exact(P, gpu, oracle, n, model, steps):
    """The benchmark sizes themselves: Taylor-Green + U(+-1e-3) noise (seed 1 at 4096^2 as
    the 4096^2 TRT roofline case, seed 7 at 8192^2 as the headline case), FP64, a few
    steps; GPU buffers equal the C oracle's (OpenMP on the host cores) bit for bit."""
    import os
    print(f"case full size {n} {model} {steps} steps", flush=True)
    nu, u0, seed = (0.005, 0.04, 1) if n == 4096 else (0.01, 0.05, 7)
    rho, ux, uy = tgv_fields(P, n, u0, nu)
    P.add_noise(ux, uy, seed, 1e-3)
    omega = 1.0 / (3 * nu + 0.5)
    lat = P.Lattice(n, n, 2, "f64")
    lat.set_params(omega, 0.25)
    lat.equilibrium_init(0, rho, ux, uy)
    iold, _ = lat.run(model, steps, 0, 1)
    got = lat.download_f(iold)
    lat.close()
    latest, prev = oracle_run(oracle, model, n, n, rho, ux, uy, steps, omega)
    del prev
    assert np.array_equal(got, latest), "GPU and oracle populations differ"


@pytest.mark.parametrize("model", ["trt", "bgk"])
def test_triple_step_matches_reference(P, gpu, model):
    """perform_triple_step (src/fvm_bardow.F90:322-340) through the reference-named Python
    API on the device, against the reference's three buffers and ids after 1 and 5 steps
    (ref_triple.npz; the reference's buffers start zeroed, as the device buffers here)."""
    G = np.load(GOLDEN / "ref_triple.npz")
    nx, ny = 16, 12
    nu, dt, magic = G["params"]
    g = P.alloc_grid(nx, ny, 3, log=False)
    zeros = np.zeros((P.padded_ld(ny), nx, 9), order="F")
    for b in range(3):
        g.lattice.upload_f(b, zeros)
    g.mf[:, :, 0], g.mf[:, :, 1], g.mf[:, :, 2] = G["rho"], G["ux"], G["uy"]
    P.set_properties(g, nu, dt, magic)
    g.streaming = P.lbm_stream
    g.collision = P.collide_trt if model == "trt" else P.collide_bgk
    P.set_pdf_to_equilibrium(g)
    done = 0
    for n in (1, 5):
        for _ in range(n - done):
            P.perform_triple_step(g)
        done = n
        assert [g.iold, g.inew, g.imid] == G[f"{model}_{n}_ids"].tolist()
        f = g.f
        for b in (1, 2, 3):
            np.testing.assert_array_equal(f[:ny, :, :, b - 1], G[f"{model}_{n}_buf{b}"],
                                          err_msg=f"step {n} buffer {b}")
    P.dealloc_grid(g)
