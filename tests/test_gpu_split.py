"""collide_bgk of the reference's -DSPLIT build (bgk_kernel_cache, src/collision_bgk.F90:
84-176) as collision model PLBM_BGK_SPLIT on the GPU: bit-exact to that build's own run
(ref_split.npz) and to the oracle, through the fused step, the unfused operators, the
fused FV/FD step, the batched engine and the reference-named Python API."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import GOLDEN

import pyoracle as po

pytestmark = pytest.mark.gpu


def fields(nx, ny, seed):
    rng = np.random.default_rng(seed)
    return (np.asfortranarray(1 + 0.01 * rng.standard_normal((ny, nx))),
            np.asfortranarray(0.05 * rng.standard_normal((ny, nx))),
            np.asfortranarray(0.05 * rng.standard_normal((ny, nx))))


@pytest.mark.parametrize("nx,ny", [(16, 16), (12, 20)])
def test_split_bgk_golden(P, gpu, nx, ny):
    G = np.load(GOLDEN / "ref_split.npz")
    key = f"{nx}x{ny}"
    rho, ux, uy = (np.asfortranarray(G[f"{key}_{c}"]) for c in ("rho", "ux", "uy"))
    lat = P.Lattice(nx, ny, 2, "f64")
    lat.set_params(float(G[f"{key}_omega"][0]), 0.25)
    lat.equilibrium_init(0, rho, ux, uy)
    iold, inew, done = 0, 1, 0
    for n in (1, 10, 100):
        iold, inew = lat.run("bgk_split", n - done, iold, inew)
        done = n
        np.testing.assert_array_equal(lat.download_f(iold)[:ny], G[f"{key}_bgk_{n}_latest"])


@pytest.mark.parametrize("nx,ny", [(300, 257), (1024, 768)])
def test_split_bgk_fused_and_unfused_vs_oracle(P, gpu, oracle, nx, ny):
    print(f"case bgk_split {nx}x{ny}", flush=True)
    rho, ux, uy = fields(nx, ny, nx + ny)
    omega = 1.0 / (3 * 0.004 + 0.5)
    fa, fb = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, fa)
    want, _ = oracle.run("bgk_split", nx, ny, fa, fb, 12, omega, 0.25)
    lat = P.Lattice(nx, ny, 2, "f64")
    lat.set_params(omega, 0.25)
    lat.equilibrium_init(0, rho, ux, uy)
    iold, _ = lat.run("bgk_split", 12, 0, 1)
    np.testing.assert_array_equal(lat.download_f(iold)[:ny], want[:ny])
    lat.equilibrium_init(0, rho, ux, uy)
    for s in range(12):
        lat.stream(s % 2, 1 - s % 2)
        lat.collide("bgk_split", 1 - s % 2)
    np.testing.assert_array_equal(lat.download_f(0)[:ny], want[:ny])


def test_split_bgk_with_fv_streaming(P, gpu, oracle):
    nx, ny, dt = 1100, 1000, 0.8
    rho, ux, uy = fields(nx, ny, 2)
    omega = 1.0 / (3 * 0.01 / dt + 0.5)
    lat = P.Lattice(nx, ny, 2, "f64")
    lat.set_params(omega, 0.25)
    lat.equilibrium_init(0, rho, ux, uy)
    a, b = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, a)
    for s in range(3):
        lat.step_op("fvm_bardow", "bgk_split", s % 2, 1 - s % 2, dt)
        oracle.stream_op("fvm_bardow", nx, ny, a, b, dt)
        oracle.collide("bgk_split", nx, ny, b, omega)
        a, b = b, a
    np.testing.assert_array_equal(lat.download_f(1)[:ny], a[:ny])


def test_split_bgk_batch_equals_single_lattice(P, gpu):
    cases = P.sweep_cases(8, seed=9, sizes=(16, 32, 48, 64))
    b = P.Batch.from_cases(cases, "f64", "bgk_split")
    b.run(20)
    b.sync()
    for i, c in enumerate(cases):
        lat = P.Lattice(c.n, c.n, 2, "f64")
        lat.set_params(c.omega, 0.25)
        lat.equilibrium_init(0, *c.fields())
        iold, _ = lat.run("bgk_split", 20, 0, 1)
        np.testing.assert_array_equal(b.download_f(i), lat.download_f(iold))


def test_split_bgk_reference_named_api(P, gpu, oracle):
    """grid.collision = collide_bgk_split through perform_step (fused) and autotune."""
    n = 96
    rho, ux, uy = fields(n, n, 4)
    g = P.alloc_grid(n, n, 2, log=False)
    g.mf[:, :, 0], g.mf[:, :, 1], g.mf[:, :, 2] = rho, ux, uy
    P.set_properties(g, 0.02, 1.0, 0.25)
    g.streaming, g.collision = P.lbm_stream, P.collide_bgk_split
    g.lattice.autotune("bgk_split", g.iold - 1, reps=2)
    P.set_pdf_to_equilibrium(g)
    for _ in range(15):
        P.perform_step(g)
    fa, fb = po.new_f(n, n), po.new_f(n, n)
    oracle.set_pdf_to_equilibrium(n, n, rho, ux, uy, fa)
    want, _ = oracle.run("bgk_split", n, n, fa, fb, 15, g.omega, 0.25)
    P.sync_host(g)
    np.testing.assert_array_equal(g.f[:n, :, :, g.iold - 1], want[:n])
