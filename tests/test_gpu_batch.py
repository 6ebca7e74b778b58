"""Batched engine on the GPU: every case equals the single-lattice engine bit for bit
(same per-cell arithmetic), FP64 cases equal the oracle bit for bit, and the graph,
direct and resident (persistent) paths agree."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu


def small_cases(P, seed=11, n=10):
    return P.sweep_cases(n, seed=seed, sizes=(16, 24, 32, 48, 64))


@pytest.mark.parametrize("storage,model", [("f32s", "bgk"), ("f64", "trt"), ("f16s", "rr"),
                                           ("f64", "rr")])
def test_batch_equals_single_lattice(P, gpu, storage, model):
    cases = small_cases(P)
    steps = 30
    print(f"case batch {storage} {model}: {[c.n for c in cases]}", flush=True)
    b = P.Batch.from_cases(cases, storage, model)
    b.run(steps)
    b.sync()
    for i, c in enumerate(cases):
        lat = P.Lattice(c.n, c.n, 2, storage)
        lat.set_params(c.omega, 0.25)
        lat.equilibrium_init(0, *c.fields())
        iold, _ = lat.run(model, steps, 0, 1)
        np.testing.assert_array_equal(b.download_f(i), lat.download_f(iold))
        lat.close()


def test_batch_fp64_equals_oracle(P, gpu, oracle):
    cases = small_cases(P, seed=3, n=6)
    steps = 25
    b = P.Batch.from_cases(cases, "f64", "bgk")
    b.run(steps)
    for i, c in enumerate(cases):
        fa, fb = po.new_f(c.n, c.n), po.new_f(c.n, c.n)
        rho, ux, uy = c.fields()
        oracle.set_pdf_to_equilibrium(c.n, c.n, rho, ux, uy, fa)
        latest, _ = oracle.run("bgk", c.n, c.n, fa, fb, steps, c.omega)
        np.testing.assert_array_equal(b.download_f(i)[:c.n], latest[:c.n])


@pytest.mark.parametrize("ty,tx,res,rb", [(64, 1, 0, 256), (32, 8, 0, 256), (128, 2, 4096, 256),
                                          (64, 4, 1 << 20, 1024), (256, 1, 2304, 64)])
def test_launch_plans_agree(P, gpu, ty, tx, res, rb):
    cases = small_cases(P, seed=5, n=9)
    ref = P.Batch.from_cases(cases, "f32s", "trt")
    ref.set_launch(128, 1, 0, 256)
    ref.run(21, graph=False)
    b = P.Batch.from_cases(cases, "f32s", "trt")
    b.set_launch(ty, tx, res, rb)
    assert b.get_launch() == (ty, tx, res, rb)
    b.run(21, graph=True)
    for i in range(len(cases)):
        np.testing.assert_array_equal(b.download_f(i), ref.download_f(i))


def test_batch_autotune_and_reductions(P, gpu):
    cases = small_cases(P, seed=8, n=12)
    b = P.Batch.from_cases(cases, "f32s", "bgk")
    before = [b.download_f(i) for i in range(len(cases))]
    ms = b.autotune(4)
    assert ms > 0
    for i in range(len(cases)):  # the tuner ran on a scratch arena
        np.testing.assert_array_equal(b.download_f(i), before[i])
    red = b.reduce()
    for i, c in enumerate(cases):
        r, u, v = b.macros(i)
        assert red[i, P._abi.RED_MASS] == pytest.approx(r.sum(), rel=1e-12)
        assert red[i, P._abi.RED_UMAX] == pytest.approx(np.hypot(u, v).max(), rel=1e-12)
        assert red[i, P._abi.RED_KE] == pytest.approx(0.5 * np.sum(r * (u * u + v * v)),
                                                      rel=1e-12)
        assert red[i, P._abi.RED_U2] == pytest.approx(np.sum(u * u + v * v), rel=1e-12)
        assert red[i, P._abi.RED_NONFINITE] == 0
    np.testing.assert_array_equal(b.reduce(), red)  # deterministic
    with pytest.raises(P.PlbmError):
        b.set_launch(48, 1, 0, 256)


@pytest.mark.parametrize("storage,model,ns,xs", [("f64", "trt", 4, 64), ("f64", "rr", 3, 128),
                                                 ("f64", "bgk_split", 2, 64), ("f64", "trt", 6, 256),
                                                 ("f64", "bgk", 8, 37), ("f64", "trt", 5, 1),
                                                 ("f32s", "bgk", 8, 128),
                                                 ("f16s", "trt", 6, 64)])
@pytest.mark.parametrize("nsteps", [7, 30])
def test_batch_passes(P, gpu, storage, model, ns, xs, nsteps):
    check_batch_passes(P, storage, model, ns, xs, nsteps)


@pytest.mark.parametrize("ns", [5, 6, 7])
def test_batch_passes_over_several_chunks(P, gpu, ns):
    """250 steps: graph chunks of 2 ns x k steps (100, 96, 98) and a shorter last chunk,
    each split into passes of nearly equal depth; FP64 TRT equals the single lattice."""
    check_batch_passes(P, "f64", "trt", ns, 64, 250)


def check_batch_passes(P, storage, model, ns, xs, nsteps):
    """Temporally blocked passes for the tiled cases: FP64 cases equal the single lattice
    bit for bit, shifted storages to roundoff; the resident cases keep stepping one step
    at a time in parallel, and the two agree on the buffer holding the latest state."""
    cases = P.sweep_cases(12, seed=13, sizes=(16, 48, 64, 96, 130))
    print(f"case batch passes {storage} {model} ns={ns} xs={xs} n={nsteps}", flush=True)
    b = P.Batch.from_cases(cases, storage, model)
    b.set_launch(128, 1, 2400, 256)  # cases up to 48^2 resident, the rest in passes
    b.set_steps_per_pass(ns)
    assert b.get_steps_per_pass() == ns
    b.set_pass_columns(xs)
    assert b.get_pass_columns() == xs
    b.run(nsteps)
    b.sync()
    tol = {"f64": 0.0, "f32s": 1e-6, "f16s": 2e-4}[storage]
    for i, c in enumerate(cases):
        lat = P.Lattice(c.n, c.n, 2, storage)
        lat.set_params(c.omega, 0.25)
        lat.set_steps_per_launch(model, 1)
        lat.equilibrium_init(0, *c.fields())
        iold, _ = lat.run(model, nsteps, 0, 1)
        if tol == 0.0:
            np.testing.assert_array_equal(b.download_f(i), lat.download_f(iold))
        else:
            for got, want in zip(b.macros(i), lat.macros(iold)):
                np.testing.assert_allclose(got, want, rtol=0, atol=tol)
