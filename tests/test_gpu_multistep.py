"""Temporal blocking of plbm_run (passes of 2..8 fused steps per launch, intermediate
states in registers): in FP64 storage bit-identical to the single-step kernel (and the
oracle) for every model, pass depth and grid shape (ragged strips and wavefronts, grids
smaller than one wavefront), with the reference's slot contents after the run (latest
state in iold, the state one step earlier in inew, which update_macros reads).  The
shifted storages' collisions allow FMA contraction, which the compiler places per
kernel, so there the two kernels agree to roundoff."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu

MODELS = ["bgk", "trt", "rr", "bgk_split"]


def fields(nx, ny, seed):
    rng = np.random.default_rng(seed)
    return (np.asfortranarray(1 + 0.01 * rng.standard_normal((ny, nx))),
            np.asfortranarray(0.05 * rng.standard_normal((ny, nx))),
            np.asfortranarray(0.05 * rng.standard_normal((ny, nx))))


def run(P, storage, model, ns, nx, ny, nsteps, seed=1):
    rho, ux, uy = fields(nx, ny, seed)
    lat = P.Lattice(nx, ny, 2, storage)
    lat.set_params(1.0 / (3 * 0.004 + 0.5), 0.25)
    lat.set_steps_per_launch(model, ns)
    assert lat.get_steps_per_launch(model) == ns
    lat.equilibrium_init(0, rho, ux, uy)
    iold, inew = lat.run(model, nsteps, 0, 1)
    return lat, iold, inew


# 200 x 130: ragged last strip (200 = 3*64 + 8) and ragged last wavefront rows;
# 5 x 3 and 1 x 7: the whole grid inside one wavefront, rows wrapping several times
@pytest.mark.parametrize("nx,ny", [(200, 130), (64, 64), (5, 3), (1, 7), (70, 1)])
@pytest.mark.parametrize("model", MODELS)
def test_passes_equal_single_steps_fp64(P, gpu, oracle, model, nx, ny):
    print(f"case multistep f64 {model} {nx}x{ny}", flush=True)
    nsteps = 11
    ref, ri, rn = run(P, "f64", model, 1, nx, ny, nsteps)
    want_new, want_prev = ref.download_f(ri), ref.download_f(rn)
    assert (ri, rn) == (1, 0)  # 11 steps from (0, 1): the reference's parity
    for ns in range(2, 9):
        lat, i, n = run(P, "f64", model, ns, nx, ny, nsteps)
        assert (i, n) == (ri, rn)
        np.testing.assert_array_equal(lat.download_f(i), want_new)
        np.testing.assert_array_equal(lat.download_f(n), want_prev)
        np.testing.assert_array_equal(np.stack(lat.macros(n)), np.stack(ref.macros(rn)))
    # and the single-step kernel is the oracle
    rho, ux, uy = fields(nx, ny, 1)
    fa, fb = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, fa)
    latest, prev = oracle.run(model, nx, ny, fa, fb, nsteps, 1.0 / (3 * 0.004 + 0.5), 0.25)
    np.testing.assert_array_equal(want_new[:ny], latest[:ny])


# velocity tolerance after 10 steps: FP64 roundoff, FP32 roundoff (FMA placement), and
# for FP16 storage the FP16 roundings (scale 2^10) that single steps make of the states
# a pass keeps in FP32 registers
SHIFTED_TOL = {"f64s": 1e-14, "f32s": 1e-6, "f16s": 2e-4}


@pytest.mark.parametrize("storage", ["f64s", "f32s", "f16s"])
@pytest.mark.parametrize("model", ["bgk", "trt", "rr"])
def test_passes_match_single_steps_shifted(P, gpu, storage, model):
    nx, ny, nsteps = 150, 100, 10
    ref, ri, rn = run(P, storage, model, 1, nx, ny, nsteps)
    r0, u0, v0 = ref.macros(ri)
    for ns in range(2, 9):
        lat, i, n = run(P, storage, model, ns, nx, ny, nsteps)
        assert (i, n) == (ri, rn)
        r, u, v = lat.macros(i)
        tol = SHIFTED_TOL[storage]
        np.testing.assert_allclose(u, u0, rtol=0, atol=tol)
        np.testing.assert_allclose(v, v0, rtol=0, atol=tol)
        np.testing.assert_allclose(r, r0, rtol=0, atol=tol)


@pytest.mark.parametrize("lag_buffer", [True, False])
@pytest.mark.parametrize("ns", [2, 3, 4, 8])
def test_step_counts_and_slots(P, gpu, ns, lag_buffer):
    """Every step count 1..20 (one pass, several, uneven splits, a single step) keeps the
    reference's slot contents, with the last pass writing the lagged state into the third
    buffer or (lag buffer off) a final single step; the slots' device storage may move.
    A second run continues from the returned ids."""
    nx, ny = 96, 80
    print(f"case step counts ns={ns} lag_buffer={lag_buffer}", flush=True)
    for nsteps in range(1, 21):
        ref, ri, rn = run(P, "f64", "trt", 1, nx, ny, nsteps)
        rho, ux, uy = fields(nx, ny, 1)
        lat = P.Lattice(nx, ny, 2, "f64")
        lat.set_params(1.0 / (3 * 0.004 + 0.5), 0.25)
        lat.set_steps_per_launch("trt", ns)
        lat.set_pass_lag_buffer(lag_buffer)
        lat.equilibrium_init(0, rho, ux, uy)
        i, n = lat.run("trt", nsteps, 0, 1)
        assert (i, n) == (ri, rn) == ((1, 0) if nsteps % 2 else (0, 1)), nsteps
        np.testing.assert_array_equal(lat.download_f(i), ref.download_f(ri), err_msg=str(nsteps))
        np.testing.assert_array_equal(lat.download_f(n), ref.download_f(rn), err_msg=str(nsteps))
        i2, n2 = lat.run("trt", 7, i, n)
        r2, q2 = ref.run("trt", 7, ri, rn)
        assert (i2, n2) == (r2, q2)
        np.testing.assert_array_equal(lat.download_f(i2), ref.download_f(r2))
        np.testing.assert_array_equal(lat.download_f(n2), ref.download_f(q2))
        lat.close()
        ref.close()


@pytest.mark.parametrize("storage", ["f32s", "f16s"])
def test_lag_slot_shifted_storage(P, gpu, storage):
    """4- and 2-byte storage: the lagged slot written by the last pass holds the state one
    step before the latest (as the run with the lag buffer off, which ends on a single
    step, to the storage's roundoff)."""
    nx, ny, nsteps = 150, 100, 13
    out = []
    for lag_buffer in (True, False):
        rho, ux, uy = fields(nx, ny, 4)
        lat = P.Lattice(nx, ny, 2, storage)
        lat.set_params(1.0 / (3 * 0.004 + 0.5), 0.25)
        lat.set_steps_per_launch("trt", 8)
        lat.set_pass_lag_buffer(lag_buffer)
        lat.equilibrium_init(0, rho, ux, uy)
        i, n = lat.run("trt", nsteps, 0, 1)
        out.append(((i, n), lat.macros(i), lat.macros(n)))
    assert out[0][0] == out[1][0]
    tol = SHIFTED_TOL[storage]
    for a, b in zip(out[0][1] + out[0][2], out[1][1] + out[1][2]):
        np.testing.assert_allclose(a, b, rtol=0, atol=tol)
    # the two slots differ by one step
    assert np.abs(out[0][1][1] - out[0][2][1]).max() > 10 * tol


def test_full_size_passes_bit_exact_4096(P, gpu, oracle):
    """4096^2 TRT, 9 steps in passes of 8 (the automatic choice at this size) against
    the oracle."""
    n, nsteps = 4096, 9
    print(f"case multistep 4096 trt {nsteps}", flush=True)
    rho, ux, uy = fields(n, n, 5)
    lat = P.Lattice(n, n, 2, "f64")
    omega = 1.0 / (3 * 0.01 + 0.5)
    lat.set_params(omega, 0.25)
    assert lat.get_steps_per_launch("trt") == 8
    lat.equilibrium_init(0, rho, ux, uy)
    i, n_ = lat.run("trt", nsteps, 0, 1)
    fa, fb = po.new_f(n, n), po.new_f(n, n)
    oracle.set_pdf_to_equilibrium(n, n, rho, ux, uy, fa)
    latest, prev = oracle.run("trt", n, n, fa, fb, nsteps, omega, 0.25)
    np.testing.assert_array_equal(lat.download_f(i)[:n], latest[:n])
    np.testing.assert_array_equal(lat.download_f(n_)[:n], prev[:n])


@pytest.mark.parametrize("segments", [1, 3, 7])
@pytest.mark.parametrize("nx,ny", [(200, 130), (5, 3), (70, 1)])
def test_pass_segments_do_not_change_results(P, gpu, nx, ny, segments):
    """The work split of a pass (whole-band walks, ragged cuts, one-column segments where
    segments > nx) leaves FP64 results bit-identical to single steps."""
    print(f"case segments {segments} {nx}x{ny}", flush=True)
    for model in ("trt", "rr"):
        ref, ri, rn = run(P, "f64", model, 1, nx, ny, 9)
        rho, ux, uy = fields(nx, ny, 1)
        lat = P.Lattice(nx, ny, 2, "f64")
        lat.set_params(1.0 / (3 * 0.004 + 0.5), 0.25)
        lat.set_steps_per_launch(model, 4)
        lat.set_pass_segments(segments)
        assert lat.get_pass_segments() == segments
        lat.equilibrium_init(0, rho, ux, uy)
        i, n = lat.run(model, 9, 0, 1)
        assert (i, n) == (ri, rn)
        np.testing.assert_array_equal(lat.download_f(i), ref.download_f(ri))
        np.testing.assert_array_equal(lat.download_f(n), ref.download_f(rn))
    with pytest.raises(P.PlbmError):
        lat.set_pass_segments(-1)


@pytest.mark.parametrize("nx,ny", [(8192, 4096), (1024, 1024), (1000, 530)])
def test_automatic_work_split(P, gpu, nx, ny):
    """FP64 TRT, 5 steps, the automatic depth (passes of 8) and work split (64 segments per
    band: 128 columns at 8192, 16 at 1024, 15-16 at 1000) against single steps."""
    nsteps = 5
    print(f"case segments auto {nx}x{ny}", flush=True)
    rho, ux, uy = fields(nx, ny, 9)
    out = []
    for ns in (1, 0):
        lat = P.Lattice(nx, ny, 2, "f64")
        lat.set_params(1.0 / (3 * 0.01 + 0.5), 0.25)
        lat.set_steps_per_launch("trt", ns)
        assert lat.get_steps_per_launch("trt") == (ns or 8)
        lat.equilibrium_init(0, rho, ux, uy)
        i, n = lat.run("trt", nsteps, 0, 1)
        out.append((i, lat.download_f(i)))
        lat.close()
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_autotune_picks_a_depth(P, gpu):
    """The autotuner picks a pass depth and, separately, the depth of the pass that ends
    a run (it stores the lagged state); a run under that choice equals single steps bit
    for bit, and setting a depth afterwards applies it to the last pass too."""
    n = 2048
    lat = P.Lattice(n, n, 2, "f64")
    lat.set_params(1.9, 0.25)
    lat.equilibrium_init(0, *fields(n, n, 2))
    ms = lat.autotune("trt", 0, reps=3)
    ns = lat.get_steps_per_launch("trt")
    assert ms > 0 and ns in range(1, 9)
    ref = P.Lattice(n, n, 2, "f64")
    ref.set_params(1.9, 0.25)
    ref.equilibrium_init(0, *fields(n, n, 2))
    ref.set_steps_per_launch("trt", 1)
    ri, rn = ref.run("trt", 20, 0, 1)
    lat.set_run_timing(True)
    i, n_ = lat.run("trt", 20, 0, 1)
    depths = [d for d, _ in lat.run_launch_times()]
    print("autotuned depth", ns, "run of 20 steps", depths, flush=True)
    assert sum(depths) == 20 and (i, n_) == (ri, rn)
    if ns > 1:
        assert depths[-1] >= ns and all(d <= ns for d in depths[:-1])
    np.testing.assert_array_equal(lat.download_f(i), ref.download_f(ri))
    np.testing.assert_array_equal(lat.download_f(n_), ref.download_f(rn))
    lat.set_steps_per_launch("trt", 4)
    lat.run("trt", 20, 0, 1)
    assert [d for d, _ in lat.run_launch_times()] == [4] * 5
    lat.set_steps_per_launch("trt", 8)
    with pytest.raises(P.PlbmError):
        lat.set_steps_per_launch("trt", 9)  # passes hold at most 8 steps
    lat32 = P.Lattice(64, 64, 2, "f32s")
    lat32.set_steps_per_launch("trt", 8)
    with pytest.raises(P.PlbmError):
        lat32.set_steps_per_launch("trt", 9)


@pytest.mark.parametrize("nsteps", [9, 10])
def test_reference_named_perform_steps(P, gpu, nsteps):
    """perform_steps(grid, n) == n x perform_step (both slots, ids) through the
    reference-named API, in passes of 4 on a 300 x 200 grid."""
    nx, ny = 300, 200
    rho, ux, uy = fields(nx, ny, 8)
    grids = []
    for _ in range(2):
        g = P.alloc_grid(nx, ny, 2, log=False)
        g.mf[:, :, 0], g.mf[:, :, 1], g.mf[:, :, 2] = rho, ux, uy
        P.set_properties(g, 0.01, 1.0, 0.25)
        g.streaming, g.collision = P.lbm_stream, P.collide_rr
        P.set_pdf_to_equilibrium(g)
        grids.append(g)
    grids[0].lattice.set_steps_per_launch("rr", 4)
    P.perform_steps(grids[0], nsteps)
    for _ in range(nsteps):
        P.perform_step(grids[1])
    a, b = grids
    assert (a.iold, a.inew) == (b.iold, b.inew)
    P.sync_host(a)
    P.sync_host(b)
    np.testing.assert_array_equal(a.f[:ny], b.f[:ny])


def test_pass_and_run_errors(P, gpu):
    lat = P.Lattice(64, 64, 2, "f64")
    with pytest.raises(P.PlbmError, match="parameters not set"):
        lat.run_pass("trt", 2, 0, 1)
    lat.set_params(1.5, 0.25)
    with pytest.raises(P.PlbmError, match="distinct"):
        lat.run_pass("trt", 2, 1, 1)
    with pytest.raises(P.PlbmError, match="steps per pass"):
        lat.run_pass("trt", 9, 0, 1)
    with pytest.raises(P.PlbmError, match="unknown collision"):
        lat.run_pass(7, 2, 0, 1)
    with pytest.raises(P.PlbmError, match="unknown collision"):
        lat.run(7, 4, 0, 1)
    with pytest.raises(P.PlbmError, match="nsteps"):
        lat.run("trt", -1, 0, 1)
    with pytest.raises(P.PlbmError):
        lat.step_op("fvm_bardow", "trt", 0, 1, float("nan"))
    with pytest.raises(P.PlbmError):
        lat.step_op(9, "trt", 0, 1, 1.0)
    b = P.Batch([(32, 32)], "f64", "bgk")
    with pytest.raises(P.PlbmError, match="steps per pass"):
        b.set_steps_per_pass(9)
    # a valid call after the errors still works and equals single steps
    lat.equilibrium_init(0, *fields(64, 64, 3))
    lat.run_pass("trt", 2, 0, 1)
    ref = P.Lattice(64, 64, 2, "f64")
    ref.set_params(1.5, 0.25)
    ref.equilibrium_init(0, *fields(64, 64, 3))
    ref.set_steps_per_launch("trt", 1)
    i, _ = ref.run("trt", 2, 0, 1)
    np.testing.assert_array_equal(lat.download_f(1), ref.download_f(i))
    # zero steps: slots and contents unchanged, at the automatic and a forced depth
    for depth in (0, 8):
        lat.set_steps_per_launch("trt", depth)
        before = lat.download_f(1)
        assert lat.run("trt", 0, 1, 0) == (1, 0)
        np.testing.assert_array_equal(lat.download_f(1), before)
