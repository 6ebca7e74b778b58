"""The oracle (C restatement) against the golden fixtures written by the reference
Fortran itself (tests/golden/make_golden.py).  Everything here is bit-exact."""
from __future__ import annotations

import hashlib
import json

import numpy as np
import pytest

from conftest import GOLDEN

import pyoracle as po

MODELS = ["bgk", "trt", "rr"]
SMALL = [(8, 8), (5, 3), (16, 16), (12, 20)]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(np.asfortranarray(a)).tobytes(order="F")).hexdigest()


@pytest.mark.parametrize("nx,ny", SMALL)
@pytest.mark.parametrize("model", MODELS)
def test_small_grids_bitwise(oracle, golden_small, nx, ny, model):
    G = golden_small
    key = f"{nx}x{ny}"
    rho, ux, uy = (np.asfortranarray(G[f"{key}_{c}"]) for c in ("rho", "ux", "uy"))
    nu, dt, magic = G[f"{key}_params"]
    par = oracle.set_properties(nu, dt, magic)
    fa, fb = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, fa)
    a, b, done = fa, fb, 0
    for n in (1, 10, 100):
        latest, prev = oracle.run(model, nx, ny, a, b, n - done, par["omega"], par["trt_magic"])
        done = n
        a, b = latest, prev
        np.testing.assert_array_equal(latest[:ny], G[f"{key}_{model}_{n}_latest"])
        r, u, v = oracle.update_macros(nx, ny, prev)
        np.testing.assert_array_equal(np.stack([r, u, v], axis=2), G[f"{key}_{model}_{n}_macros"])
        # continue from the latest state: swap so that the next run starts from it
        a, b = latest, prev


@pytest.mark.parametrize("model", MODELS)
def test_config1_taylor_green_bitwise(oracle, model):
    """64x64 TGV, u0=0.05, nu=0.01, 1000 steps: the reference's final buffers, bit for bit."""
    C = json.loads((GOLDEN / "ref_config1.json").read_text())
    n, u0, nu, steps = C["n"], C["u0"], C["nu"], C["steps"]
    M = C["models"][model]
    k = 2 * np.pi / n
    p, ux, uy = oracle.tgv(n, n, k, k, u0, nu, 0.0)
    rho = np.asfortranarray(p / (1.0 / 3.0) + 1.0)
    par = oracle.set_properties(nu, 1.0, 0.25)
    assert par["omega"] == M["omega"]
    fa, fb = po.new_f(n, n), po.new_f(n, n)
    oracle.set_pdf_to_equilibrium(n, n, rho, ux, uy, fa)
    latest, prev = oracle.run(model, n, n, fa, fb, steps, par["omega"], par["trt_magic"])
    assert sha(latest[:n]) == M["sha256_latest"]
    assert sha(prev[:n]) == M["sha256_prev"]
    r, u, v = oracle.update_macros(n, n, prev)
    assert sha(np.stack([r, u, v], axis=2)) == M["sha256_macros_prev"]
    _, uxa, uya = oracle.tgv(n, n, k, k, u0, nu, float(steps))
    l2 = np.sqrt(np.sum((u - uxa) ** 2 + (v - uya) ** 2) / np.sum(uxa ** 2 + uya ** 2))
    assert l2 == M["l2_lag"]


def test_params_and_equilibrium(oracle):
    G = json.loads((GOLDEN / "ref_params.json").read_text())
    for rec in G["set_properties"]:
        par = oracle.set_properties(rec["nu"], rec["dt"], rec["magic"])
        for key in ("tau", "omega", "trt_magic", "csqr"):
            assert par[key] == rec[key], (rec, key)
    for om, x, v in G["lambda_d"]:
        assert oracle.lambda_d(om, x) == v
    for le, ld, v in G["magic_number"]:
        assert oracle.magic_number(le, ld) == v
    for rho, ux, uy, feq in G["equilibrium"]:
        np.testing.assert_array_equal(oracle.equilibrium(rho, ux, uy), np.array(feq))


def test_reference_app_record():
    """The unchanged app's printed L2 (app/main_taylor_green.f90:125-126) as recorded."""
    A = json.loads((GOLDEN / "ref_app.json").read_text())
    assert A["values"]["L2-norm"] == pytest.approx(3.7266899791616004e-05, rel=0, abs=0)
    assert A["values"]["Final time"] == 48658.0


@pytest.mark.parametrize("op", ["fvm_bardow", "fdm_bardow", "fdm_sofonea"])
@pytest.mark.parametrize("dt", [1.0, 0.5])
def test_streaming_operators_bitwise(oracle, op, dt):
    """perform_step with stream_fvm_bardow / stream_fdm_bardow / stream_fdm_sofonea and
    collide_bgk, 5 steps on 16x16 (reference fixtures)."""
    G = np.load(GOLDEN / "ref_streaming_ops.npz")
    n = 16
    rho, ux, uy = (np.asfortranarray(G[k]) for k in ("rho", "ux", "uy"))
    omega = float(G[f"{op}_dt{dt}_omega"][0])
    assert omega == oracle.set_properties(0.02, dt, 0.25)["omega"]
    a, b = po.new_f(n, n), po.new_f(n, n)
    oracle.set_pdf_to_equilibrium(n, n, rho, ux, uy, a)
    for _ in range(5):
        oracle.stream_op(op, n, n, a, b, dt)
        oracle.collide("bgk", n, n, b, omega)
        a, b = b, a
    np.testing.assert_array_equal(a[:n], G[f"{op}_dt{dt}"])


DUGKS_CASES = [(16, 16, 1.0), (12, 20, 0.5), (24, 8, 2.0)]


@pytest.mark.parametrize("nx,ny,dt", DUGKS_CASES)
def test_dugks_bitwise(oracle, nx, ny, dt):
    """perform_dugks_step (src/periodic_dugks.F90:25-38): both buffers after 1 and 7 steps
    and update_macros of f(inew), against the reference's own run (ref_dugks.npz)."""
    G = np.load(GOLDEN / "ref_dugks.npz")
    key = f"{nx}x{ny}_dt{dt}"
    rho, ux, uy = (np.asfortranarray(G[f"{key}_{c}"]) for c in ("rho", "ux", "uy"))
    nu, dt_, tau, omega = G[f"{key}_params"]
    par = oracle.set_properties(nu, dt, 0.25)
    assert (par["tau"], par["omega"]) == (tau, omega)
    om_half = 1.0 / (tau / dt + 0.5)
    bufs = [po.new_f(nx, ny), po.new_f(nx, ny)]
    inew, iold = (int(i) for i in G[f"{key}_1_ids"])  # alloc_grid's ids, one swap back
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, bufs[iold - 1])
    done = 0
    for n in (1, 7):
        for _ in range(n - done):
            oracle.dugks_collide(nx, ny, bufs[iold - 1], bufs[inew - 1], omega, om_half)
            oracle.dugks_stream(nx, ny, bufs[iold - 1], bufs[inew - 1], dt)
            iold, inew = inew, iold
        done = n
        assert [iold, inew] == list(G[f"{key}_{n}_ids"])
        np.testing.assert_array_equal(bufs[iold - 1][:ny], G[f"{key}_{n}_iold"])
        np.testing.assert_array_equal(bufs[inew - 1][:ny], G[f"{key}_{n}_inew"])
        r, u, v = oracle.update_macros(nx, ny, bufs[inew - 1])
        np.testing.assert_array_equal(np.stack([r, u, v], axis=2), G[f"{key}_{n}_macros"])


@pytest.mark.parametrize("nx,ny,dt", DUGKS_CASES)
def test_dugks_face_collision_build_bitwise(oracle, nx, ny, dt):
    """perform_dugks_step of the -DDUGKS build (src/periodic_dugks.F90:68,199,267,308): the
    half-step collision at 0.75 of 1/(tau/dt + 1/2) and the stream with face collisions of
    rate 1/(4 tau/dt + 1); both buffers after 1 and 7 steps and update_macros of f(inew),
    against that build's own run (ref_dugks_faces.npz)."""
    G = np.load(GOLDEN / "ref_dugks_faces.npz")
    key = f"{nx}x{ny}_dt{dt}"
    rho, ux, uy = (np.asfortranarray(G[f"{key}_{c}"]) for c in ("rho", "ux", "uy"))
    nu, dt_, tau, omega = G[f"{key}_params"]
    par = oracle.set_properties(nu, dt, 0.25)
    assert (par["tau"], par["omega"]) == (tau, omega)
    om_half = 0.75 * (1.0 / (tau / dt + 0.5))
    om_face = 1.0 / (4.0 * (tau / dt) + 1.0)
    bufs = [po.new_f(nx, ny), po.new_f(nx, ny)]
    inew, iold = (int(i) for i in G[f"{key}_1_ids"])
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, bufs[iold - 1])
    done = 0
    for n in (1, 7):
        for _ in range(n - done):
            oracle.dugks_collide(nx, ny, bufs[iold - 1], bufs[inew - 1], omega, om_half)
            oracle.dugks_stream_face(nx, ny, bufs[iold - 1], bufs[inew - 1], dt, om_face)
            iold, inew = inew, iold
        done = n
        assert [iold, inew] == list(G[f"{key}_{n}_ids"])
        np.testing.assert_array_equal(bufs[iold - 1][:ny], G[f"{key}_{n}_iold"])
        np.testing.assert_array_equal(bufs[inew - 1][:ny], G[f"{key}_{n}_inew"])
        r, u, v = oracle.update_macros(nx, ny, bufs[inew - 1])
        np.testing.assert_array_equal(np.stack([r, u, v], axis=2), G[f"{key}_{n}_macros"])
    # the default build's step from the same start gives another result
    a, b = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, a)
    a, b = oracle.dugks_run(nx, ny, a, b, 7, dt, omega, 1.0 / (tau / dt + 0.5))
    assert not np.array_equal(a[:ny], G[f"{key}_7_iold"])


@pytest.mark.parametrize("nx,ny", [(16, 16), (12, 20)])
def test_split_build_bitwise(oracle, nx, ny):
    """collide_bgk of the -DSPLIT build (bgk_kernel_cache, src/collision_bgk.F90:84-176)
    is the oracle's bgk_split; the build's collide_trt (trt_split) equals the default
    trt_naive bit for bit."""
    G = np.load(GOLDEN / "ref_split.npz")
    key = f"{nx}x{ny}"
    rho, ux, uy = (np.asfortranarray(G[f"{key}_{c}"]) for c in ("rho", "ux", "uy"))
    omega = float(G[f"{key}_omega"][0])
    for model, want in (("bgk_split", "bgk"), ("trt", "trt")):
        fa, fb = po.new_f(nx, ny), po.new_f(nx, ny)
        oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, fa)
        a, b, done = fa, fb, 0
        for n in (1, 10, 100):
            latest, prev = oracle.run(model, nx, ny, a, b, n - done, omega, 0.25)
            done = n
            np.testing.assert_array_equal(latest[:ny], G[f"{key}_{want}_{n}_latest"])
            a, b = latest, prev
    # and the split algebra is not the default one (differs at roundoff)
    fa, fb = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, fa)
    latest, _ = oracle.run("bgk", nx, ny, fa, fb, 100, omega, 0.25)
    d = np.abs(latest[:ny] - G[f"{key}_bgk_100_latest"]).max()
    assert 0 < d < 1e-13


def test_shear_layer_bitwise():
    """orc_shear_eval against shear_layer_t's grid evaluation run by the reference
    (src/benchmarks/shear_layer.f90:25-52; ref_shear.npz), the batched sweep's parameters
    at N = 32, 96, 512 and a non-square grid with L != N."""
    G = np.load(GOLDEN / "ref_shear.npz")
    O = po.Oracle()
    for key in ("32x32", "96x96", "512x512", "48x40"):
        nx, ny = map(int, key.split("x"))
        length, u0, k, delta = G[f"{key}_params"]
        rho, ux, uy = O.shear(nx, ny, length, u0, k, delta)
        if f"{key}_ux" in G:
            np.testing.assert_array_equal(rho, G[f"{key}_rho"])
            np.testing.assert_array_equal(ux, G[f"{key}_ux"])
            np.testing.assert_array_equal(uy, G[f"{key}_uy"])
        else:
            np.testing.assert_array_equal(ux[:, 0], G[f"{key}_ux_col"])
            np.testing.assert_array_equal(uy[0, :], G[f"{key}_uy_row"])
            assert [sha(rho), sha(ux), sha(uy)] == G[f"{key}_sha256"].tolist()


@pytest.mark.parametrize("model", ["trt", "bgk"])
def test_triple_step_bitwise(oracle, model):
    """perform_triple_step (src/fvm_bardow.F90:322-340) composed from the oracle's stream,
    copy and collision, against the reference's three buffers and ids after 1 and 5 steps
    (ref_triple.npz): the third slot holds the pre-collision (streamed) state of the last
    step, the other the previous step's."""
    G = np.load(GOLDEN / "ref_triple.npz")
    nx, ny = 16, 12
    nu, dt, magic = G["params"]
    par = oracle.set_properties(nu, dt, magic)
    bufs = {b: po.new_f(nx, ny) for b in (1, 2, 3)}
    iold, inew, imid = 2, 1, 3  # alloc_grid (src/fvm_bardow.F90:175-180)
    oracle.set_pdf_to_equilibrium(nx, ny, np.asfortranarray(G["rho"]), np.asfortranarray(G["ux"]),
                                  np.asfortranarray(G["uy"]), bufs[iold])
    done = 0
    for n in (1, 5):
        for _ in range(n - done):
            oracle.stream(nx, ny, bufs[iold], bufs[inew])
            bufs[iold][...] = bufs[inew]
            oracle.collide(model, nx, ny, bufs[inew], par["omega"], par["trt_magic"])
            iold, inew, imid = inew, imid, iold
        done = n
        assert [iold, inew, imid] == G[f"{model}_{n}_ids"].tolist()
        for b in (1, 2, 3):
            np.testing.assert_array_equal(bufs[b][:ny], G[f"{model}_{n}_buf{b}"], err_msg=f"{n} {b}")
    # a post-collision copy in the third slot would not match
    assert not np.array_equal(G[f"{model}_5_buf{imid}"], G[f"{model}_5_buf{iold}"])


@pytest.mark.parametrize("model", MODELS)
def test_config1_l2_norm_against_reference(oracle, model):
    """orc_l2_norm of the oracle's lagged macros (update_macros after 1000 steps) against
    the analytic field at t = 1000 equals the golden l2_lag computed from the reference's
    own macros (ref_config1.json)."""
    C = json.loads((GOLDEN / "ref_config1.json").read_text())
    n, u0, nu, steps = C["n"], C["u0"], C["nu"], C["steps"]
    k = 2 * np.pi / n
    p, ux, uy = oracle.tgv(n, n, k, k, u0, nu, 0.0)
    rho = np.asfortranarray(p / (1.0 / 3.0) + 1.0)
    par = oracle.set_properties(nu, 1.0, 0.25)
    fa, fb = po.new_f(n, n), po.new_f(n, n)
    oracle.set_pdf_to_equilibrium(n, n, rho, ux, uy, fa)
    latest, prev = oracle.run(model, n, n, fa, fb, steps, par["omega"], par["trt_magic"])
    _, u, v = oracle.update_macros(n, n, prev)
    _, ua, va = oracle.tgv(n, n, k, k, u0, nu, float(steps))
    assert oracle.l2_norm(u, v, ua, va) == pytest.approx(C["models"][model]["l2_lag"], rel=1e-13)


FDM_VARIANTS = ["fdm_bardow_wls", "fdm_bardow_wls_gauss_v1", "fdm_bardow_wls_gauss_v2",
                "fdm_bardow_iso"]


@pytest.mark.parametrize("key,dt", [("16x16", 1.0), ("16x16", 0.5), ("12x20", 1.0)])
@pytest.mark.parametrize("op", FDM_VARIANTS)
def test_fdm_stencil_variants_bitwise(oracle, op, key, dt):
    """stream_fdm_bardow's compile-time stencils (src/fvm_bardow.F90:591-669; -DFDM_WLS,
    -DFDM_WLS_GAUSS_V1/_V2, -DFDM_ISO) with collide_bgk, 5 steps, against the reference
    built with each option (ref_fdm_variants.npz)."""
    G = np.load(GOLDEN / "ref_fdm_variants.npz")
    nx, ny = map(int, key.split("x"))
    rho, ux, uy = (np.asfortranarray(G[f"{key}_{k}"]) for k in ("rho", "ux", "uy"))
    omega = float(G[f"{op}_{key}_dt{dt}_omega"][0])
    assert omega == oracle.set_properties(0.02, dt, 0.25)["omega"]
    a, b = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, a)
    for _ in range(5):
        oracle.stream_op(op, nx, ny, a, b, dt)
        oracle.collide("bgk", nx, ny, b, omega)
        a, b = b, a
    np.testing.assert_array_equal(a[:ny], G[f"{op}_{key}_dt{dt}"])
