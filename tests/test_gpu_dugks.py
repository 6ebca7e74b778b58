"""DUGKS (src/periodic_dugks.F90) on the GPU against the oracle, bit-exact in FP64:
the reference-form operators (plbm_dugks_collide / plbm_dugks_stream), the fused step
(plbm_dugks_step) with its slot bookkeeping, the reference-named Python API, the
Fortran module shim, and a full-size property check."""
from __future__ import annotations

import hashlib
import json
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

import pyoracle as po

pytestmark = pytest.mark.gpu


def fields(nx, ny, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    return (np.asfortranarray(1 + 0.01 * rng.standard_normal((ny, nx))),
            np.asfortranarray(amp * rng.standard_normal((ny, nx))),
            np.asfortranarray(amp * rng.standard_normal((ny, nx))))


def rates(nu, dt):
    tau = 3.0 * nu
    return dt / (tau + 0.5 * dt), 1.0 / (tau / dt + 0.5)


@pytest.mark.parametrize("nx,ny,dt", [(40, 40, 1.0), (33, 70, 0.45), (130, 5, 2.0), (300, 257, 0.8),
                                      (1100, 1000, 0.7), (3300, 2500, 1.0)])
def test_dugks_operators_bit_exact(P, gpu, oracle, nx, ny, dt):
    """dugks_collide then dugks_stream, launched one by one, 5 steps."""
    print(f"case dugks operators {nx}x{ny} dt={dt}", flush=True)
    rho, ux, uy = fields(nx, ny, nx * 3 + ny)
    om_full, om_half = rates(0.01, dt)
    lat = P.Lattice(nx, ny, 2, "f64")
    lat.equilibrium_init(0, rho, ux, uy)
    a, b = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, a)
    for s in range(5):
        src, dst = s % 2, 1 - s % 2
        lat.dugks_collide(src, dst, om_full, om_half)
        lat.dugks_stream(src, dst, dt)
        oracle.dugks_collide(nx, ny, a, b, om_full, om_half)
        oracle.dugks_stream(nx, ny, a, b, dt)
        a, b = b, a
    np.testing.assert_array_equal(lat.download_f(1)[:ny], a[:ny])
    np.testing.assert_array_equal(lat.download_f(0)[:ny], b[:ny])


# (1100, 1000) and (3300, 2500) run the flux kernel with 8- and 32-column strips whose
# last strip is ragged; the small grids use single-column strips.
@pytest.mark.parametrize("per_pass", [1, 2, 3])
@pytest.mark.parametrize("nsteps", [1, 2, 7])
@pytest.mark.parametrize("nx,ny,dt", [(64, 64, 1.0), (45, 130, 0.6), (1100, 1000, 0.9),
                                      (3300, 2500, 1.0), (200, 3, 0.7)])
def test_dugks_fused_step_slots(P, gpu, oracle, nx, ny, dt, nsteps, per_pass):
    """plbm_dugks_step: after n steps each slot holds what the reference's buffer of the
    same index holds (latest populations and the half-collided state), with the steps
    before the last one in passes of 1, 2 or 3 steps."""
    print(f"case dugks fused {nx}x{ny} dt={dt} n={nsteps} per_pass={per_pass}", flush=True)
    rho, ux, uy = fields(nx, ny, 7 + nsteps)
    om_full, om_half = rates(0.02, dt)
    lat = P.Lattice(nx, ny, 2, "f64")
    lat.dugks_set_steps_per_pass(per_pass)
    lat.equilibrium_init(0, rho, ux, uy)
    a, b = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, a)
    p0, p1 = lat.buffer_ptr(0), lat.buffer_ptr(1)
    lat.dugks_step(0, 1, nsteps, dt, om_full, om_half)
    bufs = [a, b]
    iold, inew = 0, 1
    for _ in range(nsteps):
        oracle.dugks_collide(nx, ny, bufs[iold], bufs[inew], om_full, om_half)
        oracle.dugks_stream(nx, ny, bufs[iold], bufs[inew], dt)
        iold, inew = inew, iold
    np.testing.assert_array_equal(lat.download_f(iold)[:ny], bufs[iold][:ny])
    np.testing.assert_array_equal(lat.download_f(inew)[:ny], bufs[inew][:ny])
    # the two slots may have exchanged their device storage, nothing else
    assert {lat.buffer_ptr(0), lat.buffer_ptr(1)} == {p0, p1}


@pytest.mark.parametrize("nx,ny,dt", [(16, 16, 1.0), (12, 20, 0.5), (24, 8, 2.0)])
def test_dugks_golden(P, gpu, nx, ny, dt):
    """The reference's own perform_dugks_step run (ref_dugks.npz) through the
    reference-named Python API (fused path) and through the procedure slots."""
    G = np.load(GOLDEN / "ref_dugks.npz")
    key = f"{nx}x{ny}_dt{dt}"
    nu = float(G[f"{key}_params"][0])
    for fused in (True, False):
        g = P.alloc_grid(nx, ny, 2, log=False)
        g.mf[:, :, 0], g.mf[:, :, 1], g.mf[:, :, 2] = (G[f"{key}_{c}"] for c in ("rho", "ux", "uy"))
        P.set_properties(g, nu, dt, 0.25)
        g.collision, g.streaming = P.dugks_collide, P.dugks_stream
        P.set_pdf_to_equilibrium(g)
        done = 0
        for n in (1, 7):
            for _ in range(n - done):
                if fused:
                    P.perform_dugks_step(g)
                else:
                    g.collision(g)
                    g.streaming(g)
                    g.iold, g.inew = g.inew, g.iold
            done = n
            assert [g.iold, g.inew] == [int(i) for i in G[f"{key}_{n}_ids"]]
            P.sync_host(g)
            np.testing.assert_array_equal(g.f[:ny, :, :, g.iold - 1], G[f"{key}_{n}_iold"])
            np.testing.assert_array_equal(g.f[:ny, :, :, g.inew - 1], G[f"{key}_{n}_inew"])
            P.update_macros(g)
            np.testing.assert_array_equal(g.mf, G[f"{key}_{n}_macros"])
        P.dealloc_grid(g)


@pytest.mark.parametrize("nsteps", [6, 7])
def test_dugks_perform_steps(P, gpu, nsteps):
    """perform_dugks_steps(grid, n) == n x perform_dugks_step (both slots, ids, macros),
    in passes of 3 on a 300 x 200 grid."""
    nx, ny, dt = 300, 200, 0.8
    rho, ux, uy = fields(nx, ny, 21)
    grids = []
    for _ in range(2):
        g = P.alloc_grid(nx, ny, 2, log=False)
        g.mf[:, :, 0], g.mf[:, :, 1], g.mf[:, :, 2] = rho, ux, uy
        P.set_properties(g, 0.01, dt, 0.25)
        g.collision, g.streaming = P.dugks_collide, P.dugks_stream
        P.set_pdf_to_equilibrium(g)
        grids.append(g)
    grids[0].lattice.dugks_set_steps_per_pass(3)
    P.perform_dugks_steps(grids[0], nsteps)
    for _ in range(nsteps):
        P.perform_dugks_step(grids[1])
    a, b = grids
    assert (a.iold, a.inew) == (b.iold, b.inew)
    P.sync_host(a)
    P.sync_host(b)
    np.testing.assert_array_equal(a.f[:ny], b.f[:ny])
    P.update_macros(a)
    P.update_macros(b)
    np.testing.assert_array_equal(a.mf, b.mf)


def test_dugks_errors(P, gpu):
    lat = P.Lattice(32, 32, 2, "f32s")
    with pytest.raises(P.PlbmError):
        lat.dugks_step(0, 1, 1, 1.0, 1.5, 1.5)
    lat = P.Lattice(32, 32, 2, "f64")
    with pytest.raises(P.PlbmError):
        lat.dugks_stream(0, 0, 1.0)
    with pytest.raises(P.PlbmError):
        lat.dugks_collide(0, 2, 1.5, 1.5)
    with pytest.raises(P.PlbmError):
        lat.dugks_step(0, 1, -1, 1.0, 1.5, 1.5)
    with pytest.raises(P.PlbmError):
        lat.dugks_step(0, 1, 1, float("nan"), 1.5, 1.5)
    with pytest.raises(P.PlbmError):
        lat.dugks_set_steps_per_pass(4)


def test_dugks_full_size_bit_exact_4096(P, gpu, oracle):
    """4096^2, 2 fused steps, against the oracle (OpenMP on the host cores)."""
    n, steps, dt = 4096, 2, 1.0
    print(f"case dugks fused {n}x{n} n={steps}", flush=True)
    rho, ux, uy = fields(n, n, 99, amp=0.02)
    om_full, om_half = rates(0.01, dt)
    lat = P.Lattice(n, n, 2, "f64")
    lat.equilibrium_init(0, rho, ux, uy)
    lat.dugks_step(0, 1, steps, dt, om_full, om_half)
    a, b = po.new_f(n, n), po.new_f(n, n)
    oracle.set_pdf_to_equilibrium(n, n, rho, ux, uy, a)
    a, b = oracle.dugks_run(n, n, a, b, steps, dt, om_full, om_half)
    np.testing.assert_array_equal(lat.download_f(0)[:n], a[:n])
    np.testing.assert_array_equal(lat.download_f(1)[:n], b[:n])


def test_dugks_full_size_conservation_8192(P, gpu):
    """8192^2 (the headline grid), 20 fused steps: mass and momentum of the latest
    populations are conserved to roundoff (collision and flux balance both conserve)."""
    n, dt = 8192, 1.0
    print(f"case dugks fused {n}x{n} conservation", flush=True)
    lat = P.Lattice(n, n, 2, "f64")
    k = 2 * np.pi / n
    p, ux, uy = P.TaylorGreen(n, n, k, k, 0.05, 0.01).eval(0.0)
    lat.equilibrium_init(0, np.asfortranarray(p / P.csqr + 1.0), ux, uy)
    before = lat.reduce(0)
    om_full, om_half = rates(0.01, dt)
    lat.dugks_step(0, 1, 20, dt, om_full, om_half)
    after = lat.reduce(0)  # slot 0 holds the latest state again after an even count
    assert abs(after[P._abi.RED_MASS] - before[P._abi.RED_MASS]) <= 1e-9 * n * n
    assert after[P._abi.RED_NONFINITE] == 0


def test_dugks_fortran_module(gpu, tmp_path):
    """tests/fortran/dugks_check.f90 through the engine's periodic_dugks / fvm_bardow
    modules writes the same bytes as through the reference's modules."""
    exe = ROOT / "periodic-lbm_amd" / "_build" / "tests" / "dugks_check"
    if not exe.exists():
        pytest.skip("Fortran test programs not built")
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = json.loads((GOLDEN / "ref_dugks_app.json").read_text())
    assert r.stdout.split() == ref["stdout"].split()
    for name, want in ref["files"].items():
        got = hashlib.sha256((tmp_path / "out" / name).read_bytes()).hexdigest()
        assert got == want, name


def face_rates(nu, dt):
    """The -DDUGKS build's rates (src/periodic_dugks.F90:56,68-71,179): full, half (0.75 of
    the default), face."""
    tau = 3.0 * nu
    return dt / (tau + 0.5 * dt), 0.75 * (1.0 / (tau / dt + 0.5)), 1.0 / (4.0 * tau / dt + 1.0)


@pytest.mark.parametrize("nx,ny,dt", [(40, 40, 1.0), (33, 70, 0.45), (130, 5, 2.0),
                                      (1100, 1000, 0.7)])
def test_dugks_face_collisions_bit_exact(P, gpu, oracle, nx, ny, dt):
    """The -DDUGKS build's DUGKS (face collisions in the stream, 0.75 half rate): the
    operators one by one (plbm_dugks_collide + plbm_dugks_stream_faces) and the fused
    steps (plbm_dugks_step_faces, 1 and 6 steps with the slot rules) against the oracle,
    both slots, bit for bit."""
    print(f"case dugks faces {nx}x{ny} dt={dt}", flush=True)
    rho, ux, uy = fields(nx, ny, nx * 5 + ny)
    om_full, om_half, om_face = face_rates(0.01, dt)
    lat = P.Lattice(nx, ny, 2, "f64")
    lat.equilibrium_init(0, rho, ux, uy)
    a, b = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, a)
    for s in range(4):
        src, dst = s % 2, 1 - s % 2
        lat.dugks_collide(src, dst, om_full, om_half)
        lat.dugks_stream_faces(src, dst, dt, om_face)
        oracle.dugks_collide(nx, ny, a, b, om_full, om_half)
        oracle.dugks_stream_face(nx, ny, a, b, dt, om_face)
        a, b = b, a
    np.testing.assert_array_equal(lat.download_f(0)[:ny], a[:ny])
    np.testing.assert_array_equal(lat.download_f(1)[:ny], b[:ny])
    for nsteps in (1, 6):
        f = P.Lattice(nx, ny, 2, "f64")
        f.equilibrium_init(0, rho, ux, uy)
        f.dugks_step_faces(0, 1, nsteps, dt, om_full, om_half, om_face)
        a, b = po.new_f(nx, ny), po.new_f(nx, ny)
        oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, a)
        for _ in range(nsteps):
            oracle.dugks_collide(nx, ny, a, b, om_full, om_half)
            oracle.dugks_stream_face(nx, ny, a, b, dt, om_face)
            a, b = b, a
        # the reference swaps its ids each step: slot 0 holds buffer a after an even count
        latest, other = (0, 1) if nsteps % 2 == 0 else (1, 0)
        np.testing.assert_array_equal(f.download_f(latest)[:ny], a[:ny], err_msg=str(nsteps))
        np.testing.assert_array_equal(f.download_f(other)[:ny], b[:ny], err_msg=str(nsteps))
        f.close()


def test_dugks_face_collisions_python_api(P, gpu, oracle):
    """The reference-named Python API with set_dugks_face_collisions(True): the fused
    perform_dugks_step and perform_dugks_steps against the oracle (the rates from tau and
    dt as the -DDUGKS build computes them)."""
    nx, ny, nu, dt = 96, 80, 0.01, 0.8
    rho, ux, uy = fields(nx, ny, 17)
    om_full, om_half, om_face = face_rates(nu, dt)
    a, b = po.new_f(nx, ny), po.new_f(nx, ny)
    oracle.set_pdf_to_equilibrium(nx, ny, rho, ux, uy, a)
    for _ in range(5):
        oracle.dugks_collide(nx, ny, a, b, om_full, om_half)
        oracle.dugks_stream_face(nx, ny, a, b, dt, om_face)
        a, b = b, a
    P.set_dugks_face_collisions(True)
    try:
        for mode in ("steps", "one-by-one"):
            g = P.alloc_grid(nx, ny, 2, log=False)
            g.mf[:, :, 0], g.mf[:, :, 1], g.mf[:, :, 2] = rho, ux, uy
            P.set_properties(g, nu, dt, 0.25)
            g.collision, g.streaming = P.dugks_collide, P.dugks_stream
            P.set_pdf_to_equilibrium(g)
            if mode == "steps":
                P.perform_dugks_steps(g, 5)
            else:
                for _ in range(5):
                    P.perform_dugks_step(g)
            P.sync_host(g)
            np.testing.assert_array_equal(g.f[:ny, :, :, g.iold - 1], a[:ny], err_msg=mode)
            np.testing.assert_array_equal(g.f[:ny, :, :, g.inew - 1], b[:ny], err_msg=mode)
            P.dealloc_grid(g)
    finally:
        P.set_dugks_face_collisions(False)


def test_dugks_face_collisions_fortran_module(gpu, tmp_path):
    """tests/fortran/dugks_check.f90 linked against the shim compiled with -DDUGKS writes
    the same bytes as through the reference's modules compiled with -DDUGKS."""
    exe = ROOT / "periodic-lbm_amd" / "_build" / "tests" / "dugks_check_faces"
    if not exe.exists():
        pytest.skip("Fortran test programs not built")
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = json.loads((GOLDEN / "ref_dugks_faces_app.json").read_text())
    assert r.stdout.split() == ref["stdout"].split()
    for name, want in ref["files"].items():
        got = hashlib.sha256((tmp_path / "out" / name).read_bytes()).hexdigest()
        assert got == want, name
