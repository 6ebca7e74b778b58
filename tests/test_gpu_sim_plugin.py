"""The reference's plugin protocol (sim/lbm.h) on the engine: SimPlugin drives
libplbm.so's c_plbm_* and matches the oracle's BGK run; TGVStudy norms behave."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu


def test_plugin_tgv_matches_oracle(P, gpu, oracle):
    n, steps = 48, 200
    umax, nu = 0.01, 0.01 * 48 / 100
    omega = 1.0 / (3 * nu + 0.5)
    case = P.sim.TGVStudy((n, n), umax, (2 * np.pi / n, 2 * np.pi / n), nu)
    p0, u0 = case.solution(0.0)
    sim = P.sim.SimPlugin()
    sim.init((n, n), 1.0, p0, u0)
    tm, _ = P.sim.run(sim, omega, steps)
    rho, u = sim.vars()
    sim.free()
    # oracle: same equilibrium start (rho = 1 + 3p), BGK, (ny, nx) layout = transpose
    fa, fb = po.new_f(n, n), po.new_f(n, n)
    oracle.set_pdf_to_equilibrium(n, n, np.asfortranarray((1 + 3 * p0).T),
                                  np.asfortranarray(u0[:, :, 0].T), np.asfortranarray(u0[:, :, 1].T), fa)
    latest, _ = oracle.run("bgk", n, n, fa, fb, steps, omega)
    r, ux, uy = oracle.update_macros(n, n, latest)
    np.testing.assert_allclose(rho, r.T, rtol=1e-13)
    np.testing.assert_allclose(u[:, :, 0], ux.T, atol=1e-15)
    np.testing.assert_allclose(u[:, :, 1], uy.T, atol=1e-15)
    assert tm == steps
    err = case.norm2(tm, u)
    assert err < 5e-3 and case.norm2x(tm, u[:, :, 0]) < 5e-3 and case.norm1(tm, u) < 1e-4


def test_plugin_perfsweep_and_storage(P, gpu):
    sim = P.sim.SimPlugin()
    grids, mlups = P.sim.perfsweep(sim, 64, 129, step=64, tsteps=20)
    assert list(grids) == [64, 128] and np.all(mlups > 0)
    n = 32
    case = P.sim.TGVStudy((n, n), 0.02, (2 * np.pi / n,) * 2, 0.02)
    p0, u0 = case.solution(0.0)
    out = []
    for storage in (P._abi.F64S, P._abi.F32S):
        sim.init((n, n), 1.0, p0, u0, storage=storage)
        P.sim.run(sim, 1.0 / (3 * 0.02 + 0.5), 50)
        out.append(sim.vars()[1])
        sim.free()
    assert np.max(np.abs(out[0] - out[1])) < 1e-7
