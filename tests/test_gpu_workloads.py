"""The benchmark configurations at their real sizes.

* 8192^2 (the headline grid), FP64: runs in full temporally blocked passes of 8 steps
  (the pass kernel the bench times, including the last pass that writes the lagged
  state), uneven splits and the benchmark's 20-step run (6 + 6 + 8), both slots against
  the C oracle bit for bit; 4096^2 TRT with noise (the FP64 roofline case) for 200 steps.
* 8192^2 TRT and regularized, 5000 steps, FP32- and FP16-shifted storage against the
  FP64 run (itself bit-identical to the oracle): velocity relative L2 and mass drift.
* The batched sweep: 256 cases, n from {32 .. 512}, BGK, FP32-shifted storage, 2000 steps
  under the autotuned plan, against the same cases in FP64: no non-finite case, per-case
  mass and momentum drift, per-case velocity error; the largest FP64 cases (384^2, 512^2)
  bit-identical to the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import tgv_fields

import pyoracle as po

pytestmark = pytest.mark.gpu


def headline_fields(P, n, u0=0.05, nu=0.01, seed=7):
    rho, ux, uy = tgv_fields(P, n, u0, nu)
    P.add_noise(ux, uy, seed, 1e-3)
    return rho, ux, uy, 1.0 / (3 * nu + 0.5)


@pytest.mark.parametrize("model,steps", [("trt", 16), ("rr", 16), ("trt", 17), ("trt", 20),
                                         ("trt", 15)])
def test_headline_passes_bit_exact_8192(P, gpu, oracle, model, steps):
    """8192^2 FP64 through plbm_run at the automatic depth (8): 16 steps are two full
    8-step passes (the second also writes the lagged slot), 17 steps run as 4 + 5 + 8,
    20 steps (the benchmark's run) as 6 + 6 + 8 and 15 as 7 + 8, the TRT passes of 6 and
    7 steps in their unrolled steady-state form.  Latest and lagged slot against the C
    oracle, bit for bit."""
    n = 8192
    print(f"case headline {model} {n}^2 {steps} steps", flush=True)
    rho, ux, uy, omega = headline_fields(P, n)
    lat = P.Lattice(n, n, 2, "f64")
    lat.set_params(omega, 0.25)
    assert lat.get_steps_per_launch(model) == 8
    lat.equilibrium_init(0, rho, ux, uy)
    lat.set_run_timing(True)
    iold, inew = lat.run(model, steps, 0, 1)
    assert (iold, inew) == ((1, 0) if steps % 2 else (0, 1))
    plan = {15: [7, 8], 16: [8, 8], 17: [4, 5, 8], 20: [6, 6, 8]}[steps]
    assert [d for d, _ in lat.run_launch_times()] == plan
    got_latest = lat.download_f(iold)
    got_prev = lat.download_f(inew)
    lat.close()
    fa, fb = po.new_f(n, n), po.new_f(n, n)
    oracle.set_pdf_to_equilibrium(n, n, rho, ux, uy, fa)
    latest, prev = oracle.run(model, n, n, fa, fb, steps, omega, 0.25)
    assert np.array_equal(got_latest, latest), "latest state differs from the oracle"
    assert np.array_equal(got_prev, prev), "lagged state differs from the oracle"


def test_roofline_case_4096_bit_exact(P, gpu, oracle):
    """The 4096^2 FP64 roofline case (TRT, Taylor-Green u0 = 0.04, nu = 0.005 plus
    U(+-1e-3) noise of seed 1), 200 steps through plbm_run (passes of 8), against the C
    oracle bit for bit."""
    n, steps = 4096, 200
    print(f"case roofline {n}^2 trt {steps} steps", flush=True)
    rho, ux, uy, omega = headline_fields(P, n, 0.04, 0.005, 1)
    lat = P.Lattice(n, n, 2, "f64")
    lat.set_params(omega, 0.25)
    lat.equilibrium_init(0, rho, ux, uy)
    iold, inew = lat.run("trt", steps, 0, 1)
    got = lat.download_f(iold)
    lat.close()
    fa, fb = po.new_f(n, n), po.new_f(n, n)
    oracle.set_pdf_to_equilibrium(n, n, rho, ux, uy, fa)
    latest, _ = oracle.run("trt", n, n, fa, fb, steps, omega, 0.25)
    assert np.array_equal(got, latest)


# Velocity relative L2 against the FP64 run and relative mass drift after 5000 steps at
# 8192^2, measured on MI355X (TRT / RR): FP32 velocity 8.7e-7 / 3.2e-7, mass 5.7e-9 /
# 5.0e-13; FP16 velocity 9.8e-4 / 8.4e-4, mass 2.2e-6 / 5.5e-6.  FP32 velocity is held to
# the 1e-5 bound of the precision target, the rest to about twice the worst measured.
VEL_TOL = {"f32s": 1e-5, "f16s": 2e-3}
MASS_TOL = {"f32s": 2e-8, "f16s": 1.2e-5}


@pytest.mark.parametrize("model", ["trt", "rr"])
def test_config3_shifted_storage_8192(P, gpu, model):
    """8192^2, 5000 steps, the headline start: FP32- and FP16-shifted storage against the
    FP64 run (bit-identical to the oracle, test above)."""
    import torch
    n, steps = 8192, 5000
    rho, ux, uy, omega = headline_fields(P, n)

    def run(storage):
        print(f"case config3 {model} {storage} {n}^2 {steps} steps", flush=True)
        lat = P.Lattice(n, n, 2, storage)
        lat.set_params(omega, 0.25)
        lat.equilibrium_init(0, rho, ux, uy)
        mass0 = lat.reduce(0)[P._abi.RED_MASS]
        iold, _ = lat.run(model, steps, 0, 1)
        red = lat.reduce(iold)
        u = [torch.empty((n, n), dtype=torch.float64, device="cuda") for _ in range(3)]
        lat.macros_dev(iold, *(t.data_ptr() for t in u))
        lat.sync()
        lat.close()
        return u[1], u[2], mass0, red

    u64, v64, m64_0, red64 = run("f64")
    assert red64[P._abi.RED_NONFINITE] == 0
    drift64 = abs(red64[P._abi.RED_MASS] - m64_0) / m64_0
    assert drift64 < 1e-12
    den = torch.sqrt(torch.sum(u64 ** 2 + v64 ** 2)).item()
    for storage in ("f32s", "f16s"):
        u, v, m0, red = run(storage)
        assert red[P._abi.RED_NONFINITE] == 0
        err = torch.sqrt(torch.sum((u - u64) ** 2 + (v - v64) ** 2)).item() / den
        drift = abs(red[P._abi.RED_MASS] - m0) / m0
        print(f"{model} {storage}: velocity rel L2 {err:.3e}, mass drift {drift:.3e} "
              f"(FP64 {drift64:.1e})")
        assert err < VEL_TOL[storage]
        assert drift < MASS_TOL[storage]
        del u, v
        torch.cuda.empty_cache()


def _momentum(rho, ux, uy):
    return np.array([np.sum(rho * ux), np.sum(rho * uy)])


def test_config4_sweep_full_workload(P, gpu):
    """256 cases (n from {32..512}, nu log-uniform in [1e-3, 5e-2], half Taylor-Green,
    half shear layers), BGK, FP32-shifted storage, 2000 steps under the autotuned plan,
    against the same cases in FP64."""
    cases = P.sweep_cases(256, seed=4)
    sizes = sorted({c.n for c in cases})
    assert len(cases) == 256 and sizes[0] == 32 and sizes[-1] == 512
    steps = 2000
    print(f"case config4 sweep 256 cases {steps} steps", flush=True)
    b = P.Batch.from_cases(cases, "f32s", "bgk")
    b.autotune(10)
    print("plan", b.get_launch(), b.get_steps_per_pass(), b.get_pass_columns(), flush=True)
    R = P._abi
    red0 = b.reduce()
    mom0 = [_momentum(*b.macros(i)) for i in range(len(cases))]
    b.run(steps)
    b.sync()
    red = b.reduce()
    b64 = P.Batch.from_cases(cases, "f64", "bgk")
    b64.run(steps)
    b64.sync()
    red64 = b64.reduce()
    assert np.all(red[:, R.RED_NONFINITE] == 0) and np.all(red64[:, R.RED_NONFINITE] == 0)
    worst = {"mass": (0.0, None), "mom": (0.0, None), "vel": (0.0, None)}
    for i, c in enumerate(cases):
        r, u, v = b.macros(i)
        r64, u64, v64 = b64.macros(i)
        mass_drift = abs(red[i, R.RED_MASS] - red0[i, R.RED_MASS]) / red0[i, R.RED_MASS]
        # momentum starts at (nearly) zero: drift relative to n^2 u0
        mom_drift = np.abs(_momentum(r, u, v) - mom0[i]).max() / (c.n * c.n * c.u0)
        # velocity error relative to the initial velocity scale n u0 (the decayed field of
        # a high-viscosity case is many orders smaller than its FP32 state's resolution)
        vel = np.sqrt(np.sum((u - u64) ** 2 + (v - v64) ** 2)) / (c.n * c.u0)
        for key, val in (("mass", mass_drift), ("mom", mom_drift), ("vel", vel)):
            if val > worst[key][0]:
                worst[key] = (val, c)
    for key, (val, c) in worst.items():
        print(f"worst {key}: {val:.2e} ({c.kind} n={c.n} nu={c.nu:.2e})")
    # measured on MI355X: mass 1.6e-8, momentum 1.0e-6, velocity 1.7e-5 (all in the units
    # above); held to about 2-3 times that
    assert worst["mass"][0] < 5e-8
    assert worst["mom"][0] < 3e-6
    assert worst["vel"][0] < 4e-5
    b.close()
    b64.close()


def test_config4_largest_fp64_cases_bit_exact(P, gpu, oracle):
    """The sweep's largest grids in FP64 through the batched passes (6 steps per pass),
    20 steps: a 384^2 Taylor-Green and a 512^2 shear layer (and a 96^2 one beside them),
    each case bit-identical to the oracle."""
    cases = [P.Case(384, 0.004, "taylor_green", 0.05), P.Case(512, 0.02, "shear_layer", 0.04),
             P.Case(96, 0.001, "shear_layer", 0.04)]
    steps = 20
    print(f"case config4 fp64 largest cases {steps} steps", flush=True)
    b = P.Batch.from_cases(cases, "f64", "bgk")
    b.set_steps_per_pass(6)
    b.run(steps)
    b.sync()
    for i, c in enumerate(cases):
        rho, ux, uy = c.fields()
        fa, fb = po.new_f(c.n, c.n), po.new_f(c.n, c.n)
        oracle.set_pdf_to_equilibrium(c.n, c.n, rho, ux, uy, fa)
        latest, _ = oracle.run("bgk", c.n, c.n, fa, fb, steps, c.omega)
        np.testing.assert_array_equal(b.download_f(i)[:c.n], latest[:c.n], err_msg=str(c))
    b.close()
