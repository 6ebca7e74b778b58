"""Benchmark: D2Q9 TRT (magic Lambda = 1/4) on a periodic 8192 x 8192 grid, one MI355X.

One step = one pull-stream + collide update of the whole grid (the reference's
lbm_stream + collide_trt, src/periodic_lbm.f90:32-129 + src/collision_trt.F90:41-160),
run through plbm_run, which advances the lattice in passes of several fused steps per
HBM sweep (temporal blocking; bit-identical to single steps).  The headline value is FP64
MLUPS (reference arithmetic, bit-exact); the shifted FP32 and FP16 storage forms are
timed the same way and their velocity / mass error against the FP64 run after 5000
steps is stated beside their MLUPS.  The roofline object describes the kernel that
dominates the timed region (the pass kernel): its algorithmic bytes per launch are one
read and one write of every population (18 s bytes per cell), whatever its step count.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--acc-steps S] [--no-cpu]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
N = 8192
NU, U0, NOISE, SEED = 0.01, 0.05, 1e-3, 7
MAGIC = 0.25
BYTES_PER_LU = {"f64": 144, "f64s": 144, "f32s": 72, "f16s": 36}
HBM_PEAK_GBS = 8000.0


def progress(msg):
    """One line per finished leg on stderr (the JSON line stays the only stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--acc-steps", type=int, default=5000,
                    help="steps of the FP64 vs FP32/FP16 accuracy comparison (0: skip)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=150)
    ap.add_argument("--n", type=int, default=N, help="grid size (default 8192)")
    ap.add_argument("--batch-steps", type=int, default=2000,
                    help="steps of the batched uneven-size sweep (0: skip)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-rr", action="store_true",
                    help="skip the leg of the other collisions (BGK, RR)")
    ap.add_argument("--next-steps", type=int, default=50,
                    help="timed steps of the DUGKS and FV/FD streaming legs (0 skips them)")
    return ap.parse_args()


def initial_fields(P, n):
    """Taylor-Green (u0=0.05, nu=0.01, k=2pi/n) plus U(+-1e-3) noise of seed 7."""
    import numpy as np
    k = 2 * np.pi / n
    p, ux, uy = P.TaylorGreen(n, n, k, k, U0, NU).eval(0.0)
    P.add_noise(ux, uy, SEED, NOISE)
    return np.asfortranarray(p / P.csqr + 1.0), ux, uy


def timed_passes(torch, lat, model, steps_per_pass, launches):
    """Average duration (ms) of one pass launch, HIP events on the lattice's stream."""
    lat.run_pass(model, steps_per_pass, 0, 1)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for k in range(launches):
        lat.run_pass(model, steps_per_pass, k % 2, 1 - k % 2)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / launches


def timed_run(P, torch, lat, model, steps, warmup, launches=None):
    """Warm up, then time exactly `steps` fused steps with HIP events on the stream the
    kernels are launched on.  Returns (ms per step, final iold).  With a list for
    `launches`, the engine also records an event after each of its launches in the timed
    run, and the list receives (steps advanced, ms) per launch."""
    iold, inew = 0, 1
    iold, inew = lat.run(model, warmup, iold, inew)
    torch.cuda.synchronize()
    if launches is not None:
        lat.set_run_timing(True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    iold, inew = lat.run(model, steps, iold, inew)
    e1.record()
    torch.cuda.synchronize()
    if launches is not None:
        launches.extend(lat.run_launch_times())
        lat.set_run_timing(False)
    return e0.elapsed_time(e1) / steps, iold


def dominant_launch(launches):
    """Of the timed region's launches, the depth whose launches take the most time in
    total: (depth, launches of it, their mean ms)."""
    by = {}
    for d, ms in launches:
        by.setdefault(d, []).append(ms)
    d = max(by, key=lambda k: sum(by[k]))
    return d, len(by[d]), sum(by[d]) / len(by[d])


def gpu_leg(args):
    import numpy as np
    import torch
    import __graft_entry__ as g

    P = g.load_package()
    torch.cuda.set_device(0)
    dev_name = torch.cuda.get_device_name(0)
    # one explicit stream for the engine and the timing events (torch's default stream
    # would be passed as NULL, which the engine reads as "use the lattice's own stream")
    ts = torch.cuda.Stream()
    torch.cuda.set_stream(ts)
    stream = ts.cuda_stream
    n = args.n
    cells = n * n
    rho, ux, uy = initial_fields(P, n)
    d_rho, d_ux, d_uy = (torch.from_numpy(np.ravel(a, order="F")).cuda() for a in (rho, ux, uy))
    omega = 1.0 / (3.0 * NU + 0.5)  # set_properties(nu, dt=1): omega = dt/(3 nu + dt/2)

    results = {}
    for storage in ("f64", "f32s", "f16s"):
        lat = P.Lattice(n, n, 2, storage)
        lat.set_stream(stream)
        lat.set_params(omega, MAGIC)
        lat.equilibrium_init_dev(0, d_rho.data_ptr(), d_ux.data_ptr(), d_uy.data_ptr())
        torch.cuda.synchronize()
        lat.autotune("trt", 0, reps=5)
        block, remap = lat.get_launch("trt")
        ns = lat.get_steps_per_launch("trt")
        launches = []
        ms, _ = timed_run(P, torch, lat, "trt", args.steps, args.warmup, launches)
        mlups = cells / (ms * 1e-3) / 1e6
        # the timed region's launches (HIP events between them on the engine's stream),
        # the kernel that dominates it, and, timed alone, a pass of the tuned depth and the
        # single-step kernel
        dom, dom_n, dom_ms = dominant_launch(launches)
        pass_ms = timed_passes(torch, lat, "trt", ns, 20)
        single_ms = pass_ms if ns == 1 else timed_passes(torch, lat, "trt", 1, 20)
        bpl = BYTES_PER_LU[storage] * cells  # algorithmic bytes of one pass launch
        results[storage] = {"ms_per_step": ms, "mlups": mlups,
                            "steps_per_pass": ns, "pass_ms": pass_ms,
                            "pass_gbs": bpl / (pass_ms * 1e-3) / 1e9,
                            "bytes_per_lu": BYTES_PER_LU[storage] / ns,
                            "timed_launches": [[d, round(t, 5)] for d, t in launches],
                            "dominant_depth": dom, "dominant_launches": dom_n,
                            "dominant_ms": dom_ms,
                            "dominant_gbs": bpl / (dom_ms * 1e-3) / 1e9,
                            "single_step_ms": single_ms,
                            "single_step_gbs": bpl / (single_ms * 1e-3) / 1e9,
                            "launch": {"block": block, "xcd_remap": remap}}
        lat.close()
        torch.cuda.empty_cache()
        progress(f"trt {storage}: {mlups:.0f} MLUPS")

    # accuracy of the shifted forms vs the FP64 run (reference arithmetic) at acc-steps
    if args.acc_steps > 0:
        acc = accuracy(P, torch, stream, "trt", omega, (d_rho, d_ux, d_uy), n, args.acc_steps)
        for storage, v in acc.items():
            results[storage].update(v)
        torch.cuda.empty_cache()
        progress("trt accuracy")

    copy_gbs = P.copy_bandwidth(4 << 30, 20)
    legs = [
        ("config1", lambda: config1_leg(P, torch, stream)),
        ("config2", lambda: config2_leg(P, torch, stream) if n == N else None),
        ("collisions", lambda: None if args.no_rr else collisions_leg(
            P, torch, args, stream, omega, (d_rho, d_ux, d_uy))),
        ("schemes", lambda: schemes_leg(P, torch, args, stream, (d_rho, d_ux, d_uy))
         if args.next_steps > 0 else None),
        ("batched_sweep", lambda: batch_leg(P, torch, args) if args.batch_steps > 0 else None),
        ("adjoint", lambda: None if args.no_adjoint else adjoint_leg(P, torch, args)),
    ]
    extras = {}
    for name, leg in legs:
        extras[name] = leg()
        progress(f"{name} done")
    return results, copy_gbs, dev_name, extras


def accuracy(P, torch, stream, model, omega, fields, n, steps):
    """FP64, FP32- and FP16-shifted runs of `steps` steps from the same start: each
    shifted run's velocity relative L2 against the FP64 run (bit-identical to the
    reference arithmetic), every run's relative mass drift against its own start, and the
    shifted runs' relative mass difference to the FP64 run (FP64 reductions, plbm_reduce)."""
    cells = n * n
    macros, out = {}, {}
    for storage in ("f64", "f32s", "f16s"):
        lat = P.Lattice(n, n, 2, storage)
        lat.set_stream(stream)
        lat.set_params(omega, MAGIC)
        lat.equilibrium_init_dev(0, *(t.data_ptr() for t in fields))
        m0 = lat.reduce(0)[P._abi.RED_MASS]
        iold, _ = lat.run(model, steps, 0, 1)
        red = lat.reduce(iold)
        m = [torch.empty(cells, dtype=torch.float64, device="cuda") for _ in range(3)]
        lat.macros_dev(iold, *(t.data_ptr() for t in m))
        torch.cuda.synchronize()
        macros[storage] = m
        out[storage] = {"mass_drift_rel": abs(red[P._abi.RED_MASS] - m0) / m0,
                        "nonfinite_cells": int(red[P._abi.RED_NONFINITE]),
                        "acc_steps": steps, "_mass": red[P._abi.RED_MASS]}
        lat.close()
    r64, u64, v64 = macros["f64"]
    den = torch.sum(u64 * u64 + v64 * v64)
    for storage in ("f32s", "f16s"):
        r, u, v = macros[storage]
        out[storage]["vel_rel_l2_vs_f64"] = torch.sqrt(
            torch.sum((u - u64) ** 2 + (v - v64) ** 2) / den).item()
        out[storage]["mass_rel_diff_vs_f64"] = abs(
            out[storage]["_mass"] - out["f64"]["_mass"]) / out["f64"]["_mass"]
    for v in out.values():
        del v["_mass"]
    del macros
    torch.cuda.empty_cache()
    return out


def collisions_leg(P, torch, args, stream, omega, fields):
    """The other collision models on the same 8192^2 start: BGK (src/collision_bgk.F90)
    and recursive-regularized (src/collision_regularized.F90), FP64, FP32- and
    FP16-shifted MLUPS through plbm_run (autotuned), and each shifted form's velocity
    error and mass drift against the FP64 run after acc_steps steps (5000, as the TRT
    leg)."""
    n = args.n
    cells = n * n
    out = {}
    for model in ("bgk", "rr"):
        res = {"grid": [n, n], "steps": args.steps}
        for storage in ("f64", "f32s", "f16s"):
            lat = P.Lattice(n, n, 2, storage)
            lat.set_stream(stream)
            lat.set_params(omega, MAGIC)
            lat.equilibrium_init_dev(0, *(t.data_ptr() for t in fields))
            torch.cuda.synchronize()
            lat.autotune(model, 0, reps=5)
            ms, _ = timed_run(P, torch, lat, model, args.steps, args.warmup)
            res[storage] = {"ms_per_step": round(ms, 5), "mlups": round(cells / ms / 1e3, 1),
                            "steps_per_pass": lat.get_steps_per_launch(model)}
            lat.close()
            torch.cuda.empty_cache()
        for storage in ("f32s", "f16s"):
            res[f"{storage}_speedup_vs_f64"] = round(
                res[storage]["mlups"] / res["f64"]["mlups"], 3)
        if args.acc_steps > 0:
            for storage, v in accuracy(P, torch, stream, model, omega, fields, n,
                                       args.acc_steps).items():
                res[storage].update(v)
        out[model] = res
    return out


def config1_leg(P, torch, stream):
    """The correctness configuration: 64^2 Taylor-Green (u0 = 0.05, nu = 0.01, k = 2 pi/64,
    rho = 1 + 3p, equilibrium start), 1000 steps through plbm_run, FP64, for BGK (the
    configuration's model) and TRT / RR.  Reports whether both slots are bit-identical to
    the reference Fortran's buffers (sha256 fingerprints in tests/golden/ref_config1.json,
    written by the flang-built reference), the app's relative L2 norm of the lagged state
    against the analytic field at t = 1000 (app/main_taylor_green.f90:174-212) beside the
    reference's value, the kinetic-energy decay against the analytic exp(-2 nu k^2 t) per
    velocity component, and the mass drift."""
    import hashlib
    import numpy as np
    G = json.loads((ROOT / "tests" / "golden" / "ref_config1.json").read_text())
    n, u0, nu, steps = G["n"], G["u0"], G["nu"], G["steps"]
    k = 2 * np.pi / n
    p, ux, uy = P.TaylorGreen(n, n, k, k, u0, nu).eval(0.0)
    rho = np.asfortranarray(p / P.csqr + 1.0)
    _, uxa, uya = P.TaylorGreen(n, n, k, k, u0, nu).eval(float(steps))
    omega = 1.0 / (3.0 * nu + 0.5)
    sha = lambda a: hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()
    out = {"grid": [n, n], "steps": steps, "u0": u0, "nu": nu}
    for model in ("bgk", "trt", "rr"):
        lat = P.Lattice(n, n, 2, "f64")
        lat.set_stream(stream)
        lat.set_params(omega, MAGIC)
        lat.equilibrium_init(0, rho, ux, uy)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        iold, inew = lat.run(model, steps, 0, 1)
        e1.record()
        torch.cuda.synchronize()
        latest, prev = lat.download_f(iold)[:n], lat.download_f(inew)[:n]
        r, u, v = lat.macros(inew)  # update_macros reads the lagged slot
        lat.close()
        M = G["models"][model]
        l2 = float(np.linalg.norm(np.hypot(u - uxa, v - uya)) / np.linalg.norm(np.hypot(uxa, uya)))
        ke = float(np.sum(u ** 2 + v ** 2) / np.sum(ux ** 2 + uy ** 2))
        out[model] = {
            "bit_identical_to_reference": sha(latest) == M["sha256_latest"]
            and sha(prev) == M["sha256_prev"],
            "l2_lag_vs_analytic": l2, "reference_l2_lag": M["l2_lag"],
            # the lagged state is at step steps - 1: u decays as exp(-2 nu k^2 t)
            "ke_ratio": ke, "analytic_ke_ratio": float(np.exp(-4 * nu * k * k * (steps - 1))),
            "mass_drift_rel": abs(float(np.sum(r)) - float(np.sum(rho))) / float(np.sum(rho)),
            "ms_per_step": round(e0.elapsed_time(e1) / steps, 5)}
    return out


def config2_leg(P, torch, stream, steps=2000):
    """The FP64 roofline configuration: 4096^2 TRT (Lambda = 1/4), Taylor-Green u0 = 0.04,
    nu = 0.005 plus U(+-1e-3) noise of seed 1, equilibrium start, `steps` steps through
    plbm_run (autotuned).  MLUPS, the pass kernel's and the single-step kernel's achieved
    algorithmic bandwidth (144 B per cell per launch) against the 8 TB/s peak."""
    import numpy as np
    n, u0, nu = 4096, 0.04, 0.005
    k = 2 * np.pi / n
    p, ux, uy = P.TaylorGreen(n, n, k, k, u0, nu).eval(0.0)
    P.add_noise(ux, uy, 1, NOISE)
    rho = np.asfortranarray(p / P.csqr + 1.0)
    omega = 1.0 / (3.0 * nu + 0.5)
    lat = P.Lattice(n, n, 2, "f64")
    lat.set_stream(stream)
    lat.set_params(omega, MAGIC)
    lat.equilibrium_init(0, rho, ux, uy)
    torch.cuda.synchronize()
    lat.autotune("trt", 0, reps=5)
    ns = lat.get_steps_per_launch("trt")
    lat.equilibrium_init(0, rho, ux, uy)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    iold, _ = lat.run("trt", steps, 0, 1)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    red = lat.reduce(iold)
    pass_ms = timed_passes(torch, lat, "trt", ns, 20)
    single_ms = timed_passes(torch, lat, "trt", 1, 20)
    bpl = BYTES_PER_LU["f64"] * n * n
    lat.close()
    torch.cuda.empty_cache()
    return {"grid": [n, n], "steps": steps, "ms_per_step": round(ms, 5),
            "mlups": round(n * n / ms / 1e3, 1), "steps_per_pass": ns,
            "pass_kernel": f"k_multi<TRT, F64, {ns}>", "pass_ms": round(pass_ms, 5),
            "pass_gbs": round(bpl / pass_ms / 1e6, 1),
            "pass_frac_of_peak": round(bpl / pass_ms / 1e6 / HBM_PEAK_GBS, 4),
            "single_step_ms": round(single_ms, 5),
            "single_step_gbs": round(bpl / single_ms / 1e6, 1),
            "single_step_frac_of_peak": round(bpl / single_ms / 1e6 / HBM_PEAK_GBS, 4),
            "nonfinite_cells": int(red[P._abi.RED_NONFINITE])}


def schemes_leg(P, torch, args, stream, fields):
    """The other schemes on the same 8192^2 FP64 start: DUGKS (src/periodic_dugks.F90) as
    the fused two-kernel step and as the reference's collide + stream sequence, and the
    finite-volume / finite-difference streaming operators (src/fvm_bardow.F90:393-893)
    with collide_bgk as perform_step runs them (one fused pass per step), and the same
    pair through plbm_run_op (temporally blocked passes).  Algorithmic bytes per cell and
    step: fused DUGKS 360 (collide: read 72 + write 72; flux: read ft 72 + f 72, write
    72), reference-form DUGKS 432, stream operator + BGK 144; a temporally blocked pass
    moves 144 per cell for all of its steps, and the pass forms report the bytes per
    cell and step averaged over the timed run."""
    n = args.n
    cells = n * n
    tau = 3.0 * NU
    om_full, om_half = 1.0 / (tau + 0.5), 1.0 / (tau + 0.5)
    lat = P.Lattice(n, n, 2, "f64")
    lat.set_stream(stream)
    lat.equilibrium_init_dev(0, *(t.data_ptr() for t in fields))
    ev = lambda: torch.cuda.Event(enable_timing=True)

    def timed(fn, steps, warm):
        fn(warm)
        torch.cuda.synchronize()
        e0, e1 = ev(), ev()
        e0.record()
        fn(steps)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    def rec(ms, bpc):
        """bpc: algorithmic bytes per cell and step, averaged over the timed call."""
        return {"ms_per_step": round(ms, 5), "mlups": round(cells / ms / 1e3, 1),
                "bytes_per_cell": round(bpc, 2), "gbs": round(bpc * cells / ms / 1e6, 1),
                "frac_of_peak": round(bpc * cells / ms / 1e6 / HBM_PEAK_GBS, 4)}

    def run_bytes(nsteps, ns, pass_b, single_b, last_b):
        """Bytes per cell and step of a run that does nsteps - 1 steps in passes of up to
        ns steps (a remainder of one as a single step) and the last as a single step."""
        total, done = 0, 0
        while done < nsteps - 1:
            k = min(ns, nsteps - 1 - done)
            total += single_b if k == 1 else pass_b
            done += k
        return (total + last_b) / nsteps

    steps = args.next_steps
    out = {"grid": [n, n], "storage": "f64", "steps": steps}
    ms = timed(lambda k: lat.dugks_step(0, 1, k, 1.0, om_full, om_half), steps, 4)
    dns = lat.dugks_get_steps_per_pass()
    out["dugks_fused"] = rec(ms, run_bytes(steps, dns, 144, 360, 360))
    out["dugks_fused"]["steps_per_pass"] = dns
    out["dugks_fused"]["note"] = ("all steps but the last in temporally blocked passes "
                                  "(144 B per cell per pass), the last as collide + flux "
                                  "(360 B per cell)")

    def dugks_ref(k):
        for s in range(k):
            lat.dugks_collide(s % 2, 1 - s % 2, om_full, om_half)
            lat.dugks_stream(s % 2, 1 - s % 2, 1.0)
    out["dugks_reference_form"] = rec(timed(dugks_ref, 10, 2), 432)
    lat.set_params(om_full, MAGIC)
    for op in ("fvm_bardow", "fdm_bardow", "fdm_sofonea"):
        def run(k, op=op):
            for s in range(k):
                lat.step_op(op, "bgk", s % 2, 1 - s % 2, 1.0)
        out[f"stream_{op}+bgk"] = rec(timed(run, steps, 2), 144)
        ons = lat.get_op_steps_per_pass(op, "bgk")
        out[f"stream_{op}+bgk_passes"] = rec(
            timed(lambda k, op=op: lat.run_op(op, "bgk", k, 0, 1, 1.0), steps, 4),
            run_bytes(steps, ons, 144, 144, 144))
        out[f"stream_{op}+bgk_passes"]["steps_per_pass"] = ons
    lat.close()
    torch.cuda.empty_cache()
    return out


def batch_leg(P, torch, args):
    """Uneven-size sweep: 256 periodic cases (n from {32..512}, nu log-uniform in
    [1e-3, 5e-2], half Taylor-Green, half shear layers), BGK, FP32-shifted storage, one
    autotuned launch plan (tile shape or steps per pass, resident threshold), hipGraph
    replay; the same cases in FP64 give the error.  Bytes per lattice update: 72 / steps
    per pass (one read and one write of the FP32 populations per pass)."""
    import numpy as np
    cases = P.sweep_cases(256, seed=4)
    out = {"ncases": len(cases), "steps": args.batch_steps, "storage": "f32s", "model": "bgk"}
    b = P.Batch.from_cases(cases, "f32s", "bgk")
    info = b.info()
    out["cells"] = info["cells"]
    out["tune_ms_per_step"] = b.autotune(10)
    out["launch"] = dict(zip(("ty", "tx", "resident_cells", "resident_block"), b.get_launch()))
    out["launch"]["steps_per_pass"] = b.get_steps_per_pass()
    out["launch"]["pass_columns"] = b.get_pass_columns()
    b.run(20)
    b.sync()
    t0 = time.perf_counter()
    b.run(args.batch_steps)
    b.sync()
    dt = time.perf_counter() - t0
    out["seconds"] = dt
    out["mlups"] = info["cells"] * args.batch_steps / dt / 1e6
    out["bytes_per_lu"] = 72 / out["launch"]["steps_per_pass"]
    out["algorithmic_gbs"] = out["mlups"] * 1e6 * out["bytes_per_lu"] / 1e9
    red32 = b.reduce()
    b.close()
    b64 = P.Batch.from_cases(cases, "f64", "bgk")
    b64.run(20 + args.batch_steps)
    b64.sync()
    red64 = b64.reduce()
    R = P._abi
    errs = []
    for i in range(len(cases)):
        # velocity error through the kinetic-energy-free measure: |sum u^2| relative
        if red64[i, R.RED_NONFINITE] or red32[i, R.RED_NONFINITE]:
            continue
        errs.append(abs(red32[i, R.RED_U2] - red64[i, R.RED_U2]) / max(red64[i, R.RED_U2], 1e-300))
    out["nonfinite_cases_f64"] = int(np.sum(red64[:, R.RED_NONFINITE] > 0))
    out["nonfinite_cases_f32s"] = int(np.sum(red32[:, R.RED_NONFINITE] > 0))
    out["max_rel_diff_sum_u2_f32s_vs_f64"] = float(max(errs)) if errs else None
    b64.close()
    return out


def adjoint_leg(P, torch, args):
    """Adjoint inversion setup: 512^2 BGK, tau = 0.6, 500 steps, checkpoint every 50,
    band-limited initial velocity (|k| <= 8, amplitude 0.02, seed 1), target from seed 2.
    Times forward + reverse sweeps (FP32 and FP64) and checks <grad, d> against FP64
    central differences (h = 1e-6) on 16 random directions and for tau."""
    import numpy as np
    n, steps, every, tau, h = 512, 500, 50, 0.6, 1e-6
    t = lambda a: torch.from_numpy(np.ascontiguousarray(np.ravel(a, order="F"))).cuda().view(n, n)
    ux0, uy0 = (t(a) for a in P.band_limited_field(n, n, 8, 0.02, 1))
    uxt, uyt = (t(a) for a in P.band_limited_field(n, n, 8, 0.02, 2))
    rho0 = torch.ones_like(ux0)
    out = {"grid": [n, n], "steps": steps, "checkpoint_every": every, "tau": tau}
    eng64 = P.Adjoint(n, n, steps, every, "f64")
    gen = torch.Generator(device="cpu").manual_seed(123)
    dirs = [(torch.randn(n, n, generator=gen, dtype=torch.float64).cuda(),
             torch.randn(n, n, generator=gen, dtype=torch.float64).cuda()) for _ in range(16)]
    fds = []
    for dx, dy in dirs:
        lp = eng64.forward(rho0, (ux0 + h * dx).contiguous(), (uy0 + h * dy).contiguous(),
                           1.0 / tau, uxt, uyt)
        lm = eng64.forward(rho0, (ux0 - h * dx).contiguous(), (uy0 - h * dy).contiguous(),
                           1.0 / tau, uxt, uyt)
        fds.append((lp - lm) / (2 * h))
    fd_tau = (eng64.forward(rho0, ux0, uy0, 1.0 / (tau + h), uxt, uyt)
              - eng64.forward(rho0, ux0, uy0, 1.0 / (tau - h), uxt, uyt)) / (2 * h)
    for storage in ("f32s", "f64"):
        eng = eng64 if storage == "f64" else P.Adjoint(n, n, steps, every, storage)
        eng.forward(rho0, ux0, uy0, 1.0 / tau, uxt, uyt)  # warm-up
        eng.backward()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss = eng.forward(rho0, ux0, uy0, 1.0 / tau, uxt, uyt)
        t1 = time.perf_counter()
        _, gux, guy, gom = eng.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        errs = [abs(float(torch.sum(gux * dx) + torch.sum(guy * dy)) - fd) / abs(fd)
                for (dx, dy), fd in zip(dirs, fds)]
        out[storage] = {"loss": loss, "forward_s": t1 - t0, "backward_s": t2 - t1,
                        "max_rel_err_field_dirs_vs_fd64": max(errs),
                        "rel_err_tau_vs_fd64": abs(-gom / tau ** 2 - fd_tau) / abs(fd_tau)}
    return out


def host_cpu():
    """CPU model, the host's physical cores (distinct (package, core) pairs) and logical
    CPUs, and the CPUs this process may run on, from /proc/cpuinfo and the affinity mask."""
    model, pairs, logical = "unknown", set(), 0
    try:
        phys = core = None
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            key, _, val = line.partition(":")
            key, val = key.strip(), val.strip()
            if key == "processor":
                logical += 1
            elif key == "model name" and model == "unknown":
                model = val
            elif key == "physical id":
                phys = val
            elif key == "core id":
                core = val
                pairs.add((phys, core))
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"model": model, "host_physical_cores": len(pairs) or None,
            "host_logical_cpus": logical or os.cpu_count(), "cpus_allowed": allowed}


def cpu_leg(args):
    """The reference Fortran (oracle/_ref, flang -O2 -fopenmp) on the host cores, over a
    bounded sample of the same workload: the 8192^2 TRT grid for --cpu-steps steps, with
    OpenMP over this process's share of the host (OMP_NUM_THREADS; 16 per GPU on a box
    shared per GPU), CPU model and core counts stated."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po
    import __graft_entry__ as g
    P = g.load_package()
    cpu = host_cpu()
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, cpu["cpus_allowed"])
    n = args.n
    rho, ux, uy = initial_fields(P, n)
    if po.have_ref():
        R = po.RefLib()
        grid = R.grid(n, n, "trt")
        # the omega set_properties(NU, dt=1, MAGIC) computes, without its stdout print
        grid.set_omega(1.0 / (3.0 * NU + 0.5), MAGIC)
        grid.set_macros(rho, ux, uy)
        grid.set_pdf_to_equilibrium()
        grid.step(1)  # warm-up
        t0 = time.perf_counter()
        grid.step(args.cpu_steps)
        dt = time.perf_counter() - t0
        grid.close()
        kind = "reference"
    else:
        O = po.Oracle()
        fa, fb = po.new_f(n, n), po.new_f(n, n)
        O.set_pdf_to_equilibrium(n, n, rho, ux, uy, fa)
        omega = 1.0 / (3.0 * NU + 0.5)
        O.run("trt", n, n, fa, fb, 1, omega, MAGIC)
        t0 = time.perf_counter()
        O.run("trt", n, n, fa, fb, args.cpu_steps, omega, MAGIC)
        dt = time.perf_counter() - t0
        kind = "port"
    mlups = n * n * args.cpu_steps / dt / 1e6
    return {"value": round(mlups, 2), "unit": "MLUPS", "cores": cores, "kind": kind,
            "sample": f"D2Q9 TRT {n}x{n} FP64, {args.cpu_steps} steps after 1 warm-up step "
                      f"({'flang-built reference Fortran' if kind == 'reference' else 'C oracle'}, "
                      f"OpenMP {cores} threads), {dt:.1f} s",
            "cpu": cpu}


def traffic_from_profiles(n, ns):
    """HBM bytes per launch of the FP64 TRT kernel that runs `ns` steps per pass, from the
    committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration)."""
    name = f"traffic_trt_f64_{n}.json" if ns == 1 else f"traffic_trt_f64_{n}_pass{ns}.json"
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get("bytes_per_launch")
    except Exception:
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the "
              "multi-GPU slab decomposition is design-only", file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("OMP_NUM_THREADS", str(min(16, os.cpu_count() or 1)))
    results, copy_gbs, dev_name, extras = gpu_leg(args)
    cpu = None if args.no_cpu else cpu_leg(args)
    progress("cpu baseline done")

    f64 = results["f64"]
    n = args.n
    ns = f64["steps_per_pass"]
    dom = f64["dominant_depth"]
    bytes_per_launch = BYTES_PER_LU["f64"] * n * n  # one read + one write per population
    achieved = bytes_per_launch / (f64["dominant_ms"] * 1e-3) / 1e9
    line = {
        "metric": "MLUPS & % HBM roofline, D2Q9 TRT 8192^2: FP64 vs FP32/FP16-shifted at stated err",
        "value": round(f64["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f64["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Taylor-Green u0=0.05 nu=0.01 + U(+-1e-3) noise, seed 7)",
        "config": {"workload": f"D2Q9 TRT Lambda=1/4 periodic {n}x{n}, FP64 populations, "
                               f"fused pull-stream+collide, passes of up to {ns} steps per "
                               f"HBM sweep",
                   "grid": [n, n], "model": "trt", "storage": "f64",
                   "omega": 1.0 / (3.0 * NU + 0.5), "device": dev_name},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(n, dom),
                     "kernel": f"k_multi<TRT, F64, {dom}>" if dom > 1 else "k_fused<TRT, F64>",
                     "launch_ms": round(f64["dominant_ms"], 5),
                     "launches_in_timed_region": f64["dominant_launches"],
                     "lu_per_launch": dom * n * n, "bytes_per_launch": bytes_per_launch,
                     "timed_launches": f64["timed_launches"]},
        "tuned_depth_pass": {
            "kernel": f"k_multi<TRT, F64, {ns}>" if ns > 1 else "k_fused<TRT, F64>",
            "launch_ms": round(f64["pass_ms"], 5), "achieved": round(f64["pass_gbs"], 1),
            "frac": round(f64["pass_gbs"] / HBM_PEAK_GBS, 4), "ms_per_step": round(
                f64["pass_ms"] / ns, 5)},
        "single_step_kernel": {
            "kernel": "k_fused<TRT, F64>", "launch_ms": round(f64["single_step_ms"], 5),
            "achieved": round(f64["single_step_gbs"], 1),
            "frac": round(f64["single_step_gbs"] / HBM_PEAK_GBS, 4),
            "mlups": round(n * n / f64["single_step_ms"] / 1e3, 1)},
        "cpu_baseline": cpu,
        "measured_copy_gbs": round(copy_gbs, 1),
        "frac_of_measured_copy": round(achieved / copy_gbs, 4),
        "single_step_frac_of_measured_copy": round(f64["single_step_gbs"] / copy_gbs, 4),
        # "% roofline" in the lattice-update sense: MLUPS x one step's 144 B per lattice
        # update over the measured copy bandwidth (above 1 once several steps share a sweep)
        "lu_bytes_rate_over_measured_copy": round(
            f64["mlups"] * 1e6 * BYTES_PER_LU["f64"] / 1e9 / copy_gbs, 3),
        "storage_forms": {
            k: {kk: (round(vv, 6) if isinstance(vv, float) and vv > 1e-3 else vv)
                for kk, vv in v.items()} for k, v in results.items()},
        "roofline_by_storage": {
            k: {"steps_per_pass": v["steps_per_pass"],
                "bytes_per_lu": round(BYTES_PER_LU[k] / v["steps_per_pass"], 2),
                "pass_achieved_gbs": round(v["pass_gbs"], 1),
                "pass_frac_of_peak": round(v["pass_gbs"] / HBM_PEAK_GBS, 4),
                "single_step_gbs": round(v["single_step_gbs"], 1),
                "single_step_frac_of_peak": round(v["single_step_gbs"] / HBM_PEAK_GBS, 4)}
            for k, v in results.items()},
        "f32s_speedup_vs_f64": round(results["f32s"]["mlups"] / f64["mlups"], 3),
        "f16s_speedup_vs_f64": round(results["f16s"]["mlups"] / f64["mlups"], 3),
        "config1_taylor_green_64": extras["config1"],
        "config2_roofline_4096": extras["config2"],
        "other_collisions": extras["collisions"],
        "batched_sweep": extras["batched_sweep"],
        "adjoint": extras["adjoint"],
        "schemes": extras["schemes"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
