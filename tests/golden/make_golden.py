"""Write the golden fixtures from the reference itself.

Runs the reference Fortran hot path (oracle/_ref/libplbm_ref.so, flang-built from the
ivan-pi/periodic-lbm sources by oracle/Makefile) and stores inputs and outputs as small
data files in this directory:

  ref_small.npz   full populations of tiny grids (8x8, 5x3, 16x16, 12x20) after 1, 10
                  and 100 steps of BGK / TRT / RR, plus update_macros of the lagged buffer
  ref_config1.json  the 64x64 Taylor-Green case (ν=0.01, u0=0.05, 1000 steps, Λ=1/4) for
                  BGK / TRT / RR: sha256 of the final two population buffers (bitwise
                  fingerprint), the relative L2 errors vs the analytic solution, mass
  ref_params.json set_properties / lambda_d / magic_number / equilibrium values
  ref_app.json    the unchanged app's printed results (main_taylor_green 1.0)
  ref_streaming_ops.npz / ref_vortex_app.json   FV/FD streaming operators, main_vortex
  ref_outputs.json  sha256 of the VTK / npy / gnuplot files of tests/fortran/outputs_check
  ref_dugks.npz   DUGKS (src/periodic_dugks.F90): both buffers after 1 and 7 steps and
                  the update_macros fields, square and non-square, dt = 1 and 0.5
  ref_dugks_app.json  sha256 of the raw fields tests/fortran/dugks_check.f90 writes
  ref_split.npz   the reference built with -DSPLIT (oracle/_ref/split): collide_bgk
                  (bgk_kernel_cache) and collide_trt (trt_split), 1/10/100 steps
  ref_shear.npz   shear_layer_t's grid evaluation (src/benchmarks/shear_layer.f90:25-52)
                  with the batched sweep's parameters (L = N, k = 80, delta = 0.05,
                  U0 = 0.04) at N = 32, 96, 512 and on a non-square grid with L != N:
                  full fields up to 96^2, the x- and y-profiles and sha256 at 512^2
  ref_fdm_variants.npz  stream_fdm_bardow built with -DFDM_WLS / -DFDM_WLS_GAUSS_V1 /
                  -DFDM_WLS_GAUSS_V2 / -DFDM_ISO (oracle/_ref/fdm_*): perform_step with BGK,
                  5 steps, 16x16 (dt 1, 0.5) and 12x20
  ref_dugks_faces.npz  DUGKS built with -DDUGKS (oracle/_ref/dugks): both buffers after 1
                  and 7 steps and update_macros, the cases of ref_dugks.npz
  ref_dugks_faces_app.json  sha256 of the fields tests/fortran/dugks_check.f90 writes
                  through the -DDUGKS reference modules
  ref_triple.npz  perform_triple_step (src/fvm_bardow.F90:322-340): all three buffers and
                  the ids after 1 and 5 steps, TRT and BGK, 16 x 12 (nf = 3)

Usage (in a container that has the reference tree and flang):
  make -C oracle ref ref-variants && python tests/golden/make_golden.py [--app | --pins-only |
  --variants-only]
"""
from __future__ import annotations

import hashlib
import json
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import pyoracle as po  # noqa: E402

MODELS = ["bgk", "trt", "rr"]


def tgv_init(R, nx, ny, u0, nu):
    k = 2 * np.pi / nx
    p, ux, uy = R.tgv(nx, ny, k, k, u0, nu, 0.0)
    return p / (1.0 / 3.0) + 1.0, ux, uy


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(np.asfortranarray(a)).tobytes(order="F")).hexdigest()


def small(R):
    out = {}
    cases = [(8, 8, 0.05, 0.01), (5, 3, 0.03, 0.02), (16, 16, 0.05, 0.01), (12, 20, 0.04, 0.005)]
    for nx, ny, u0, nu in cases:
        # non-square grids: the reference's TGV eval only covers square grids
        # (src/benchmarks/taylor_green.f90:68), so the field is a seeded random one
        if nx == ny:
            rho, ux, uy = tgv_init(R, nx, ny, u0, nu)
        else:
            rng = np.random.default_rng(nx * 100 + ny)
            rho = np.asfortranarray(1.0 + 0.01 * rng.standard_normal((ny, nx)))
            ux = np.asfortranarray(u0 * rng.standard_normal((ny, nx)))
            uy = np.asfortranarray(u0 * rng.standard_normal((ny, nx)))
        key = f"{nx}x{ny}"
        out[f"{key}_rho"], out[f"{key}_ux"], out[f"{key}_uy"] = rho, ux, uy
        out[f"{key}_params"] = np.array([nu, 1.0, 0.25])
        for m in MODELS:
            g = R.grid(nx, ny, m)
            g.set_properties(nu, 1.0, 0.25)
            g.set_macros(rho, ux, uy)
            g.set_pdf_to_equilibrium()
            done = 0
            for n in (1, 10, 100):
                g.step(n - done)
                done = n
                iold, inew, _, _ = g.indices()
                out[f"{key}_{m}_{n}_latest"] = g.get_f(iold)[:ny]
                g.update_macros()
                r, u, v = g.macros()
                out[f"{key}_{m}_{n}_macros"] = np.stack([r, u, v], axis=2)
            g.close()
    np.savez_compressed(HERE / "ref_small.npz", **out)


def config1(R):
    n, u0, nu, steps = 64, 0.05, 0.01, 1000
    k = 2 * np.pi / n
    res = {"n": n, "u0": u0, "nu": nu, "steps": steps, "magic": 0.25, "dt": 1.0, "models": {}}
    rho, ux, uy = tgv_init(R, n, n, u0, nu)
    _, uxa, uya = R.tgv(n, n, k, k, u0, nu, float(steps))
    for m in MODELS:
        g = R.grid(n, n, m)
        g.set_properties(nu, 1.0, 0.25)
        g.set_macros(rho, ux, uy)
        g.set_pdf_to_equilibrium()
        mass0 = float(np.sum(rho))
        g.step(steps)
        iold, inew, _, _ = g.indices()
        latest, prev = g.get_f(iold)[:n], g.get_f(inew)[:n]
        g.update_macros()
        r, u, v = g.macros()
        l2lag = float(np.sqrt(np.sum((u - uxa) ** 2 + (v - uya) ** 2) / np.sum(uxa ** 2 + uya ** 2)))
        res["models"][m] = {
            "sha256_latest": sha(latest), "sha256_prev": sha(prev),
            "sha256_macros_prev": sha(np.stack([r, u, v], axis=2)),
            "l2_lag": l2lag, "mass0": mass0, "mass_prev": float(np.sum(r)),
            "omega": g.params()["omega"],
            "sample_latest": latest[::8, ::8, :].ravel(order="F").tolist(),
        }
        g.close()
    (HERE / "ref_config1.json").write_text(json.dumps(res, indent=1))


def params(R):
    g = R.grid(4, 4, "trt")
    out = {"set_properties": [], "lambda_d": [], "magic_number": [], "equilibrium": []}
    for nu, dt, magic in [(0.01, 1.0, 0.25), (0.005, 1.0, None), (0.0554, 1.0, 0.25),
                          (0.1, 0.5, None), (1e-3, 2.0, 3 / 16)]:
        g.set_properties(nu, dt, magic)
        out["set_properties"].append({"nu": nu, "dt": dt, "magic": magic, **g.params()})
    for om, x in [(1.9, 0.25), (1.0, 1 / 12), (0.5, 0.25), (1.99, 3 / 16), (1.2, 1.0)]:
        out["lambda_d"].append([om, x, R.lambda_d(om, x)])
    for le, ld in [(1.9, 0.1), (1.0, 1.0), (0.6, 1.7)]:
        out["magic_number"].append([le, ld, R.magic_number(le, ld)])
    rng = np.random.default_rng(3)
    for _ in range(6):
        rho, ux, uy = 1 + 0.1 * rng.standard_normal(), 0.1 * rng.standard_normal(), 0.1 * rng.standard_normal()
        out["equilibrium"].append([rho, ux, uy, R.equilibrium(rho, ux, uy).tolist()])
    g.close()
    (HERE / "ref_params.json").write_text(json.dumps(out, indent=1))


def shear(R):
    out = {}
    cases = [(32, 32, 32.0), (96, 96, 96.0), (512, 512, 512.0), (48, 40, 64.0)]
    u0, k, delta = 0.04, 80.0, 0.05
    for nx, ny, length in cases:
        rho, ux, uy = R.shear(nx, ny, length, u0, k, delta)
        key = f"{nx}x{ny}"
        out[f"{key}_params"] = np.array([length, u0, k, delta])
        if nx * ny <= 96 * 96:
            out[f"{key}_rho"], out[f"{key}_ux"], out[f"{key}_uy"] = rho, ux, uy
        else:
            # ux depends on y only and uy on x only: the profiles plus a fingerprint
            out[f"{key}_ux_col"], out[f"{key}_uy_row"] = ux[:, 0].copy(), uy[0, :].copy()
            out[f"{key}_sha256"] = np.array([sha(rho), sha(ux), sha(uy)])
    np.savez_compressed(HERE / "ref_shear.npz", **out)


def triple(R):
    out = {}
    nx, ny = 16, 12
    rng = np.random.default_rng(2024)
    rho = np.asfortranarray(1.0 + 0.01 * rng.standard_normal((ny, nx)))
    ux = np.asfortranarray(0.04 * rng.standard_normal((ny, nx)))
    uy = np.asfortranarray(0.04 * rng.standard_normal((ny, nx)))
    out["rho"], out["ux"], out["uy"], out["params"] = rho, ux, uy, np.array([0.01, 1.0, 0.25])
    for m in ("trt", "bgk"):
        g = R.grid(nx, ny, m, nf=3)  # all buffers zero before the first step
        g.set_properties(0.01, 1.0, 0.25)
        g.set_macros(rho, ux, uy)
        g.set_pdf_to_equilibrium()
        done = 0
        for n in (1, 5):
            g.triple_step(n - done)
            done = n
            iold, inew, imid, _ = g.indices()
            out[f"{m}_{n}_ids"] = np.array([iold, inew, imid])
            for b in (1, 2, 3):
                out[f"{m}_{n}_buf{b}"] = g.get_f(b)[:ny]
        g.close()
    np.savez_compressed(HERE / "ref_triple.npz", **out)


def app():
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([str(po.REF_APP), "1.0"], cwd=d, capture_output=True, text=True,
                           check=True)
    vals = {}
    for line in r.stdout.splitlines():
        s = line.strip()
        for key in ("L2-norm =", "Final time =", "nsteps =", "omega =", "tc   ="):
            if s.startswith(key):
                vals[key.rstrip(" =").strip()] = float(s.split("=")[1])
    (HERE / "ref_app.json").write_text(json.dumps({"args": ["1.0"], "stdout": r.stdout,
                                                   "values": vals}, indent=1))


def streaming_ops(R):
    """perform_step with the finite-volume / finite-difference streaming operators and
    BGK on a 16x16 Taylor-Green start, dt = 1 and 0.5: populations after 5 steps."""
    out = {}
    n = 16
    rho, ux, uy = tgv_init(R, n, n, 0.05, 0.02)
    for op in ("fvm_bardow", "fdm_bardow", "fdm_sofonea"):
        for dt in (1.0, 0.5):
            g = R.grid(n, n, "bgk")
            g.set_properties(0.02, dt, 0.25)
            g.set_streaming(op)
            g.set_macros(rho, ux, uy)
            g.set_pdf_to_equilibrium()
            g.perform_step(5)
            iold, _, _, _ = g.indices()
            out[f"{op}_dt{dt}"] = g.get_f(iold)[:n]
            out[f"{op}_dt{dt}_omega"] = np.array([g.params()["omega"]])
            g.close()
    out["rho"], out["ux"], out["uy"] = rho, ux, uy
    np.savez_compressed(HERE / "ref_streaming_ops.npz", **out)


FDM_VARIANTS = {"fdm_bardow_wls": "fdm_wls", "fdm_bardow_wls_gauss_v1": "fdm_gauss_v1",
                "fdm_bardow_wls_gauss_v2": "fdm_gauss_v2", "fdm_bardow_iso": "fdm_iso"}


def fdm_variants():
    """stream_fdm_bardow as the reference builds it with -DFDM_WLS, -DFDM_WLS_GAUSS_V1,
    -DFDM_WLS_GAUSS_V2 and -DFDM_ISO (oracle/_ref/<variant>, make -C oracle ref-fdm):
    perform_step with collide_bgk for 5 steps on a 16x16 Taylor-Green start (dt = 1 and
    0.5) and on a seeded 12x20 field (dt = 1)."""
    out = {}
    rng = np.random.default_rng(77)
    fields = {
        "16x16": tgv_init(po.RefLib(), 16, 16, 0.05, 0.02),
        "12x20": (np.asfortranarray(1.0 + 0.01 * rng.standard_normal((20, 12))),
                  np.asfortranarray(0.04 * rng.standard_normal((20, 12))),
                  np.asfortranarray(0.04 * rng.standard_normal((20, 12)))),
    }
    for key, (rho, ux, uy) in fields.items():
        out[f"{key}_rho"], out[f"{key}_ux"], out[f"{key}_uy"] = rho, ux, uy
    for op, d in FDM_VARIANTS.items():
        RV = po.RefLib(po.HERE / "_ref" / d / "libplbm_ref.so")
        for key, dt in (("16x16", 1.0), ("16x16", 0.5), ("12x20", 1.0)):
            nx, ny = map(int, key.split("x"))
            rho, ux, uy = fields[key]
            g = RV.grid(nx, ny, "bgk")
            g.set_properties(0.02, dt, 0.25)
            g.set_streaming("fdm_bardow")
            g.set_macros(rho, ux, uy)
            g.set_pdf_to_equilibrium()
            g.perform_step(5)
            iold, _, _, _ = g.indices()
            out[f"{op}_{key}_dt{dt}"] = g.get_f(iold)[:ny]
            out[f"{op}_{key}_dt{dt}_omega"] = np.array([g.params()["omega"]])
            g.close()
    np.savez_compressed(HERE / "ref_fdm_variants.npz", **out)


def dugks_faces():
    """perform_dugks_step as the reference builds it with -DDUGKS (oracle/_ref/dugks): the
    half-step collision at 0.75 of its rate and face collisions in the stream (rate
    1/(4 tau/dt + 1)); both buffers after 1 and 7 steps and update_macros of f(inew), on
    the cases of ref_dugks.npz."""
    RD = po.RefLib(po.HERE / "_ref" / "dugks" / "libplbm_ref.so")
    out = {}
    cases = [(16, 16, 0.02, 1.0), (12, 20, 0.01, 0.5), (24, 8, 0.005, 2.0)]
    for nx, ny, nu, dt in cases:
        key = f"{nx}x{ny}_dt{dt}"
        rng = np.random.default_rng(nx * 31 + ny + 1)
        rho = np.asfortranarray(1.0 + 0.01 * rng.standard_normal((ny, nx)))
        ux = np.asfortranarray(0.05 * rng.standard_normal((ny, nx)))
        uy = np.asfortranarray(0.05 * rng.standard_normal((ny, nx)))
        out[f"{key}_rho"], out[f"{key}_ux"], out[f"{key}_uy"] = rho, ux, uy
        g = RD.grid(nx, ny, "bgk")
        g.set_properties(nu, dt, 0.25)
        g.set_dugks()
        g.set_macros(rho, ux, uy)
        g.set_pdf_to_equilibrium()
        par = g.params()
        out[f"{key}_params"] = np.array([nu, dt, par["tau"], par["omega"]])
        done = 0
        for n in (1, 7):
            g.dugks_step(n - done)
            done = n
            iold, inew, _, _ = g.indices()
            out[f"{key}_{n}_iold"] = g.get_f(iold)[:ny]
            out[f"{key}_{n}_inew"] = g.get_f(inew)[:ny]
            out[f"{key}_{n}_ids"] = np.array([iold, inew])
            g.update_macros()
            r, u, v = g.macros()
            out[f"{key}_{n}_macros"] = np.stack([r, u, v], axis=2)
        g.close()
    np.savez_compressed(HERE / "ref_dugks_faces.npz", **out)


def vortex_app():
    """The unchanged app/main_vortex.f90 (128^2, stream_fvm_bardow + collide_bgk) run as
    `main_vortex 1.0`: its printout and the sha256 of the files it writes."""
    exe = po.HERE / "_ref" / "main_vortex_ref"
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([str(exe), "1.0"], cwd=d, capture_output=True, text=True, check=True)
        files = {p.name: hashlib.sha256(p.read_bytes()).hexdigest()
                 for p in sorted(Path(d, "vortex").iterdir())}
    (HERE / "ref_vortex_app.json").write_text(json.dumps(
        {"args": ["1.0"], "stdout": r.stdout, "files": files}, indent=1))


def dugks(R):
    """perform_dugks_step with dugks_collide + dugks_stream in the slots."""
    out = {}
    cases = [(16, 16, 0.02, 1.0), (12, 20, 0.01, 0.5), (24, 8, 0.005, 2.0)]
    for nx, ny, nu, dt in cases:
        key = f"{nx}x{ny}_dt{dt}"
        rng = np.random.default_rng(nx * 31 + ny)
        rho = np.asfortranarray(1.0 + 0.01 * rng.standard_normal((ny, nx)))
        ux = np.asfortranarray(0.05 * rng.standard_normal((ny, nx)))
        uy = np.asfortranarray(0.05 * rng.standard_normal((ny, nx)))
        out[f"{key}_rho"], out[f"{key}_ux"], out[f"{key}_uy"] = rho, ux, uy
        g = R.grid(nx, ny, "bgk")
        g.set_properties(nu, dt, 0.25)
        g.set_dugks()
        g.set_macros(rho, ux, uy)
        g.set_pdf_to_equilibrium()
        par = g.params()
        out[f"{key}_params"] = np.array([nu, dt, par["tau"], par["omega"]])
        done = 0
        for n in (1, 7):
            g.dugks_step(n - done)
            done = n
            iold, inew, _, _ = g.indices()
            out[f"{key}_{n}_iold"] = g.get_f(iold)[:ny]
            out[f"{key}_{n}_inew"] = g.get_f(inew)[:ny]
            out[f"{key}_{n}_ids"] = np.array([iold, inew])
            g.update_macros()
            r, u, v = g.macros()
            out[f"{key}_{n}_macros"] = np.stack([r, u, v], axis=2)
        g.close()
    np.savez_compressed(HERE / "ref_dugks.npz", **out)


def dugks_app():
    """sha256 of the raw fields that tests/fortran/dugks_check.f90 writes through the
    reference modules (oracle/_ref/dugks_check_ref)."""
    exe = po.HERE / "_ref" / "dugks_check_ref"
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([str(exe)], cwd=d, capture_output=True, text=True, check=True)
        files = {p.name: hashlib.sha256(p.read_bytes()).hexdigest()
                 for p in sorted(Path(d, "out").iterdir())}
    (HERE / "ref_dugks_app.json").write_text(json.dumps({"stdout": r.stdout, "files": files},
                                                        indent=1))


def dugks_faces_app():
    """sha256 of the raw fields that tests/fortran/dugks_check.f90 writes through the
    reference modules built with -DDUGKS (oracle/_ref/dugks/dugks_check_ref)."""
    exe = po.HERE / "_ref" / "dugks" / "dugks_check_ref"
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([str(exe)], cwd=d, capture_output=True, text=True, check=True)
        files = {p.name: hashlib.sha256(p.read_bytes()).hexdigest()
                 for p in sorted(Path(d, "out").iterdir())}
    (HERE / "ref_dugks_faces_app.json").write_text(
        json.dumps({"stdout": r.stdout, "files": files}, indent=1))


def split():
    """perform_lbm_step with the -DSPLIT build's collide_bgk / collide_trt."""
    RS = po.RefLib(po.LIBREF_SPLIT)
    out = {}
    for nx, ny in [(16, 16), (12, 20)]:
        rng = np.random.default_rng(nx * 5 + ny)
        rho = np.asfortranarray(1.0 + 0.01 * rng.standard_normal((ny, nx)))
        ux = np.asfortranarray(0.05 * rng.standard_normal((ny, nx)))
        uy = np.asfortranarray(0.05 * rng.standard_normal((ny, nx)))
        key = f"{nx}x{ny}"
        out[f"{key}_rho"], out[f"{key}_ux"], out[f"{key}_uy"] = rho, ux, uy
        for m in ("bgk", "trt"):
            g = RS.grid(nx, ny, m)
            g.set_properties(0.01, 1.0, 0.25)
            g.set_macros(rho, ux, uy)
            g.set_pdf_to_equilibrium()
            out[f"{key}_omega"] = np.array([g.params()["omega"]])
            done = 0
            for n in (1, 10, 100):
                g.step(n - done)
                done = n
                iold, _, _, _ = g.indices()
                out[f"{key}_{m}_{n}_latest"] = g.get_f(iold)[:ny]
            g.close()
    np.savez_compressed(HERE / "ref_split.npz", **out)


def outputs():
    """sha256 of the files tests/fortran/outputs_check.f90 writes through the reference's
    output_vtk / output_npy / output_gnuplot (oracle/_ref/outputs_check_ref)."""
    exe = po.HERE / "_ref" / "outputs_check_ref"
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([str(exe)], cwd=d, capture_output=True, text=True, check=True)
        rec = {p.name: hashlib.sha256(p.read_bytes()).hexdigest()
               for p in sorted(Path(d, "out").iterdir())}
    (HERE / "ref_outputs.json").write_text(json.dumps(rec, indent=1))


if __name__ == "__main__":
    R = po.RefLib()
    if "--split-only" in sys.argv:
        split()
        sys.exit(0)
    if "--dugks-only" in sys.argv:
        dugks(R)
        dugks_app()
        sys.exit(0)
    if "--streaming-only" in sys.argv:
        streaming_ops(R)
        vortex_app()
        sys.exit(0)
    if "--variants-only" in sys.argv:
        fdm_variants()
        dugks_faces()
        dugks_faces_app()
        sys.exit(0)
    if "--pins-only" in sys.argv:
        shear(R)
        triple(R)
        sys.exit(0)
    if "--outputs-only" not in sys.argv:
        shear(R)
        triple(R)
        fdm_variants()
        dugks_faces()
        dugks_faces_app()
        params(R)
        small(R)
        config1(R)
        streaming_ops(R)
        dugks(R)
        dugks_app()
        split()
        if "--app" in sys.argv:
            app()
            vortex_app()
    outputs()
    print("golden fixtures written to", HERE)
