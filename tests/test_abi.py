"""C ABI checks that need no GPU: the library loads, exports every function that
include/plbm.h declares, rejects bad arguments before touching the device, and its
host-side routines (flow cases, generator, TRT rate helpers) equal the oracle."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT


def header_functions():
    text = (ROOT / "include" / "plbm.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(plbm_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree(P):
    declared = header_functions()
    assert declared == P._abi.exported_symbols()


def test_library_exports_every_declared_symbol(P):
    L = P._abi.lib()
    for name in header_functions():
        assert hasattr(L, name), name


def test_argument_errors_without_device(P):
    L = P._abi.lib()
    h = ctypes.c_void_p()
    assert L.plbm_create(0, 4, 2, 0, ctypes.byref(h)) == P._abi.ERR_ARG
    assert b"positive" in L.plbm_last_error()
    assert L.plbm_create(4, 4, 9, 0, ctypes.byref(h)) == P._abi.ERR_ARG
    assert L.plbm_create(4, 4, 2, 7, ctypes.byref(h)) == P._abi.ERR_ARG
    assert L.plbm_step(None, 0, 0, 1) == P._abi.ERR_ARG
    assert L.plbm_sync(None) == P._abi.ERR_ARG
    assert L.plbm_destroy(None) == P._abi.OK
    with pytest.raises(P.PlbmError):
        P._abi.check(L.plbm_create(-3, 4, 2, 0, ctypes.byref(h)))


def test_trt_rate_helpers_equal_oracle(P, oracle):
    for om, x in [(1.9, 0.25), (1.0, 1 / 12), (0.5, 3 / 16), (1.99, 1.0)]:
        assert P.lambda_d(om, x) == oracle.lambda_d(om, x)
        assert P.magic_number(om, 0.7) == oracle.magic_number(om, 0.7)


@pytest.mark.parametrize("n,u0,nu,t", [(64, 0.05, 0.01, 0.0), (64, 0.05, 0.01, 1000.0),
                                       (320, 0.01 / 3 ** 0.5, 0.018475208614068, 48658.0)])
def test_taylor_green_field_equals_oracle(P, oracle, n, u0, nu, t):
    k = 2 * np.pi / n
    for a, b in zip(P.TaylorGreen(n, n, k, k, u0, nu).eval(t), oracle.tgv(n, n, k, k, u0, nu, t)):
        np.testing.assert_array_equal(a, b)


def test_shear_layer_field_equals_reference(P):
    """The engine's shear-layer field (config 4's initial state) against the reference's
    own evaluation (ref_shear.npz, src/benchmarks/shear_layer.f90:25-52), bit for bit."""
    import hashlib
    G = np.load(ROOT / "tests" / "golden" / "ref_shear.npz")
    sha = lambda a: hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()
    for key in ("32x32", "96x96", "512x512", "48x40"):
        nx, ny = map(int, key.split("x"))
        length, u0, k, delta = G[f"{key}_params"]
        rho, ux, uy = P.ShearLayer(length, u0, k, delta).eval(nx, ny)
        if f"{key}_ux" in G:
            for a, name in ((rho, "rho"), (ux, "ux"), (uy, "uy")):
                np.testing.assert_array_equal(a, G[f"{key}_{name}"])
        else:
            assert [sha(rho), sha(ux), sha(uy)] == G[f"{key}_sha256"].tolist()


def test_noise_published_splitmix64_vector(P):
    """plbm_noise(seed, c) = output c + 1 of SplitMix64 from state `seed`, top 53 bits:
    the published outputs from state 1234567 (Rosetta Code, Pseudo-random numbers/
    Splitmix64)."""
    outs = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    for c, z in enumerate(outs):
        assert P.noise(1234567, c) == (z >> 11) * 2.0 ** -53


def test_shear_layer_field_equals_oracle(P, oracle):
    for n in (32, 96, 512):
        a = P.ShearLayer(float(n), 0.04, 80.0, 0.05).eval(n, n)
        b = oracle.shear(n, n, float(n), 0.04, 80.0, 0.05)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


def test_noise_equals_oracle(P, oracle):
    for seed in (1, 7, 12345):
        for c in (0, 1, 2, 99, 2 ** 33):
            assert P.noise(seed, c) == oracle.noise(seed, c)
    ux = np.zeros((4, 3), order="F")
    uy = np.zeros((4, 3), order="F")
    P.add_noise(ux, uy, 7, 1e-3)
    flat = np.ravel(ux, order="F")
    assert flat[5] == 1e-3 * (2 * oracle.noise(7, 10) - 1)
    assert np.ravel(uy, order="F")[5] == 1e-3 * (2 * oracle.noise(7, 11) - 1)


def test_fortran_shim_built():
    fb = ROOT / "periodic-lbm_amd" / "fortran" / "_build"
    if not (fb / "libplbm_fortran.a").exists():
        pytest.skip("Fortran shim not built (no flang at build time)")
    for mod in ("fvm_bardow", "periodic_lbm", "collision_bgk", "collision_trt",
                "collision_regularized", "taylor_green", "precision"):
        assert (fb / "mod" / f"{mod}.mod").exists(), mod


def test_sim_plugin_protocol_exported(P):
    """c_plbm_{init,step,vars,free,norm} (include/plbm_sim.h) for the reference's ctypes
    plugin loader; norm needs no GPU."""
    import ctypes
    text = (ROOT / "include" / "plbm_sim.h").read_text()
    names = sorted(set(re.findall(r"\b(c_plbm_[a-z]+)\s*\(", text)))
    assert names == ["c_plbm_free", "c_plbm_init", "c_plbm_norm", "c_plbm_step", "c_plbm_vars"]
    plugin = P.sim.SimPlugin()
    u = np.asfortranarray(np.arange(12.0).reshape(4, 3))
    ua = np.asfortranarray(u + 0.5)
    assert plugin.my_norm(u, ua) == pytest.approx(np.linalg.norm(u - ua) / np.linalg.norm(ua),
                                                  rel=1e-15)
    # dt != 1 is refused before any device work, as the reference plugin does
    p0 = np.zeros((4, 3), order="F")
    u0 = np.zeros((4, 3, 2), order="F")
    with pytest.raises(RuntimeError):
        plugin.init((4, 3), 0.5, p0, u0)


@pytest.mark.parametrize("define,op", [("", 2), ("-DFDM_WLS", 4), ("-DFDM_WLS_GAUSS_V1", 5),
                                       ("-DFDM_WLS_GAUSS_V2", 6), ("-DFDM_ISO", 7)])
def test_fortran_shim_fdm_stencil_options(tmp_path, define, op):
    """The shim's fvm_bardow module compiles with each of the reference's stream_fdm_bardow
    build options (src/fvm_bardow.F90:591-669) and binds stream_fdm_bardow to the engine's
    operator of that stencil (the FDM_BARDOW_OP constant after preprocessing)."""
    import shutil
    import subprocess
    flang = "/opt/rocm/lib/llvm/bin/flang"
    if not Path(flang).exists():
        pytest.skip("flang not available")
    src = ROOT / "periodic-lbm_amd" / "fortran"
    args = [flang, "-E", "-cpp"] + ([define] if define else []) + [str(src / "fvm_bardow.F90")]
    pre = subprocess.run(args, capture_output=True, text=True, check=True).stdout
    pre = re.sub(r"&[ \t]*\n[ \t]*&?", "", pre)  # join free-form continuation lines
    names = {2: "PLBM_STREAM_FDM_BARDOW", 4: "PLBM_STREAM_FDM_BARDOW_WLS",
             5: "PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V1", 6: "PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V2",
             7: "PLBM_STREAM_FDM_BARDOW_ISO"}
    decl = [l for l in pre.splitlines() if "FDM_BARDOW_OP =" in l]
    assert len(decl) == 1 and decl[0].split("=")[1].strip() == names[op]
    text = (src / "plbm_capi.f90").read_text()
    assert f"{names[op]} = {op}" in text
    for mod in ("plbm_capi.f90", "precision.F90"):
        shutil.copy(src / mod, tmp_path / mod)
    for f in ("plbm_capi.f90", "precision.F90", "fvm_bardow.F90"):
        cmd = [flang, "-c", "-module-dir", str(tmp_path), f"-I{tmp_path}"] + \
              ([define] if define else []) + [str(src / f), "-o", str(tmp_path / (f + ".o"))]
        subprocess.run(cmd, capture_output=True, text=True, check=True)
