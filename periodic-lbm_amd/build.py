"""Build recipe for the engine (run by __graft_entry__.build(); no cmake, no JIT cache).

Outputs, all in-tree so they travel with the source tree:
  periodic-lbm_amd/libplbm.so          HIP kernels + C ABI (include/plbm.h), gfx950
  periodic-lbm_amd/fortran/_build/     Fortran shim: module files + libplbm_fortran.a
  periodic-lbm_amd/_build/app/         the reference's app/main_taylor_green.f90, compiled
                                       unmodified against the shim (only where the
                                       reference tree is present at build time)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
FSRC = PKG / "fortran"
OBJ = PKG / "_build" / "obj"
FBUILD = FSRC / "_build"
APPDIR = PKG / "_build" / "app"
LIB = PKG / "libplbm.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
ARCH = os.environ.get("PLBM_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["lattice.hip", "batch.hip", "adjoint.hip", "streaming_ops.hip",
               "streaming_ops_fdm.hip", "dugks.hip", "multistep.hip"]
CXX_SOURCES = ["fields.cpp", "sim_plugin.cpp"]
# -ffp-contract=off: the FP64 reference path must not contract a*b+c into an FMA (the
# reference is built without FMA); kernels that want FMA enable it locally by pragma.
# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{ROOT / 'include'}",
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-Wall",
             f"-I{ROOT / 'include'}"]
# Per-source extras.  The temporally blocked pass kernels issue long dependent FP64 chains
# with two wavefronts per SIMD; the ILP-maximising scheduler interleaves them better (the
# arithmetic is unchanged; same-box A/B, TRT 8 steps per pass, GLUPS: 8192^2 FP64
# 208.0 -> 212.4, FP16 storage 398.3 -> 403.6, FP32 386.0 -> 387.9; 4096^2 FP64 187.1 ->
# 191.5; the DUGKS passes 32.3 -> 32.8).  For the batched and FV/FD kernels it measured
# within -0.6 .. +0.3% and is not used.
_MAX_ILP = ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]
HIP_EXTRA = {"multistep.hip": _MAX_ILP, "dugks.hip": _MAX_ILP}
# Fortran shim modules in dependency order
FORTRAN_SOURCES = [
    "plbm_capi.f90", "precision.F90", "fvm_bardow.F90", "periodic_lbm.f90",
    "collision_bgk.F90", "collision_trt.f90", "collision_regularized.f90",
    "taylor_green.f90", "shear_layer.f90", "barotropic_vortex_case.f90", "periodic_dugks.f90",
]


def _run(cmd, cwd=None):
    r = subprocess.run([str(c) for c in cmd], cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(map(str, cmd))}\n"
                           f"{r.stdout}\n{r.stderr}")
    return r


def build_hip(verbose: bool = True) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        jobs.append([HIPCC, *HIP_FLAGS, *HIP_EXTRA.get(s, []), "-c", CSRC / s, "-o",
                     OBJ / (Path(s).stem + ".o")])
    for s in CXX_SOURCES:
        jobs.append(["g++", *CXX_FLAGS, "-c", CSRC / s, "-o", OBJ / (Path(s).stem + ".o")])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    objs = [OBJ / (Path(s).stem + ".o") for s in HIP_SOURCES + CXX_SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print(f"[build] {LIB.relative_to(ROOT)}")
    return LIB


# Preprocessor options of the reference's Fortran build that the shim honours, e.g.
# PLBM_FORTRAN_DEFINES="-DSPLIT" (collide_bgk in split-weight algebra, as the reference's
# src/collision_bgk.F90:23 selects it).
FORTRAN_DEFINES = os.environ.get("PLBM_FORTRAN_DEFINES", "").split()


def build_fortran(verbose: bool = True, outdir: Path | None = None,
                  defines: list[str] | None = None) -> Path:
    """Compile the Fortran module shim (ISO_C_BINDING over libplbm.so) into outdir
    (default fortran/_build) with the given preprocessor options (default
    PLBM_FORTRAN_DEFINES)."""
    outdir = FBUILD if outdir is None else outdir
    defines = FORTRAN_DEFINES if defines is None else defines
    mod = outdir / "mod"
    mod.mkdir(parents=True, exist_ok=True)
    objs = []
    for s in FORTRAN_SOURCES:
        o = outdir / (Path(s).stem + ".o")
        _run([FLANG, "-O2", "-fPIC", "-fopenmp", *defines, "-module-dir", mod,
              f"-I{mod}", "-c", FSRC / s, "-o", o])
        objs.append(o)
    lib = outdir / "libplbm_fortran.a"
    if lib.exists():
        lib.unlink()
    _run(["ar", "rcs", lib, *objs])
    if verbose:
        print(f"[build] {lib.relative_to(ROOT)}")
    return lib


def link_fortran_program(src: Path, out: Path, moddir: Path | None = None,
                         shim: Path | None = None) -> Path:
    """Compile a Fortran program that `use`s the reference module names against the
    shim (default fortran/_build) and libplbm.so.  The internal-procedure logger of the
    reference app needs an executable stack for its trampoline."""
    shim = FBUILD if shim is None else shim
    mod = shim / "mod"
    out.parent.mkdir(parents=True, exist_ok=True)
    pmod = moddir or (out.parent / "mod")
    pmod.mkdir(parents=True, exist_ok=True)
    _run([FLANG, "-O2", "-fopenmp", f"-I{mod}", "-module-dir", pmod, src,
          shim / "libplbm_fortran.a", f"-L{PKG}", "-lplbm", f"-Wl,-rpath,{PKG}",
          "-Wl,-rpath,$ORIGIN/../..", "-Wl,-z,execstack", "-o", out])
    return out


# The shim compiled with the reference's -DDUGKS (DUGKS with face collisions), for the
# DUGKS test program built against it.
FBUILD_DUGKS = FSRC / "_build_dugks"


def build_fortran_tests(verbose: bool = True) -> None:
    """The Fortran test programs under tests/fortran, linked against the shim; and
    dugks_check once more against the shim compiled with -DDUGKS (dugks_check_faces)."""
    out = PKG / "_build" / "tests"
    for name in ("shim_check", "outputs_check", "dugks_check"):
        src = ROOT / "tests" / "fortran" / f"{name}.f90"
        if src.exists():
            link_fortran_program(src, out / name, moddir=out / f"mod_{name}")
            if verbose:
                print(f"[build] {(out / name).relative_to(ROOT)}")
    src = ROOT / "tests" / "fortran" / "dugks_check.f90"
    if src.exists():
        build_fortran(False, FBUILD_DUGKS, [*FORTRAN_DEFINES, "-DDUGKS"])
        link_fortran_program(src, out / "dugks_check_faces", moddir=out / "mod_dugks_faces",
                             shim=FBUILD_DUGKS)
        if verbose:
            print(f"[build] {(out / 'dugks_check_faces').relative_to(ROOT)}")


def build_app(ref_root: Path, verbose: bool = True) -> None:
    """The reference's driver programs, compiled unmodified against the shim."""
    for name in ("main_taylor_green", "main_vortex"):
        src = ref_root / "app" / f"{name}.f90"
        if not src.exists():
            continue
        out = link_fortran_program(src, APPDIR / name, moddir=APPDIR / f"mod_{name}")
        if verbose:
            print(f"[build] {out.relative_to(ROOT)} (unchanged {src.name} on the HIP engine)")


def build_oracle(ref_root: Path | None, verbose: bool = True) -> None:
    _run(["make", "-C", ROOT / "oracle", "liboracle"])
    if verbose:
        print("[build] oracle/liboracle.so")
    if ref_root is not None and (ref_root / "src" / "fvm_bardow.F90").exists():
        _run(["make", "-C", ROOT / "oracle", "ref", f"REF={ref_root}"])
        if verbose:
            print("[build] oracle/_ref (reference Fortran, flang)")


def build_all(verbose: bool = True) -> None:
    ref = Path(os.environ.get("PLBM_REFERENCE", "/root/reference"))
    ref_root = ref if (ref / "src").exists() else None
    build_hip(verbose)
    if shutil.which(FLANG) or Path(FLANG).exists():
        build_fortran(verbose)
        build_fortran_tests(verbose)
        if ref_root is not None:
            build_app(ref_root, verbose)
    build_oracle(ref_root, verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
