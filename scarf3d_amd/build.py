"""Build recipe of the SRM library for gfx950 (hipcc, in-tree).

Produces scarf3d_amd/lib/libscarf3df.so (real = float) and
scarf3d_amd/lib/libscarf3d.so (real = double, -DDOUBLE_PREC), the two library
flavours of the reference (src/Makefile:9-13). The kernels are compiled once
(they are templated on the field type); only the C ABI layer differs.

Usage: python -m scarf3d_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "lib"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-I", str(CSRC), "-I", str(ROOT / "include")]
# Packed FP32 instructions are kept off for every HIP translation unit.
DEVICE = [f"--offload-arch={ARCH}", "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# The sampler reproduces the reference's roundings: no FMA contraction, and
# sin/cos stay separate libm calls (a fused sincos may differ in the last bit).
HOST_EXACT = ["-O2", "-ffp-contract=off", "-fno-fast-math", "-fno-builtin-sin", "-fno-builtin-cos",
              "-fno-builtin-sinf", "-fno-builtin-cosf", "-fno-builtin-sincos", "-fno-builtin-sincosf"]

UNITS = [
    # (source, extra flags, object name)
    ("srm_kernels.hip", DEVICE, "srm_kernels.o"),
    ("srm_mfma.hip", DEVICE, "srm_mfma.o"),
    ("srm_host.cpp", DEVICE + HOST_EXACT, "srm_host.o"),
    ("quadpack.cpp", DEVICE + HOST_EXACT, "quadpack.o"),
    ("fim.hip", DEVICE, "fim.o"),
    ("fim_host.cpp", DEVICE + HOST_EXACT, "fim_host.o"),
    ("scarf_capi.cpp", DEVICE + HOST_EXACT, "scarf_capi_f.o"),
    ("scarf_capi.cpp", DEVICE + HOST_EXACT + ["-DDOUBLE_PREC"], "scarf_capi_d.o"),
]


def _compile(src: str, flags: list[str], obj: str, verbose: bool) -> None:
    cmd = [HIPCC, "-x", "hip", *COMMON, *flags, "-c", str(CSRC / src), "-o", str(BUILD / obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    msgs = [ln for ln in (res.stdout + res.stderr).splitlines() if noise not in ln]
    if res.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n" + "\n".join(msgs))
    if msgs and verbose:
        print("\n".join(msgs))


def _link(objs: list[str], out: Path, verbose: bool) -> None:
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(out), *[str(BUILD / o) for o in objs],
           "-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"linking {out.name} failed:\n{res.stdout}{res.stderr}")


FC = os.environ.get("FC_SCARF", "/opt/rocm/bin/amdflang")


def build_fortran(verbose: bool = False) -> list[Path]:
    """Fortran module m_scarflib for both library flavours and the test driver
    (tests/fortran/srm_driver.f90), when amdflang is available."""
    if not Path(FC).exists():
        return []
    outs = []
    driver = ROOT / "tests" / "fortran" / "srm_driver.f90"
    for flav, defs, libname in (("f", [], "scarf3df"), ("d", ["-DDOUBLE_PREC"], "scarf3d")):
        moddir = LIB / f"fortran_{flav}"
        moddir.mkdir(parents=True, exist_ok=True)
        obj = moddir / "m_scarflib.o"
        cmds = [[FC, "-cpp", *defs, "-O2", "-fPIC", "-J", str(moddir), "-c", str(HERE / "fortran" / "m_scarflib.f90"),
                 "-o", str(obj)]]
        if driver.exists():
            exe = LIB / f"srm_driver_{flav}"
            cmds.append([FC, "-cpp", *defs, "-O2", "-I", str(moddir), str(driver), str(obj), "-L", str(LIB),
                         f"-l{libname}", f"-Wl,-rpath,{LIB}", "-o", str(exe)])
            outs.append(exe)
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
            res = subprocess.run(cmd, capture_output=True, text=True, cwd=str(moddir))
            if res.returncode != 0:
                raise RuntimeError(f"Fortran build failed:\n{res.stdout}{res.stderr}")
        outs.append(obj)
    return outs


CXX = os.environ.get("CXX_SCARF", "g++")


def build_standalone(verbose: bool = False) -> list[Path]:
    """The standalone program (standalone/scarf3d_main.cpp) for both flavours:
    host code only, linked against the library next to it."""
    outs = []
    src = HERE / "standalone" / "scarf3d_main.cpp"
    for flav, defs, libname in (("f", [], "scarf3df"), ("d", ["-DDOUBLE_PREC"], "scarf3d")):
        exe = LIB / f"scarf3d_standalone_{flav}"
        cmd = [CXX, "-std=c++17", "-O2", "-Wall", *defs, "-I", str(ROOT / "include"), str(src), "-L", str(LIB),
               f"-l{libname}", "-Wl,-rpath,$ORIGIN", "-o", str(exe)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"standalone build failed:\n{res.stdout}{res.stderr}")
        outs.append(exe)
    return outs


CC = os.environ.get("CC_SCARF", "gcc")


def build_drivers(verbose: bool = False) -> list[Path]:
    """The C and C++ drop-in drivers (tests/c/srm_driver.c,
    tests/cpp/srm_driver.cpp) for both flavours: callers written against
    include/scarf3d.h as the reference's examples are, linked with the
    library."""
    outs = []
    jobs = []
    for flav, defs, libname in (("f", [], "scarf3df"), ("d", ["-DDOUBLE_PREC"], "scarf3d")):
        c_src, cpp_src = ROOT / "tests" / "c" / "srm_driver.c", ROOT / "tests" / "cpp" / "srm_driver.cpp"
        if c_src.exists():
            jobs.append(([CC, "-std=c11", "-O2", "-Wall", *defs, "-I", str(ROOT / "include"), str(c_src)],
                         LIB / f"srm_driver_c_{flav}", libname))
        if cpp_src.exists():
            jobs.append(([CXX, "-std=c++17", "-O2", "-Wall", *defs, "-I", str(ROOT / "include"), str(cpp_src)],
                         LIB / f"srm_driver_cpp_{flav}", libname))
    for head, exe, libname in jobs:
        cmd = [*head, "-L", str(LIB), f"-l{libname}", "-Wl,-rpath,$ORIGIN", "-o", str(exe)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"driver build failed:\n{res.stdout}{res.stderr}")
        outs.append(exe)
    return outs


def build(jobs: int = 4, verbose: bool = False) -> list[Path]:
    """Compile every HIP translation unit for gfx950 and link both libraries."""
    BUILD.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = [ex.submit(_compile, s, f, o, verbose) for s, f, o in UNITS]
        for fut in futs:
            fut.result()
    shared = ["srm_kernels.o", "srm_mfma.o", "srm_host.o", "quadpack.o", "fim.o", "fim_host.o"]
    outs = []
    for obj, name in (("scarf_capi_f.o", "libscarf3df.so"), ("scarf_capi_d.o", "libscarf3d.so")):
        out = LIB / name
        tmp = LIB / (name + ".tmp")
        _link([*shared, obj], tmp, verbose)
        os.replace(tmp, out)
        outs.append(out)
    outs.extend(build_standalone(verbose))
    outs.extend(build_drivers(verbose))
    outs.extend(build_fortran(verbose))
    return outs


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args(argv)
    for p in build(args.jobs, args.verbose):
        print(p)
    return 0


if __name__ == "__main__":
    sys.exit(main())
