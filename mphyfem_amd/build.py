"""Build libmpf.so (HIP, gfx950) in-tree, and the CPU oracle used by the tests.

    python -m mphyfem_amd.build            # product library + oracle
    python -m mphyfem_amd.build --lib      # product library only

Every .hip file under csrc/ is compiled for gfx950 with packed-FP32 instructions disabled,
host-only .cpp files with the same compiler, and the objects are linked into
mphyfem_amd/libmpf.so.  The oracle (oracle/oracle.cpp) is built with g++ into
oracle/liboracle.so using the reference's optimisation flags.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libmpf.so"
HOST_CALLER_SRC = ROOT / "tests" / "host_caller.cpp"
HOST_CALLER = ROOT / "tests" / "host_caller"
ORACLE_SRC = ROOT / "oracle" / "oracle.cpp"
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-Wno-unused-but-set-variable", "-Wno-unused-result"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"command failed: {cmd[0]} ... {cmd[-1]}")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ("_hip" if src.suffix == ".hip" else "_cpp") + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *COMMON, "-x", "c++", "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build_lib(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_host_caller() -> Path:
    """The compiled C++ caller of include/mpf_host.hpp (a test of the boundary), linked to libmpf.so."""
    gxx = shutil.which("g++") or "g++"
    tmp = HOST_CALLER.with_name(HOST_CALLER.name + ".tmp")
    _run([gxx, "-std=c++17", "-O2", "-Wall", "-I", str(ROOT / "include"), str(HOST_CALLER_SRC), "-o", str(tmp),
          "-L", str(PKG), "-lmpf", "-Wl,-rpath,$ORIGIN/../mphyfem_amd", "-Wl,-rpath-link,/opt/rocm/lib"])
    os.replace(tmp, HOST_CALLER)
    return HOST_CALLER


def build_oracle() -> Path:
    gxx = shutil.which("g++") or "g++"
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run([gxx, "-std=c++17", "-O3", "-fopenmp", "-frounding-math", "-fPIC", "-shared",
          "-o", str(tmp), str(ORACLE_SRC)])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="build the product library only")
    ap.add_argument("--oracle", action="store_true", help="build the oracle only")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    if not a.oracle:
        print("built", build_lib(a.j))
        print("built", build_host_caller())
    if not a.lib:
        print("built", build_oracle())
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
