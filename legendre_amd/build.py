"""Build recipe for the HIP backend: compiles legendre_amd/csrc/*.hip for gfx950
into the in-tree shared library legendre_amd/liblegendre_gpu.so.

    python -m legendre_amd.build            # build (always recompiles)

Flags that matter for parity:
  -ffp-contract=off   the reference (Rust) never fuses a*b+c; explicit fma() calls
                      stand for its mul_add, nothing else may be contracted.
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keeps the compiler from emitting packed-FP32 VALU ops.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "liblegendre_gpu.so"
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and pathlib.Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP backend cannot be built")


# Host-only sources (*.cpp: the Parquet writer) build with the host compiler.
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-pthread"]


def _compile(src: pathlib.Path, obj: pathlib.Path) -> None:
    if src.suffix == ".cpp":
        cmd = [os.environ.get("CXX", "g++"), *CXX_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [hipcc(), *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stderr}")
    if res.stderr.strip():
        # Surface real warnings; the host pass's note about the packed-fp32
        # target feature is expected and harmless.
        lines = [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            print("\n".join(lines), file=sys.stderr)


def build(verbose: bool = False) -> pathlib.Path:
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    objs = [objdir / (s.name + ".o") for s in sources]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        list(ex.map(lambda so: _compile(*so), zip(sources, objs)))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *map(str, objs), "-o",
           str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


HARNESS_SRC = HERE.parent / "tests" / "cabi" / "harness.cpp"
HARNESS = HERE.parent / "tests" / "cabi" / "harness"


def build_cabi_harness(verbose: bool = False) -> pathlib.Path:
    """The C++ caller of the C ABI used by the tests (host code only, linked against
    the in-tree library with a relative rpath so the tree can move)."""
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", str(HARNESS_SRC),
           f"-I{HERE.parent / 'include'}", f"-L{HERE}", "-llegendre_gpu",
           "-Wl,-rpath,$ORIGIN/../../legendre_amd", "-o", str(HARNESS)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"C ABI harness build failed:\n{res.stderr}")
    if verbose:
        print(f"built {HARNESS}")
    return HARNESS


if __name__ == "__main__":
    build(verbose=True)
    build_cabi_harness(verbose=True)
