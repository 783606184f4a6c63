"""Build libicu_hip.so (HIP kernels for gfx950 + the C ABI) in-tree.

    python -m imagecodingutils_amd.build

Each source compiles to an object under imagecodingutils_amd/_build/ in
parallel, then one shared library is linked next to this file.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
SRC = PKG / "csrc"
OUT = PKG / "libicu_hip.so"
BUILD = PKG / "_build"
SOURCES = ["me_search.hip", "me_decide.hip", "me_vbs.hip", "it_kernels.hip", "capi.cpp", "server.cpp", "vbs.cpp", "mc.cpp", "io.cpp", "cavlc.hip", "me_umhs.hip", "me_cross.hip", "me_fhs.hip", "noalloc.cpp"]

ARCH = os.environ.get("ICU_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # no packed fp32 instructions in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value",
    "-I", str(PKG.parent / "include"),
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: Path, obj: Path) -> None:
    lang = [] if src.suffix == ".hip" else ["-x", "hip"]
    cmd = [hipcc(), *FLAGS, *lang, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if warn:
        print(warn, file=sys.stderr)


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = [BUILD / (s.rsplit(".", 1)[0] + ".o") for s in SOURCES]
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        list(ex.map(lambda p: _compile(SRC / p[0], p[1]), zip(SOURCES, objs)))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build()


def build_dropin_test() -> Path:
    """Compile tests/cpp/dropin_test.cpp (the C++ drop-in classes driven like
    the reference classes) against the built library."""
    src = PKG.parent / "tests" / "cpp" / "dropin_test.cpp"
    exe = src.with_suffix("")
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-pthread", "-I", str(PKG.parent / "include"), str(src), "-o", str(exe),
           "-L", str(PKG), "-licu_hip", "-Wl,-rpath,$ORIGIN/../../imagecodingutils_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"drop-in test build failed:\n{r.stderr}")
    return exe
