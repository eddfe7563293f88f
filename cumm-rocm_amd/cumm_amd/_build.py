"""AOT build of libcumm_amd.so for gfx950 with hipcc (no JIT, no torch extension).

Every translation unit under ``csrc/`` is compiled in parallel into
``_build/`` and linked into ``cumm_amd/libcumm_amd.so`` next to this file, so
the library travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
ROOT = PKG_DIR.parent
CSRC = ROOT / "csrc"
BUILD = ROOT / "_build"
LIB = PKG_DIR / "libcumm_amd.so"
INCLUDE = ROOT.parent / "include"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed-FP32 VALU ops anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{CSRC}", f"-I{INCLUDE}",
]


def sources():
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"hipcc failed for {src.name}:\n{msg}")
    return obj


def build(verbose: bool = False, jobs: int = 0) -> Path:
    """Compile every csrc translation unit for gfx950 and link the shared library."""
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


# Host-sanitizer build: the host translation unit (abi.cpp: argument
# validation, planning, workspace carving) with AddressSanitizer and
# UndefinedBehaviorSanitizer, linked with the kernels' objects from build()
# into _build/libcumm_amd_san.so.  Device code is never instrumented (each
# -fsanitize= sits behind -Xarch_host).  Load it with CUMM_AMD_LIBRARY and the
# ASan runtime preloaded (asan_runtime()); tests/test_sanitized_host.py does.
SAN_LIB = BUILD / "libcumm_amd_san.so"
SAN_FLAGS = [
    "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
    "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer",
]


def asan_runtime() -> Path:
    """The clang ASan runtime matching hipcc (empty path if absent)."""
    found = sorted(Path("/opt/rocm/llvm/lib/clang").glob("*/lib/linux/libclang_rt.asan-x86_64.so"))
    return found[-1] if found else Path()


def build_sanitized(verbose: bool = False) -> Path:
    objs = [BUILD / (s.stem + ".o") for s in sources() if s.name != "abi.cpp"]
    if not all(o.exists() for o in objs):
        build(verbose=verbose)
    san_obj = BUILD / "abi_san.o"
    cmd = [HIPCC, *[f for f in COMMON if f != "-O3"], "-O1", "-g", *SAN_FLAGS, "-c", str(CSRC / "abi.cpp"),
           "-o", str(san_obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for abi.cpp (sanitized):\n{r.stderr}")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *SAN_FLAGS[:4], str(san_obj), *map(str, objs),
           "-o", str(SAN_LIB)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"sanitized link failed:\n{r.stderr}")
    if verbose:
        print(f"built {SAN_LIB}", file=sys.stderr)
    return SAN_LIB


if __name__ == "__main__":
    build(verbose=True)
    build_sanitized(verbose=True)
