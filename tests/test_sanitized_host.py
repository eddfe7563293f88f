"""The host layer of the library (argument validation, planning, workspace
carving in csrc/abi.cpp) under AddressSanitizer and UndefinedBehaviorSanitizer:
the CPU boundary tests run in a child process against
_build/libcumm_amd_san.so (built by cumm_amd._build.build_sanitized) with the
ASan runtime preloaded; any sanitizer report aborts the child.  No GPU."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

from cumm_amd import _build

ROOT = Path(__file__).resolve().parent.parent


def test_boundary_tests_clean_under_asan_ubsan():
    rt = _build.asan_runtime()
    if not _build.SAN_LIB.exists() or not rt.is_file():
        pytest.skip("sanitized library or ASan runtime not present (build with cumm_amd._build.build_sanitized)")
    # the host code is instrumented: it calls the ASan and UBSan runtimes
    nm = subprocess.run(["nm", "-D", "--undefined-only", str(_build.SAN_LIB)], capture_output=True, text=True)
    if nm.returncode == 0:
        assert "__asan_report" in nm.stdout and "__ubsan_handle" in nm.stdout
    env = dict(os.environ)
    env.update(CUMM_AMD_LIBRARY=str(_build.SAN_LIB), LD_PRELOAD=str(rt),
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p", "no:cacheprovider",
                        str(ROOT / "tests" / "test_boundary.py")], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=600)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in tail and "runtime error:" not in tail
