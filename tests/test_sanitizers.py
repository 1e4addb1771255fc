"""The CPU code under AddressSanitizer and UndefinedBehaviorSanitizer.

The oracle (oracle/ora_riemann.c, oracle/ora_runner.c) and the host tessellator
(hydro_rs_amd/csrc/host/voronoi.cpp) are rebuilt with gcc/g++ -fsanitize=address,undefined (every report
fatal) into a temporary folder, and a child pytest runs the CPU tests that drive them (the oracle's
golden vectors, pins and whole runs, the tessellator's 1D/2D/3D and partial builds, the snapshot layout)
against those builds, with the ASan runtime preloaded into the child's Python.  Any out-of-bounds access,
use after free or undefined operation in either library fails this test.  The HIP library is not
involved: the child runs no test that loads it.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
# the CPU test files that exercise the two libraries and nothing of the device library
CHILD_TESTS = ["tests/test_oracle_riemann.py", "tests/test_oracle_pins.py", "tests/test_oracle_step.py",
               "tests/test_voronoi.py", "tests/test_snapshot.py"]


def _runtime():
    path = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="gcc/g++ absent")
def test_cpu_libraries_clean_under_asan_ubsan(tmp_path):
    asan = _runtime()
    if asan is None:
        pytest.skip("the ASan runtime is not installed")
    ora = tmp_path / "liboracle_san.so"
    host = tmp_path / "libhfv_host_san.so"
    subprocess.check_call(["gcc", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                           *SAN, "-o", str(ora), os.path.join(ROOT, "oracle", "ora_riemann.c"),
                           os.path.join(ROOT, "oracle", "ora_runner.c"), "-lm"])
    subprocess.check_call(["g++", "-std=c++17", "-fPIC", "-shared", "-fopenmp", *SAN, "-o", str(host),
                           os.path.join(ROOT, "hydro_rs_amd", "csrc", "host", "voronoi.cpp")])
    env = dict(os.environ)
    env.update({"ORA_LIB": str(ora), "HFV_HOST_LIB": str(host), "LD_PRELOAD": asan,
                "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=0:halt_on_error=1",
                "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1", "OMP_NUM_THREADS": "4"})
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        *CHILD_TESTS], cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout[-4000:] + r.stderr[-4000:]
    assert r.returncode == 0, out
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out
    assert " passed" in r.stdout, out
