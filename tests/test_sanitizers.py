"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

* The FP64 oracle (oracle/liborc_asan.so, `make -C oracle asan`) runs its own
  pin tests with libasan preloaded into the Python process, so every index
  the oracle forms is bounds-checked while the restatement is exercised.
* The library's host-side release map and BFGS fit (csrc/release.hpp, plain
  C++) run in a g++ -fsanitize=address,undefined driver.

The device code is not sanitized here (GPU sanitizers are not available);
its bounds are covered by the GPU parity tests and the kernels' own clamps."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


def _libasan():
    gcc = shutil.which("gcc")
    if gcc is None:
        return None
    p = subprocess.run([gcc, "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def asan():
    lib = _libasan()
    if lib is None:
        pytest.skip("gcc's libasan is not installed")
    return lib


def test_oracle_pins_under_asan(asan):
    subprocess.run(["make", "-s", "-C", ORACLE, "asan"], check=True)
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, MAC3D_ORACLE_LIB=os.path.join(ORACLE, "liborc_asan.so"),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="2")
    tests = ["tests/test_oracle_closed_forms.py", "tests/test_wave_oracle.py",
             "tests/test_parcel_oracle.py", "tests/test_conc_oracle.py",
             "tests/test_reference_physics_oracle.py", "tests/test_oracle.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", *tests, "-x", "-q", "-s", "-m", "not gpu",
                        "-p", "no:cacheprovider", "-k", "not contraction_h_independent"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=1500)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]


def test_release_host_code_under_asan(tmp_path, asan):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not installed")
    exe = str(tmp_path / "release_check")
    subprocess.run([gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "sanitize", "release_check.cpp"), "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, ASAN_OPTIONS="abort_on_error=1"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
