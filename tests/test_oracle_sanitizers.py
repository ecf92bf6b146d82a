"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle's C source is built with -fsanitize=address,undefined (host code,
gcc; `make -C oracle sanitize`) and the oracle suite (tests/test_oracle.py)
runs on that build in a child process with the ASan runtime preloaded.  Any
out-of-bounds access, use after free or undefined operation in the oracle
fails the run."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    try:
        path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    except FileNotFoundError:
        return None
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_suite_under_asan_ubsan():
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, NS_ORACLE_LIB=os.path.join(ROOT, "oracle", "_build", "libns_oracle_san.so"),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4")
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_oracle.py"), "-x", "-q",
                        "-p", "no:cacheprovider"], capture_output=True, text=True, env=env, cwd=ROOT, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
