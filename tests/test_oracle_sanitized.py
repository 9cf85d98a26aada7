"""The oracle's CPU tests again, with the oracle built under AddressSanitizer and
UndefinedBehaviorSanitizer (host code; any report aborts the run).  Python itself is not
instrumented, so the ASan runtime is preloaded into the child interpreter."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


def _runtime(name):
    cc = shutil.which("gcc")
    if cc is None:
        return None
    p = subprocess.run([cc, "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.mark.timeout(600)
def test_oracle_tests_pass_under_asan_and_ubsan():
    asan = _runtime("libasan.so")
    if asan is None or os.environ.get("LD_PRELOAD"):
        pytest.skip("no ASan runtime, or another preload that ASan must precede")
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    env = dict(os.environ)
    env.update(PTN_ORACLE_LIB=os.path.join(ORACLE, "libptn_oracle_san.so"), LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(ROOT, "tests", "test_oracle.py"), os.path.join(ROOT, "tests", "test_mlp_oracle.py")],
                       env=env, capture_output=True, text=True, cwd=ROOT, timeout=580)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail
