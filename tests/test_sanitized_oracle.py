"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

Builds oracle/liborc_san.so (`make -C oracle sanitize`: -fsanitize=address,
undefined with -fno-sanitize-recover=all) and runs the oracle's own tests
against it in a child Python process with libasan preloaded: an
out-of-bounds access, use-after-free or undefined operation anywhere in the
C oracle aborts that process and fails this test.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
TESTS = ["test_oracle_csr.py", "test_oracle_jitconn.py", "test_oracle_neuron.py", "test_oracle_plast.py",
         "test_oracle_rng.py", "test_oracle_coo.py", "test_oracle_bsr.py", "test_golden.py"]


def _libasan():
    try:
        p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True)
    except (OSError, subprocess.CalledProcessError):
        return None
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_tests_pass_under_asan_ubsan():
    asan = _libasan()
    if asan is None:
        pytest.skip("gcc's libasan is not installed")
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    env = dict(os.environ, LD_PRELOAD=asan, BPX_ORC_LIB=os.path.join(ORACLE, "liborc_san.so"),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider"]
    cmd += [os.path.join(ROOT, "tests", t) for t in TESTS]
    p = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    tail = (p.stdout + p.stderr)[-4000:]
    assert p.returncode == 0, tail
    # the child really loaded the sanitized build
    probe = ("import os; from oracle import orc; orc.lib(); "
             "print(any('liborc_san.so' in l for l in open('/proc/self/maps')))")
    q = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True, timeout=120)
    assert q.returncode == 0 and q.stdout.strip().endswith("True"), q.stdout + q.stderr
