"""The CPU-side native code under AddressSanitizer and UndefinedBehaviorSanitizer: the oracle (oracle/liborc.so, GMP
bridge included), the host build of the product library's arithmetic header (tests/native/arith_host.c) and of the e
pipeline's reciprocal and quotient (tests/native/ebig_host.cpp) are rebuilt with -fsanitize=address,undefined and their
CPU test suites run again in a child process with the sanitizer runtimes preloaded; any sanitizer report fails the
run (halt on error).  Host code only: no GPU code is built with a sanitizer."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def runtime(name):
    r = subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.skipif(runtime("libasan.so") is None or runtime("libubsan.so") is None,
                    reason="the sanitizer runtimes are not installed")
def test_native_host_code_clean_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liborc_san.so"], check=True)
    env = dict(os.environ)
    env.update({"ORC_SANITIZE": "1", "NTTM_NATIVE_SANITIZE": "1",
                "LD_PRELOAD": runtime("libasan.so") + ":" + runtime("libubsan.so"),
                "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0:verify_asan_link_order=0",
                "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"})
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
           os.path.join(ROOT, "tests", "test_oracle.py"), os.path.join(ROOT, "tests", "test_arith_host.py"),
           os.path.join(ROOT, "tests", "test_ebig_host.py")]
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert " passed" in out
