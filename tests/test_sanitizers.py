"""The oracle's C restatement under AddressSanitizer and
UndefinedBehaviorSanitizer: oracle/Makefile's liboracle_san.so (every finding
aborts) runs the oracle's own tests in a child process with the ASan runtime
preloaded.  (The library's host-only VLSV writer runs under the same
sanitizers in tests/test_vlsv_host.py.)"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_oracle_tests_under_asan_and_ubsan():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle_san.so"])
    asan = subprocess.check_output(["gcc", "-print-file-name=libasan.so"], text=True).strip()
    # the ASan runtime goes first; whatever the environment preloads stays
    pre = ":".join(p for p in (asan, os.environ.get("LD_PRELOAD", "")) if p)
    env = dict(os.environ, RH_ORACLE_SANITIZED="1", LD_PRELOAD=pre, OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    files = [os.path.join(HERE, f) for f in ("test_oracle_push.py", "test_oracle_fields.py", "test_oracle_deposit.py",
                                             "test_oracle_inject.py", "test_oracle_logs.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider"] + files, env=env,
                       cwd=ROOT, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert " passed" in r.stdout
