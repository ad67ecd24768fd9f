"""The oracle's C restatement under AddressSanitizer and
UndefinedBehaviorSanitizer: oracle/sanitize_main.c drives every entry point
(singles, doubles and their FP64 forms, the exact-exponential and sparse
schemes, the sparse two-exciton step, absorption in both Hamiltonian types
and thread counts, the gradient path, the 1D FFT, and 2D items for every
technique with and without a waiting time); the run must end normally with
no sanitizer report."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")


def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([os.path.join(ORACLE, "_build", "oracle_sanitize")], capture_output=True, text=True,
                       env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert "sanitized oracle run completed" in r.stdout
