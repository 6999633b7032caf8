"""Host-code sanitizer check (CPU only, no GPU needed).

Builds the library with AddressSanitizer and UndefinedBehaviorSanitizer on
its host code (parameter validation, CA mesh, error estimates, tuning, the
C ABI) into espresso_amd/lib/sanitize/, then runs tests/test_host.py against
that build with the sanitizer runtime preloaded; any report fails the run.

    python espresso_amd/check_sanitizers.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)


def asan_runtime():
    from espresso_amd.build import hipcc
    out = subprocess.run([hipcc(), "-print-file-name=libclang_rt.asan-x86_64.so"],
                         capture_output=True, text=True, check=True).stdout.strip()
    if not os.path.isabs(out) or not os.path.exists(out):
        raise RuntimeError("AddressSanitizer runtime of the ROCm clang not found")
    return out


def main():
    from espresso_amd import build as B
    lib = B.build(sanitize=True)
    env = dict(os.environ)
    env.update(ESPRESSO_AMD_LIB=lib, ESPRESSO_AMD_NO_TORCH="1",
               LD_PRELOAD=asan_runtime(),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    cmd = [sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_host.py"),
           "-q", "-m", "not gpu", "-p", "no:cacheprovider"]
    r = subprocess.run(cmd, env=env, cwd=ROOT)
    sys.exit(r.returncode)


if __name__ == "__main__":
    main()
