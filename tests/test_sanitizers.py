"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

* The oracle (oracle/Makefile `sanitize`: g++ -fsanitize=address,undefined) runs the oracle test
  files in a child Python process that preloads GCC's ASan runtime.
* The C-ABI host code (incompflow-cuda_amd/csrc/Makefile `sanitize`: the capi*.hip sources with
  the sanitizers on their host pass, linked with the ordinary kernel objects) runs the host-side
  tests (argument and configuration checks, struct layouts, exports) in a child process that
  preloads clang's ASan runtime.  Without a GPU only those paths run; the device paths run in the
  GPU tests with the ordinary build.

Both builds use -fno-sanitize-recover=all, so any report aborts the child and fails the test.
The child first proves that the sanitized library, not the ordinary one, is the one mapped.
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SAN = os.path.join(ROOT, "oracle", "_san", "liboracle_san.so")
CSRC = os.path.join(ROOT, "incompflow-cuda_amd", "csrc")
LIB_SAN = os.path.join(CSRC, "_san", "libincompflow_san.so")

# the two slowest oracle tests (converged-solve finite differences) add no code path the others miss
SLOW = ["tests/test_oracle_adjoint.py::test_gradient_matches_central_differences_of_converged_solves",
        "tests/test_oracle_adjoint.py::test_stokes_limit_identities"]

CHECK_MAPPED = r"""
import sys
sys.path[:0] = [{root!r}, {pkg!r}]
mod = __import__({mod!r}, fromlist=["lib"])
mod.lib()
maps = open("/proc/self/maps").read()
assert {san!r} in maps, "sanitized library not mapped"
assert "asan" in maps, "ASan runtime not mapped"
print("mapped", {san!r})
"""


def _run(env, args, timeout):
    p = subprocess.run([sys.executable] + args, cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=timeout)
    assert p.returncode == 0, (p.stdout[-3000:] + "\n" + p.stderr[-6000:])
    return p.stdout


def _env(preload, **extra):
    env = dict(os.environ)
    env["LD_PRELOAD"] = preload
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"   # CPython itself keeps allocations at exit
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env.update(extra)
    return env


def test_oracle_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    rt = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isfile(rt):
        pytest.skip("GCC's ASan runtime is not installed")
    env = _env(rt, IFL_ORACLE_LIB=ORACLE_SAN)
    code = CHECK_MAPPED.format(root=ROOT, pkg=os.path.join(ROOT, "incompflow-cuda_amd"), mod="oracle.oracle",
                               san="liboracle_san.so")
    _run(env, ["-c", code], 120)
    files = sorted(glob.glob(os.path.join(ROOT, "tests", "test_oracle_*.py")))
    args = ["-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"] + files
    for t in SLOW:
        args += ["--deselect", t]
    out = _run(env, args, 900)
    assert " passed" in out and "failed" not in out


def test_capi_host_code_under_asan_ubsan():
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc is not installed")
    subprocess.run(["make", "-s", "-j", "8", "-C", CSRC, "sanitize"], check=True, capture_output=True)
    rts = glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    if not rts:
        pytest.skip("clang's ASan runtime is not installed")
    env = _env(rts[0], IFL_LIBRARY=LIB_SAN)
    code = CHECK_MAPPED.format(root=ROOT, pkg=os.path.join(ROOT, "incompflow-cuda_amd"), mod="incompflow",
                               san="libincompflow_san.so")
    _run(env, ["-c", code], 120)
    out = _run(env, ["-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                     os.path.join(ROOT, "tests", "test_capi_host.py")], 300)
    assert " passed" in out and "failed" not in out
