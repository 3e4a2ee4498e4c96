"""The CPU tests of the host library and of the oracle, run again against
their AddressSanitizer + UndefinedBehaviorSanitizer builds (make sanitize in
differentialequations.jl_amd/csrc and in oracle/): any invalid access,
leak-free use aside, or undefined behaviour aborts the child run."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "differentialequations.jl_amd", "csrc")
HOST_SAN = os.path.join(REPO, "differentialequations.jl_amd", "_build", "san", "libmideq.so")
ORACLE_SAN = os.path.join(REPO, "oracle", "_build", "libmideq_oracle_san.so")


def runtime(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    return r.stdout.strip() if r.returncode == 0 else ""


def child(env_extra, tests):
    env = dict(os.environ)
    env.update(env_extra)
    # Python itself is not instrumented: leak reports would be its own
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    cmd = [sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu"]
    cmd += [os.path.join(REPO, "tests", t) for t in tests]
    r = subprocess.run(cmd, cwd=os.path.join(REPO, "tests"), env=env, capture_output=True,
                       text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "AddressSanitizer" not in tail and "runtime error" not in tail, tail
    return tail


@pytest.mark.slow
def test_oracle_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "sanitize"], check=True)
    asan = runtime(["gcc", "-print-file-name=libasan.so"])
    if not asan or not os.path.isabs(asan):
        pytest.skip("gcc's ASan runtime is not installed")
    out = child({"LD_PRELOAD": asan, "MIDEQ_ORACLE_LIB": ORACLE_SAN},
                ["test_oracle_solver.py", "test_oracle_dense.py", "test_oracle_tableaus.py"])
    assert " passed" in out


@pytest.mark.slow
def test_host_library_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", CSRC, "sanitize"], check=True)
    clang_rt = runtime(["/opt/rocm/llvm/bin/clang", "-print-file-name=libclang_rt.asan-x86_64.so"])
    if not clang_rt or not os.path.isabs(clang_rt) or not os.path.exists(clang_rt):
        pytest.skip("clang's ASan runtime is not installed")
    out = child({"LD_PRELOAD": clang_rt, "MIDEQ_LIB": HOST_SAN},
                ["test_abi.py", "test_user_model.py"])
    assert " passed" in out
