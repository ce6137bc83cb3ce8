"""The host code of the library under ThreadSanitizer and AddressSanitizer +
UndefinedBehaviorSanitizer, on the CPU (no GPU needed): the server core and
the launch planner.

tests/host/server_stress.cpp runs the continuous-batching server core
(csrc/sarbp_server_core.cpp: admission, scheduler and completion threads,
job table, arena, watchdog, cancellation, over-budget rejection, shutdown)
against a host-memory stand-in for the device; the planning it calls is
csrc/sarbp_plan.cpp.  These are the same source files the library builds.
Each variant is compiled here and must finish every scenario with no
sanitizer report.  (The reference has no server; its host code exits on the
first CUDA error, code/SAR_BP/CUDA/BPA_CUDA.h:11-17.)

ThreadSanitizer comes from ROCm's clang: GCC 11's libtsan does not intercept
pthread_cond_clockwait, which libstdc++ uses for condition-variable waits
with a timeout, and then reports mutexes as held by two threads at once."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "sar_cambrianaccelerationcard_amd", "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
CLANG = os.path.join(ROCM, "llvm", "bin", "clang++")
PROGRAMS = {
    "server_stress": ([os.path.join(CSRC, "sarbp_plan.cpp"), os.path.join(CSRC, "sarbp_server_core.cpp"),
                       os.path.join(ROOT, "tests", "host", "server_stress.cpp")], "all scenarios passed"),
    # the planner on random grids, tracks and parameters, with its launch
    # invariants (window covers every pixel's taps, LDS budget, DF quantum,
    # grouping) checked
    "plan_fuzz": ([os.path.join(CSRC, "sarbp_plan.cpp"), os.path.join(ROOT, "tests", "host", "plan_fuzz.cpp")],
                  "all plan checks passed"),
}
COMMON = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__",
          f"-I{ROCM}/include", f"-I{ROOT}/include", f"-I{CSRC}", "-pthread"]

VARIANTS = {
    "thread": ([CLANG], ["-fsanitize=thread"],
               {"TSAN_OPTIONS": "halt_on_error=1 second_deadlock_stack=1"}),
    "address_undefined": (["g++"], ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"],
                          {"ASAN_OPTIONS": "detect_leaks=1:strict_string_checks=1",
                           "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"}),
}


def _have(cc):
    return os.path.isabs(cc) and os.access(cc, os.X_OK) or shutil.which(cc) is not None


@pytest.mark.parametrize("program,variant", [("server_stress", "thread"), ("server_stress", "address_undefined"),
                                             ("plan_fuzz", "address_undefined")])
def test_host_code_under_sanitizer(tmp_path, program, variant):
    cc, flags, env_extra = VARIANTS[variant]
    sources, passed = PROGRAMS[program]
    if not _have(cc[0]):
        pytest.skip(f"{cc[0]} not available")
    exe = str(tmp_path / f"{program}_{variant}")
    b = subprocess.run(cc + COMMON + flags + sources + ["-o", exe], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, **env_extra)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    print(r.stdout)
    assert r.returncode == 0, out[-6000:]
    assert passed in r.stdout
    for marker in ("WARNING: ThreadSanitizer", "ERROR: AddressSanitizer", "ERROR: LeakSanitizer",
                   "runtime error:"):
        assert marker not in out, out[-6000:]
