"""Builds the native libraries of the package in-tree.

    python -m sar_cambrianaccelerationcard_amd.build        # everything
    python -m sar_cambrianaccelerationcard_amd.build sarbp  # one target

Targets (all written to sar_cambrianaccelerationcard_amd/_lib/):
  libsarbp.so   gfx950 HIP back-projection, range-compression and omega-K
                kernels + C ABI (include/sarbp.h); links hipFFT
  libsarsim.so  FP64 CPU scene simulator (csrc/sarsim.h)
  sarbp_run     C++ command-line driver (mirror of the reference's main.cpp)
  sarwk_run     C++ omega-K command-line imager (mirror of wk.cc's process_wk)

Every HIP file is compiled with packed FP32 instructions disabled
(-target-feature -packed-fp32-ops) for the device.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "_lib")
INCLUDE = os.path.join(ROOT, "include")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.environ.get("HIPCC", os.path.join(ROCM, "bin", "hipcc"))
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


SARBP_SOURCES = ("sarbp_kernels_nearest.hip", "sarbp_kernels_linear.hip", "sarbp_kernels_sinc8.hip",
                 "sarbp_kernels_sinc8w.hip",
                 "sarbp_api.hip", "sarbp_batch.hip", "sarbp_server.hip",
                 "sarbp_range.hip", "sarbp_wk.hip", "sarbp_echo.hip", "sarbp_diag.hip")
# Host-only sources without HIP runtime calls (planning, the server core):
# compiled by the host compiler, so the CPU sanitizer builds of the tests
# compile the same files.
SARBP_HOST_SOURCES = ("sarbp_plan.cpp", "sarbp_server_core.cpp")
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
              f"-I{ROCM}/include", f"-I{INCLUDE}", f"-I{CSRC}"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature and says so.
    err = "\n".join(l for l in r.stderr.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{err}")


def _newer(target: str, sources: list[str]) -> bool:
    if not os.path.exists(target):
        return False
    t = os.path.getmtime(target)
    return all(os.path.getmtime(s) <= t for s in sources)


def build_sarbp(force: bool = True) -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libsarbp.so")
    srcs = [os.path.join(CSRC, f) for f in SARBP_SOURCES]
    hsrcs = [os.path.join(CSRC, f) for f in SARBP_HOST_SOURCES]
    deps = srcs + hsrcs + [os.path.join(CSRC, h) for h in (
        "sarbp_internal.h", "sarbp_host.h", "sarbp_fft.h", "sarbp_kernels.inc", "sarbp_rotation.h",
        "sarbp_server_core.h")] + [os.path.join(INCLUDE, "sarbp.h")]
    if not force and _newer(out, deps):
        return out
    from concurrent.futures import ThreadPoolExecutor
    objs = [os.path.join(LIB, os.path.basename(s) + ".o") for s in srcs + hsrcs]
    cmds = [[HIPCC, *HIP_FLAGS, "-c", s, "-o", o] for s, o in zip(srcs, objs)] + \
           [["g++", *HOST_FLAGS, "-c", s, "-o", o] for s, o in zip(hsrcs, objs[len(srcs):])]
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(_run, cmds))
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-L" + os.path.join(ROCM, "lib"), "-lhipfft", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    return out


def build_sarsim(force: bool = True) -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libsarsim.so")
    src = os.path.join(CSRC, "sarsim.cpp")
    if not force and _newer(out, [src, os.path.join(CSRC, "sarsim.h")]):
        return out
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall",
          "-Wextra", src, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_run(force: bool = True) -> str:
    """The sarbp_run command-line imager (host code only, links both libraries)."""
    out = os.path.join(LIB, "sarbp_run")
    src = os.path.join(CSRC, "sarbp_run.cpp")
    deps = [src, os.path.join(INCLUDE, "sarbp.h"), os.path.join(CSRC, "sarsim.h"),
            os.path.join(LIB, "libsarbp.so"), os.path.join(LIB, "libsarsim.so")]
    if not force and _newer(out, deps):
        return out
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
          "-I" + os.path.join(ROCM, "include"), "-I" + INCLUDE, "-I" + CSRC, src, "-o", tmp,
          "-L" + LIB, "-lsarbp", "-lsarsim", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, out)
    return out


def build_wk_run(force: bool = True) -> str:
    """The sarwk_run omega-K command-line imager (host code only)."""
    out = os.path.join(LIB, "sarwk_run")
    src = os.path.join(CSRC, "sarwk_run.cpp")
    deps = [src, os.path.join(INCLUDE, "sarbp.h"), os.path.join(LIB, "libsarbp.so")]
    if not force and _newer(out, deps):
        return out
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
          "-I" + os.path.join(ROCM, "include"), "-I" + INCLUDE, src, "-o", tmp,
          "-L" + LIB, "-lsarbp", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, out)
    return out


def build_all(force: bool = True) -> list[str]:
    return [build_sarbp(force), build_sarsim(force), build_run(force), build_wk_run(force)]


if __name__ == "__main__":
    which = sys.argv[1:] or ["all"]
    for w in which:
        if w == "all":
            print("\n".join(build_all()))
        elif w == "sarbp":
            print(build_sarbp())
        elif w == "sarsim":
            print(build_sarsim())
        elif w == "run":
            print(build_run())
            print(build_wk_run())
        else:
            raise SystemExit(f"unknown target {w}")
