"""Build the gfx950 library (librtm3d_hip.so) and the rtm3d host driver.

Every HIP translation unit is compiled with:
  * --offload-arch=gfx950 only (no other targets, no dual code paths);
  * -ffp-contract=off: no fused multiply-add, so the FP32 arithmetic is the
    reference's operation for operation;
  * -fno-gpu-flush-denormals-to-zero: FP32 denormals kept, like the
    reference's x86 build (no FTZ/DAZ);
  * -Xclang -target-feature -Xclang -packed-fp32-ops: packed FP32
    instructions disabled.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "bin")
OBJDIR = os.path.join(PKG, "_build")
LIBNAME = os.path.join(LIBDIR, "librtm3d_hip.so")
DRIVER = os.path.join(BINDIR, "rtm3d")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_SOURCES = ["rtm_stencil.hip", "rtm_ops.hip", "rtm_ae.hip", "rtm_dct.hip", "rtm_capi.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang",
    "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-I" + os.path.join(REPO, "include"),
]
HOSTLIB = os.path.join(LIBDIR, "librtm3d_host.so")
DRIVER_SOURCES = ["driver/json.cpp", "driver/param.cpp", "driver/grid.cpp",
                  "driver/platform.cpp", "driver/kernel.cpp", "driver/controller.cpp",
                  "driver/host_api.cpp"]
# The host set-up arithmetic must match the reference's x86 build: no FMA
# contraction, no fast-math, SSE2 (no -march), OpenMP for host loops.
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall",
             "-I" + os.path.join(REPO, "include"), "-I" + os.path.join(CSRC, "driver")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r


def _stale(obj, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build_library(jobs=4, incremental=False):
    """Compiles every HIP source (incremental=True: only those whose object
    is older than the source or a shared header) and links the library."""
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    headers = [os.path.join(CSRC, "rtm_internal.hpp"), os.path.join(REPO, "include", "rtm_hip.h")]
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
        objs.append(obj)
        if incremental and not _stale(obj, [os.path.join(CSRC, src)] + headers):
            continue
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIBNAME] + objs)
    return LIBNAME


def build_driver(jobs=4):
    os.makedirs(BINDIR, exist_ok=True)
    os.makedirs(os.path.join(OBJDIR, "driver"), exist_ok=True)
    objs, cmds = [], []
    for src in DRIVER_SOURCES:
        obj = os.path.join(OBJDIR, src.replace(".cpp", ".o"))
        objs.append(obj)
        cmds.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    _run(["g++", "-shared", "-fopenmp", "-o", HOSTLIB] + objs +
         ["-L" + LIBDIR, "-lrtm3d_hip", "-Wl,-rpath,$ORIGIN"])
    main_obj = os.path.join(OBJDIR, "driver", "main.o")
    _run(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, "driver", "main.cpp"), "-o", main_obj])
    _run(["g++", "-fopenmp", "-o", DRIVER, main_obj, "-L" + LIBDIR, "-lrtm3d_host",
          "-lrtm3d_hip", "-Wl,-rpath,$ORIGIN/../lib"])
    return DRIVER


def build_all(jobs=4, incremental=False):
    build_library(jobs, incremental)
    if all(os.path.exists(os.path.join(CSRC, s)) for s in DRIVER_SOURCES):
        build_driver(jobs)


if __name__ == "__main__":
    build_all(incremental="--incremental" in sys.argv)
    print("built", LIBNAME)
