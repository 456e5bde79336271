"""Build the native libraries in-tree.

  libqgw_host.so   host C (gcc): params reader, grid, modes, initial condition,
                   netCDF classic output and restart input
  libqgw_hip.so    HIP C++ for gfx950 (hipcc): the step kernels and the C-ABI
  qg_hip           the C driver (restated main loop of the reference)

Every call compiles (no up-to-date checks), so a stale library never survives
a source change.  Object files go to qgw_amd/_build/.
"""
import os
import shlex
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions are disabled for all device code
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-I" + INC,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
CC_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-fopenmp", "-I" + INC, "-Wall",
            "-march=x86-64-v2"]

HIP_SOURCES = ["qgw_ctx.hip", "qgw_launch_f32.hip", "qgw_launch_f64.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(shlex.quote(c) for c in cmd) + "\n" + r.stdout)
        raise RuntimeError("build step failed: %s" % cmd[-1])
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    host_so = os.path.join(HERE, "libqgw_host.so")
    _run(["gcc"] + CC_FLAGS + ["-shared", "-o", host_so,
                               os.path.join(HOST, "qgw_host.c"), os.path.join(HOST, "qgw_nc.c"),
                               os.path.join(HOST, "qgw_stoch.c"), "-lm"])

    def comp(src):
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=len(HIP_SOURCES)) as ex:
        objs = list(ex.map(comp, HIP_SOURCES))
    hip_so = os.path.join(HERE, "libqgw_hip.so")
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", hip_so] + objs +
         ["-L" + HERE, "-lqgw_host", "-lrocblas", "-Wl,-rpath,$ORIGIN"])
    drv = os.path.join(HERE, "qg_hip")
    if os.path.exists(os.path.join(HOST, "qg_main.c")):
        _run(["gcc"] + CC_FLAGS + ["-o", drv, os.path.join(HOST, "qg_main.c"),
                                   "-L" + HERE, "-lqgw_hip", "-lqgw_host", "-lm",
                                   "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link," + "/opt/rocm/lib"])
    if verbose:
        print("built", host_so, hip_so)
    return host_so, hip_so


if __name__ == "__main__":
    build(verbose=True)
