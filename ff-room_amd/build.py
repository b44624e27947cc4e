"""Build recipe of the MI355X library and its Python module.

    python ff-room_amd/build.py            # everything
    python ff-room_amd/build.py --lib      # HIP library only

Outputs (in-tree, next to the Python package so they travel with it):
  ff_room/libffroom.so                      HIP kernels + C ABI, gfx950
  ff_room/_ffroom_core.<ext-suffix>         host C++ classes + pybind11 module
Object files go to ff-room_amd/_build/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
PKG = os.path.join(HERE, "ff_room")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_SOURCES = [
    "ffr_capi.hip", "ffr_scene.hip", "ffr_task.hip", "ffr_mg.hip", "ffr_serve.hip", "ffr_gang.hip",
    "ffr_snapshot.hip", "ffr_lbm.hip", "ffr_noalloc.hip", "ffr_pool.hip",
    "kernels/step_kernels.hip", "kernels/cg_kernels.hip", "kernels/mg_kernels.hip",
    "kernels/lbm_kernels.hip",
]
HOST_SOURCES = [
    "host/Grid.cpp", "host/BoundaryManager.cpp", "host/Device.cpp",
    "host/PoissonSolver.cpp", "host/FluidSolver.cpp", "host/LBMSolver.cpp", "bindings/module.cpp",
]

HIPCC_FLAGS = [
    "-std=c++17", "-O3", "--offload-arch=gfx950", "-fPIC",
    # no packed-FP32 instructions (v_pk_*_f32) in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
    f"-I{INCLUDE}",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def _obj(src: str) -> str:
    return os.path.join(BUILD, src.replace("/", "_") + ".o")


def build_lib(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    cmds = [[hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, s), "-o", _obj(s)] for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(PKG, "libffroom.so")
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *[_obj(s) for s in srcs],
          "-o", tmp, f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, out)
    return out


def build_module() -> str:
    os.makedirs(BUILD, exist_ok=True)
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    py_inc = sysconfig.get_paths()["include"]
    flags = ["-std=c++17", "-O2", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{pybind11.get_include()}",
             f"-I{py_inc}", "-fvisibility=hidden"]
    cmds = [["g++", *flags, "-c", os.path.join(CSRC, s), "-o", _obj(s)] for s in HOST_SOURCES]
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(PKG, "_ffroom_core" + suffix)
    tmp = out + ".tmp"
    _run(["g++", "-shared", "-fPIC", *[_obj(s) for s in HOST_SOURCES], "-o", tmp,
          f"-L{PKG}", "-lffroom", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_lib()
    build_module()


if __name__ == "__main__":
    if "--lib" in sys.argv:
        print(build_lib())
    else:
        build_all()
        print("built", PKG)
