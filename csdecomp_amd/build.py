"""Build recipe for the native parts of csdecomp_amd.

* ``lib/libcsdecomp_hip.so``: the gfx950 kernels and the C-ABI of
  ``include/csdecomp_amd.h`` (hipcc, one object per translation unit).
* ``_pycsdecomp*.so``: the C++ host API and its pybind11 module, linked
  against the HIP library (g++; the host code only sees the C-ABI header).

Run ``python csdecomp_amd/build.py`` (or ``__graft_entry__.build()``).
Every call rebuilds from source.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HIP_DIR = os.path.join(PKG, "csrc", "hip")
HOST_DIR = os.path.join(PKG, "csrc", "host")
LIB_DIR = os.path.join(PKG, "lib")
OBJ_DIR = os.path.join(ROOT, "_build", "obj")

HIP_SOURCES = ["csd_plant.hip", "csd_check.hip", "csd_eizo.hip", "csd_stages.hip", "csd_edges.hip",
               "csd_mlp.hip", "csd_field.hip", "csd_screen.hip", "csd_bisect8.hip",
               "csd_bisect16.hip"]
HOST_SOURCES = ["csdecomp.cpp", "cpu_checker.cpp", "bindings.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops: no v_pk_*_f32 in generated code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def build_hip(jobs=None):
    os.makedirs(OBJ_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(HIP_DIR, s))]

    def one(src):
        obj = os.path.join(OBJ_DIR, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(HIP_DIR, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=jobs or min(8, len(srcs))) as ex:
        objs = list(ex.map(one, srcs))
    lib = os.path.join(LIB_DIR, "libcsdecomp_hip.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, lib)
    return lib


def build_host():
    import pybind11
    inc = [pybind11.get_include(), sysconfig.get_paths()["include"]]
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(PKG, "_pycsdecomp" + suffix)
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
           *[f"-I{i}" for i in inc],
           *[os.path.join(HOST_DIR, s) for s in HOST_SOURCES],
           "-L" + LIB_DIR, "-lcsdecomp_hip", "-Wl,-rpath,$ORIGIN/lib", "-o", tmp]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_all():
    lib = build_hip()
    mod = build_host()
    return lib, mod


if __name__ == "__main__":
    for p in build_all():
        print("built", p)
    sys.exit(0)
