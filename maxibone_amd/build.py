"""Build the MI355X diffusion library and its Python module in-tree.

* ``libmxb_diffusion.so`` - the C ABI (include/maxibone_diffusion.h) and the
  gfx950 kernels, compiled by hipcc.
* ``diffusion<ext-suffix>`` - the pybind11 drop-in module, compiled by g++
  against the C ABI only and linked to the library with an $ORIGIN rpath.

Every HIP source is compiled with packed-FP32 instructions disabled and with
FP contraction off (FMAs appear only where the source asks for them).

Usage: ``python -m maxibone_amd.build``.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")

HIP_SOURCES = ["mxb_kernels_f32.hip", "mxb_kernels_f32_lds.hip", "mxb_kernels_lowp.hip", "mxb_kernels_lowp_lds.hip", "mxb_fused.hip", "mxb_batched.hip", "mxb_fields.hip", "mxb_edt.hip", "mxb_probe.hip", "mxb_capi.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I" + INCLUDE, "-I" + CSRC,
]

LIB = os.path.join(PKG, "libmxb_diffusion.so")
EXT = os.path.join(PKG, "diffusion" + sysconfig.get_config_var("EXT_SUFFIX"))


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in line)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def _compile_hip(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_library(jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    sources = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_hip, sources))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", LIB + ".tmp"])
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_extension():
    import pybind11

    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall",
           "-isystem", pybind11.get_include(), "-isystem", sysconfig.get_paths()["include"],
           "-I" + INCLUDE, os.path.join(CSRC, "diffusion_pybind.cpp"),
           "-L" + PKG, "-lmxb_diffusion", "-Wl,-rpath,$ORIGIN", "-o", EXT + ".tmp"]
    _run(cmd)
    os.replace(EXT + ".tmp", EXT)
    return EXT


def build(jobs=4):
    lib = build_library(jobs)
    ext = build_extension()
    return lib, ext


if __name__ == "__main__":
    for p in build():
        print(p)
    sys.exit(0)
