"""Build recipe for the oracle (test infrastructure only).

* ``liboracle.so``: the plain-C restatements in ``diffusion_oracle.c`` (the
  diffusion) and ``edt_oracle.c`` (the EDT field's distance transform).
  Flags follow the reference CPU build (src/Makefile:9-18: -O3, OpenMP, no
  -march) plus ``-ffp-contract=off`` so that no FMA is ever formed.
* ``_ref/<module><ext-suffix>``: the reference's own CPU modules, compiled
  directly from its sources where they lie
  (src/pybind/diffusion-pybind.cc, which #includes
  src/lib/cpp/cpu/diffusion.cc and src/lib/cpp/include/*.hh) with the
  reference's compile flags (src/Makefile:9,18) and the image's pybind11
  headers.  Output goes only to ``oracle/_ref/`` (git-ignored).  Skipped when
  the reference tree is absent (the prebuilt file then travels as is).

Usage: ``python oracle/build.py``.
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SRC = "/root/reference/src"


def _run(cmd):
    subprocess.run(cmd, check=True)


def build_oracle() -> str:
    out = os.path.join(HERE, "liboracle.so")
    srcs = [os.path.join(HERE, f) for f in ("diffusion_oracle.c", "edt_oracle.c")]
    _run(["gcc", "-O3", "-fopenmp", "-ffp-contract=off", "-std=gnu11", "-fPIC", "-shared",
          "-Wall", "-Wextra", *srcs, "-o", out + ".tmp", "-lm"])
    os.replace(out + ".tmp", out)
    return out


def ref_module_path(name: str = "diffusion") -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    return os.path.join(HERE, "_ref", name + suffix)


# Reference CPU modules compiled from their own sources: the diffusion path and
# its downstream consumers (histograms: src/pybind/histograms-pybind.cc with
# src/lib/cpp/cpu/histograms.cc; general: src/pybind/general-pybind.cc with
# src/lib/cpp/cpu/general.cc, for the C++ to_int / normalized_convert).
REF_MODULES = ("diffusion", "histograms", "general")


def build_ref(name: str = "diffusion") -> str | None:
    """Compile the reference CPU module `name` from its own sources."""
    pyb = os.path.join(REF_SRC, "pybind", name + "-pybind.cc")
    if not os.path.exists(pyb):
        return None
    import pybind11  # header-only, part of the image

    out = ref_module_path(name)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    inc = [
        "-isystem", pybind11.get_include(),
        "-isystem", sysconfig.get_paths()["include"],
        "-I" + os.path.join(REF_SRC, "lib", "cpp", "include"),
        "-I" + os.path.join(REF_SRC, "lib", "cpp", "cpu"),
    ]
    # src/Makefile:9 (CXXFLAGS), :16 and :18 (cpu platform flags), :31 (-mno-tbm)
    flags = ["-Wno-unknown-pragmas", "-Wno-comment", "-fopenmp", "-shared", "-fPIC",
             "-std=c++20", "-O3", "-mno-tbm", "-w"]
    _run(["g++", *flags, *inc, pyb, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build_oracle())
    for m in REF_MODULES:
        print(build_ref(m))
    sys.exit(0)
