"""Build libmxb_diffusion with AddressSanitizer and UndefinedBehaviorSanitizer
on its host code, and the host_check driver that exercises it.

The HIP sources compile with the library's own flags plus the sanitizers
behind -Xarch_host (device code is built as usual); the library links with
ROCm's clang, and host_check links the static ASan runtime into the
executable, so nothing has to be preloaded.  Outputs go to
tests/sanitize/_build/.

    python tests/sanitize/build.py        # build (incremental by mtime)
    tests/sanitize/_build/host_check cpu  # no-device checks
    tests/sanitize/_build/host_check gpu  # full checks on a GPU
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from maxibone_amd import build as B  # noqa: E402

OUT = os.path.join(HERE, "_build")
OBJ = os.path.join(OUT, "obj")  # intermediate objects (not needed to run)
LIB = os.path.join(OUT, "libmxb_diffusion_san.so")
EXE = os.path.join(OUT, "host_check")
DRIVER = os.path.join(HERE, "host_check.cpp")
CLANG = "/opt/rocm/llvm/bin/clang++"
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer",
            "-Xarch_host", "-gline-tables-only"]
SAN_LINK = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]


def _inputs():
    srcs = [os.path.join(B.CSRC, s) for s in B.HIP_SOURCES]
    deps = [os.path.join(B.CSRC, f) for f in os.listdir(B.CSRC) if f.endswith(".h")]
    deps += [os.path.join(B.INCLUDE, f) for f in os.listdir(B.INCLUDE) if f.endswith(".h")]
    return srcs, deps


def stale():
    """True when the sanitized build is missing or older than a source."""
    srcs, deps = _inputs()
    if not (os.path.exists(LIB) and os.path.exists(EXE)):
        return True
    built = min(os.path.getmtime(LIB), os.path.getmtime(EXE))
    return any(os.path.getmtime(p) > built for p in srcs + deps + [DRIVER, __file__])


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(os.path.basename(src))[0] + ".o")
    B._run([B.HIPCC, *B.HIP_FLAGS, *SAN_HOST, "-c", src, "-o", obj])
    return obj


def build(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    srcs, _ = _inputs()
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    B._run([CLANG, "-shared", "-fPIC", *SAN_LINK, *objs, "-L/opt/rocm/lib", "-lamdhip64",
            "-Wl,-rpath,/opt/rocm/lib", "-o", LIB + ".tmp"])
    os.replace(LIB + ".tmp", LIB)
    B._run([CLANG, "-std=c++17", "-O1", *SAN_LINK, "-fno-omit-frame-pointer", "-gline-tables-only",
            "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + B.INCLUDE, DRIVER,
            "-L" + OUT, "-lmxb_diffusion_san", "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64",
            "-Wl,-rpath,/opt/rocm/lib", "-o", EXE + ".tmp"])
    os.replace(EXE + ".tmp", EXE)
    return EXE


def run(mode, timeout=600):
    """Run host_check in `mode` with the sanitizers set to stop at the first
    report; returns (exit status, output)."""
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "halt_on_error=1:detect_leaks=0:verify_asan_link_order=0"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    r = subprocess.run([EXE, mode], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       env=env, timeout=timeout)
    return r.returncode, r.stdout


if __name__ == "__main__":
    print(build())
