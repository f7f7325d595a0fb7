"""Build libelsim_hip.so (HIP kernels for gfx950 + the C-ABI) in-tree.

Usage: python -m elsim_amd.build  (or elsim_amd.build.build()).
Objects go to elsim_amd/_build/, the library to elsim_amd/lib/.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libelsim_hip.so")
FORTRAN = os.path.join(HERE, "fortran")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
ARCH = os.environ.get("ELSIM_ARCH", "gfx950")

SOURCES = [
    "elsim_runtime.hip",
    "elsim_stencil.hip",
    "elsim_sbm.hip",
    "elsim_levelset.hip",
    "elsim_connect.hip",
    "elsim_nn.hip",
    "elsim_noalloc.hip",
    "elsim_gen.cpp",
]

# No FMA contraction in the solver kernels: the FP64 arithmetic is meant to be
# rounded exactly as the reference writes it.  Packed FP32 instructions are
# disabled (they are emitted by -O3 from ordinary float code).
COMMON = ["-O3", "-fPIC", "-std=c++17", "-I" + os.path.join(ROOT, "include")]
HIP_FLAGS = [
    "--offload-arch=" + ARCH,
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# the learned warm start is not bound by bit-parity: contraction allowed
PER_FILE = {"elsim_nn.hip": ["-ffp-contract=fast"]}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    flags = list(COMMON)
    if src.endswith(".hip"):
        flags += ["-x", "hip"] + HIP_FLAGS + PER_FILE.get(src, [])
    _run([HIPCC] + flags + ["-c", path, "-o", obj])
    return obj


def build_fortran(lib=LIB):
    """Fortran host layer: the bind(C) module, the reference-named wrappers and
    the drop-in driver program (needs flang; skipped if absent)."""
    if not os.path.exists(FLANG):
        return None
    out = os.path.join(BUILD, "fortran")
    os.makedirs(out, exist_ok=True)
    srcs = ["elsim_gpu.f90", "elsim_dropin.f90", "sbm_run.f90"]
    objs = []
    for s in srcs:
        o = os.path.join(out, s + ".o")
        _run([FLANG, "-O2", "-fPIC", "-module-dir", out, "-c", os.path.join(FORTRAN, s), "-o", o])
        objs.append(o)
    flib = os.path.join(LIBDIR, "libelsim_fortran.so")
    _run([FLANG, "-shared", "-fPIC", "-o", flib] + objs[:2] + ["-L" + LIBDIR, "-lelsim_hip",
                                                               "-Wl,-rpath,$ORIGIN"])
    # RM_UCNCT_3D's copies of the level-set subroutines (REAL(4) time step)
    orm = os.path.join(out, "elsim_dropin_rm.f90.o")
    _run([FLANG, "-O2", "-fPIC", "-module-dir", out, "-c", os.path.join(FORTRAN, "elsim_dropin_rm.f90"), "-o", orm])
    _run([FLANG, "-shared", "-fPIC", "-o", os.path.join(LIBDIR, "libelsim_rm.so"), objs[0], orm, "-L" + LIBDIR,
          "-lelsim_hip", "-Wl,-rpath,$ORIGIN"])
    exe = os.path.join(LIBDIR, "sbm_run")
    _run([FLANG, "-o", exe, objs[2], "-L" + LIBDIR, "-lelsim_fortran", "-lelsim_hip",
          "-Wl,-rpath,$ORIGIN"])
    _run([FLANG, "-O2", "-module-dir", out, "-o", os.path.join(LIBDIR, "geom_run"),
          os.path.join(FORTRAN, "geom_run.f90"), "-L" + LIBDIR, "-lelsim_fortran", "-lelsim_hip",
          "-Wl,-rpath,$ORIGIN"])
    # the test programs of the drop-ins that no driver calls
    for name in ("dropin_check", "solve_check"):
        chk = os.path.join(ROOT, "tests", "fortran", name + ".f90")
        if os.path.exists(chk):
            _run([FLANG, "-O2", "-module-dir", out, "-o", os.path.join(LIBDIR, name), chk,
                  "-L" + LIBDIR, "-lelsim_fortran", "-lelsim_hip", "-Wl,-rpath,$ORIGIN"])
    chk = os.path.join(ROOT, "tests", "fortran", "rm_check.f90")
    if os.path.exists(chk):
        _run([FLANG, "-O2", "-o", os.path.join(LIBDIR, "rm_check"), chk, "-L" + LIBDIR, "-lelsim_rm", "-lelsim_hip",
              "-Wl,-rpath,$ORIGIN"])
    # the rank stitcher is file I/O only: no GPU library
    _run([FLANG, "-O2", "-module-dir", out, "-o", os.path.join(LIBDIR, "sbm_asmb"),
          os.path.join(FORTRAN, "sbm_asmb.f90")])
    return flib


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    if all(os.path.exists(os.path.join(FORTRAN, s)) for s in ("elsim_gpu.f90", "elsim_dropin.f90", "sbm_run.f90")):
        build_fortran()
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
