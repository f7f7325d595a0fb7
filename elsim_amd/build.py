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
    "elsim_mg.hip",
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
# named ranges of a step's phases (rocprofv3 --marker-trace)
ROCTX = ["-L" + os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(HIPCC))), "lib"),
         "-lrocprofiler-sdk-roctx"]
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
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ROCTX + ["-lpthread"])
    os.replace(tmp, LIB)
    if all(os.path.exists(os.path.join(FORTRAN, s)) for s in ("elsim_gpu.f90", "elsim_dropin.f90", "sbm_run.f90")):
        build_fortran()
    if verbose:
        print("built", LIB)
    return LIB


# Host code of the library under AddressSanitizer and UndefinedBehaviorSanitizer:
# each sanitizer flag goes to the host compilation only (the device code is
# built as usual, at -O1).
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
SAN_DIR = os.path.join(BUILD, "sanitize")


def build_sanitized():
    """libelsim_hip_san.so (host code only) and the host driver of
    tests/sanitize/host_driver.cpp linked against it; returns the driver."""
    os.makedirs(SAN_DIR, exist_ok=True)

    def comp(src):
        obj = os.path.join(SAN_DIR, src + ".o")
        flags = ["-O1", "-g", "-fPIC", "-std=c++17", "-I" + os.path.join(ROOT, "include")] + SAN_FLAGS
        if src.endswith(".hip"):
            flags += ["-x", "hip", "--offload-arch=" + ARCH, "-ffp-contract=off"]
        _run([HIPCC] + flags + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(comp, SOURCES))
    lib = os.path.join(SAN_DIR, "libelsim_hip_san.so")
    _run([HIPCC, "-shared", "-fPIC", "-shared-libsan"] + SAN_FLAGS + ["-o", lib] + objs + ROCTX + ["-lpthread"])
    drv = os.path.join(SAN_DIR, "host_driver")
    # the sanitizer runtimes of the compiler (a shared runtime, so that the
    # library and the driver share one)
    rt = subprocess.run([os.path.join(os.path.dirname(os.path.realpath(HIPCC)), "..", "lib", "llvm", "bin", "clang"),
                         "--print-file-name=libclang_rt.asan-x86_64.so"], stdout=subprocess.PIPE, text=True)
    rtdir = os.path.dirname(rt.stdout.strip()) if rt.returncode == 0 else ""
    rpath = [SAN_DIR] + ([rtdir] if rtdir and os.path.isdir(rtdir) else [])
    _run([HIPCC, "-O1", "-g", "-std=c++17", "-shared-libsan"] + SAN_FLAGS +
         ["-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "sanitize", "host_driver.cpp"),
          "-o", drv, "-L" + SAN_DIR, "-lelsim_hip_san"] + ["-Wl,-rpath," + r for r in rpath])
    return drv


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        print("built", build_sanitized())
    else:
        build(verbose=True)
    sys.exit(0)
