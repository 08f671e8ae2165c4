"""Build recipe of the HIP library libfastiso.so (gfx950), in-tree.

    python -m fastisostasy_amd.build        # or fastisostasy_amd.build.build()
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfastiso.so")
SOURCES = ["rhs.hip", "rhs_fast.hip", "sealevel.hip", "init.hip", "conv.hip", "kelvin.cpp", "rheogrid.cpp", "api.cpp",
           "output.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    # packed FP32 instructions are disabled on purpose: some of their forms
    # have returned wrong results on this hardware without raising an error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose=False, jobs=None):
    """Compile every source (in parallel, `jobs` at a time) and link libfastiso.so."""
    from concurrent.futures import ThreadPoolExecutor

    builddir = os.path.join(HERE, "_build")
    os.makedirs(builddir, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(builddir, src + ".o")
        cmds.append([HIPCC] + FLAGS + ["-x", "hip", "-c", path, "-o", obj])
        objs.append(obj)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    jobs = jobs or min(len(cmds), int(os.environ.get("MAX_JOBS", "0") or 0) or os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for f in [ex.submit(run, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, check=True)
    os.replace(tmp, LIB)
    return LIB


FORTRAN_DIR = os.path.join(HERE, "fortran")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")


def build_fortran(verbose=False, single=False):
    """Compile the iso_c_binding module (fortran/fastisostasy.f90) and the
    Fortran caller drive_disc against libfastiso.so, with wp = dp, or with
    wp = sp (the reference's default precision) into drive_disc_sp when
    `single`.  Returns the driver path or None when amdflang is not installed."""
    if not os.path.exists(FLANG):
        return None
    mod = os.path.join(HERE, "_build", "fortran_sp" if single else "fortran")
    os.makedirs(mod, exist_ok=True)
    obj = os.path.join(mod, "fastisostasy.o")
    exe = os.path.join(FORTRAN_DIR, "drive_disc_sp" if single else "drive_disc")
    defs = ["-cpp"] + (["-DFI_WP_SP"] if single else [])
    cmds = [[FLANG, "-O2"] + defs + ["-c", os.path.join(FORTRAN_DIR, "fastisostasy.f90"), "-J", mod, "-o", obj],
            [FLANG, "-O2", os.path.join(FORTRAN_DIR, "drive_disc.f90"), "-J", mod, obj, "-L" + HERE,
             "-lfastiso", "-Wl,-rpath,$ORIGIN/..", "-o", exe]]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv, single=True))
