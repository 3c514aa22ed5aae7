"""Build libtealeaf_mi355x.so (gfx950) in-tree with hipcc.

Usage: python -m tealeaf_opencl_amd.build   (or build() from Python)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIBNAME = "libtealeaf_mi355x.so"
ARCH = os.environ.get("TL_OFFLOAD_ARCH", "gfx950")

SOURCES = ["tl_kernels.hip", "tl_refine.hip", "tl_adjoint.hip", "tl_chunk.hip", "tl_driver.hip", "tl_deck.cpp",
           "tl_eigen.cpp"]
HEADERS = ["tl_common.h", "tl_device.h", "tl_kernels.h", "tl_chunk.h", "tl_deck.h", "tl_eigen.h"]

# packed FP32 instructions are disabled on every device compile
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# an ignored hipError_t (a nodiscard value) fails the build
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Werror=shadow", "-Werror=unused-value",
          "-I", os.path.join(HERE, "..", "include")]


def lib_path():
    return os.path.join(LIBDIR, LIBNAME)


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if proc.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def build(verbose=False, jobs=4):
    """Compile every source for gfx950 and link the shared library."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    cmds = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        path = os.path.join(CSRC, src)
        if src.endswith(".hip"):
            cmd = ["hipcc", "--offload-arch=" + ARCH, "-x", "hip"] + COMMON + DEVICE_FLAGS + ["-c", path, "-o", obj]
        else:
            cmd = ["g++"] + COMMON + ["-c", path, "-o", obj]
        cmds.append(cmd)
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    errors = []
    for cmd, p in procs:
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines()
                        if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
        if p.returncode != 0:
            errors.append(" ".join(cmd) + "\n" + out)
        elif verbose and out.strip():
            print(out)
    if errors:
        raise RuntimeError("build failed:\n" + "\n\n".join(errors))
    tmp = lib_path() + ".tmp"
    link = ["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl"]
    if verbose:
        print(" ".join(link), flush=True)
    _run(link)
    os.replace(tmp, lib_path())
    return lib_path()


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
