"""Build libflip3d_amd.so (all HIP sources, gfx950) in-tree with hipcc."""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libflip3d_amd.so")
SHIM_SRC = os.path.join(HERE, "shim", "flip3d_shim.cpp")
SHIM_OUT = os.path.join(HERE, "libflip3d_shim.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FLIP_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # bit-exact rounding: no FMA contraction; IEEE division / sqrt are the default
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    r = subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, OUT)
    build_shim()
    build_shim_check()
    if verbose:
        print("built", OUT, SHIM_OUT)
    return OUT


def build_shim():
    """The reference-signature C++ shim: plain g++ over the C ABI, linked
    against libflip3d_amd.so next to it."""
    tmp = SHIM_OUT + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           SHIM_SRC, "-L", HERE, "-lflip3d_amd", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("shim build failed:\n" + r.stderr)
    os.replace(tmp, SHIM_OUT)
    return SHIM_OUT


def build_shim_check():
    """tests/shim/shim_check: the GPU test program for the shim (built here,
    run by the GPU test suite)."""
    src = os.path.join(ROOT, "tests", "shim", "shim_check.cpp")
    out = os.path.join(ROOT, "tests", "shim", "shim_check")
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), src, "-L", HERE,
           "-lflip3d_shim", "-lflip3d_amd", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../../flip3d_amd", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("shim check build failed:\n" + r.stderr)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
