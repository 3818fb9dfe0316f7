"""Builds the HIP shared library libphinum_gpu.so in-tree for gfx950.

Every .hip file under phinum_amd/csrc is compiled with hipcc and linked
against rocFFT.  Floating-point contraction is off so that the FP64 stencil
arithmetic keeps the reference's rounding (dynamics.cpp:80,84); packed FP32
instructions are disabled.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libphinum_gpu.so")
DROPIN_LIB = os.path.join(HERE, "libphinum_dropin.so")
OBJ = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIPFLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _hipcc():
    h = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    if not os.path.exists(h):
        raise RuntimeError("hipcc not found")
    return h


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    hipcc = _hipcc()

    def comp(f):
        obj = os.path.join(OBJ, f[:-4] + ".o")
        cmd = [hipcc] + HIPFLAGS + ["-I", os.path.join(HERE, "..", "include"), "-c",
                                    os.path.join(CSRC, f), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(comp, srcs))
    tmp = OUT + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L", os.path.join(ROCM, "lib"), "-lrocfft", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, OUT)
    build_dropin(verbose)
    return OUT


def build_dropin(verbose=False):
    """Host C++ drop-in layer (reference names) over the C ABI."""
    d = os.path.join(CSRC, "dropin")
    if not os.path.isdir(d):
        return None
    srcs = [os.path.join(d, f) for f in sorted(os.listdir(d)) if f.endswith(".cpp") and f != "phinum_main.cpp"]
    if not srcs:
        return None
    tmp = DROPIN_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-I",
           os.path.join(HERE, "..", "include"), "-o", tmp] + srcs + [
        "-L", HERE, "-lphinum_gpu", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(tmp, DROPIN_LIB)
    main = os.path.join(d, "phinum_main.cpp")
    if os.path.exists(main):
        exe = os.path.join(HERE, "phinum_main")
        _run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(HERE, "..", "include"),
              "-o", exe + ".tmp", main, "-L", HERE, "-lphinum_dropin", "-lphinum_gpu",
              "-Wl,-rpath,$ORIGIN"])
        os.replace(exe + ".tmp", exe)
    return DROPIN_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
