"""Compile the HIP library for gfx950 in place (altar_amd/libaltar_amd.so).

Every source under altar_amd/csrc is built with hipcc into one shared
object.  Packed FP32 instructions are disabled (-packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libaltar_amd.so")
SOURCES = ["sources.hip", "gemm.hip", "sampler.hip", "mlp.hip", "cov.hip", "trajectory.hip", "capi.hip"]

# Per-source options.  trajectory.hip: the fused kernel keeps network weights
# in registers across its sub-step loop, and machine-level loop-invariant code
# motion would otherwise pin the FP64 constants of log and sincos in vector
# registers for the whole loop, leaving too few for the weights.
EXTRA = {"trajectory.hip": ["-mllvm", "-disable-machine-licm"]}
ARCH = os.environ.get("ALTAR_OFFLOAD_ARCH", "gfx950")


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def flags():
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-I", os.path.join(HERE, "..", "include")]


def build(verbose=False, jobs=4):
    """Compile each source to an object (in parallel), then link."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [hipcc()] + flags() + EXTRA.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(text)
        elif verbose and text.strip():
            sys.stderr.write(text)
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB + ".tmp"
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + \
        ["-L", os.path.join(rocm, "lib"), "-Wl,-rpath," + os.path.join(rocm, "lib")]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
