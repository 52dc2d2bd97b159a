"""Build qlens_amd/libqlx.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

    python -m qlens_amd.build

Flags: -ffp-contract=off keeps the FP64 path's operation order (no implicit fused multiply-add
outside the MFMA instructions and explicit fma calls); packed-FP32 instructions are disabled.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libqlx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QLX_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=%s" % ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip"))) + sorted(glob.glob(os.path.join(CSRC, "*.cpp")))


def build(verbose=True):
    srcs = sources()
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout[-20000:] + r.stderr[-20000:])
        raise RuntimeError("hipcc failed (%d)" % r.returncode)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
