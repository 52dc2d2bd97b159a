"""Build qlens_amd/libqlx.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

    python -m qlens_amd.build            # recompile sources whose inputs changed, then link
    python -m qlens_amd.build --force    # recompile everything

Each source compiles to its own object under _build/ (non-RDC: every kernel is launched from its
own translation unit), then hipcc links the shared library.  Flags: -ffp-contract=off keeps the
FP64 path's operation order (no implicit fused multiply-add outside MFMA and explicit fma calls);
packed-FP32 instructions are disabled.
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(REPO, "_build", "qlx")
OUT = os.path.join(HERE, "libqlx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QLX_ARCH", "gfx950")

CFLAGS = [
    "--offload-arch=%s" % ARCH, "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip"))) + sorted(glob.glob(os.path.join(CSRC, "*.cpp")))


def _headers():
    return glob.glob(os.path.join(CSRC, "*.h")) + [os.path.join(REPO, "include", "qlx.h")]


def _stale(src, obj, hdr_mtime):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return os.path.getmtime(src) > t or hdr_mtime > t


def _compile(src, obj, verbose):
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout[-20000:] + r.stderr[-20000:])
        raise RuntimeError("hipcc failed on %s (%d)" % (os.path.basename(src), r.returncode))
    return obj


def build(verbose=True, force=False, jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    hdr = max(os.path.getmtime(h) for h in _headers())
    work, objs = [], []
    for src in sources():
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        objs.append(obj)
        if force or _stale(src, obj, hdr):
            work.append((src, obj))
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(work) or 1))) as ex:
        list(ex.map(lambda a: _compile(a[0], a[1], verbose), work))
    if work or force or not os.path.exists(OUT):
        tmp = OUT + ".tmp"
        cmd = [HIPCC, "--offload-arch=%s" % ARCH, "-shared", "-fPIC", "-o", tmp] + objs
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout[-20000:] + r.stderr[-20000:])
            raise RuntimeError("link failed (%d)" % r.returncode)
        os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(force="--force" in sys.argv)
