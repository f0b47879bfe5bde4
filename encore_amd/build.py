"""Builds libencore_rq.so in-tree with hipcc for gfx950.

`python -m encore_amd.build` or `__graft_entry__.build()`.  Every source is
compiled to an object in parallel, then linked into encore_amd/lib/.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
LIBNAME = "libencore_rq.so"

ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    # Bit-exact FP32: no contraction into FMA; IEEE division stays the default.
    "-ffp-contract=off",
    "-I" + os.path.join(ROOT, "include"),
]
DEVICE = [
    "--offload-arch=" + ARCH,
    # Packed FP32 instructions are kept out of all device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(h):
        raise RuntimeError("hipcc not found")
    return h


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".cpp", ".hip")))


def _compile(src):
    obj = os.path.join(OBJDIR, src + ".o")
    cmd = [hipcc()] + COMMON
    if src.endswith(".hip"):
        cmd += DEVICE + ["-x", "hip"]
    else:
        cmd += ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-x", "c++"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    workers = max(1, min(len(srcs), os.cpu_count() or 4, 8))
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, srcs))
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_oracle(verbose=False):
    """Builds the parity oracle (oracle/liboracle.so) and, when the reference
    sources are present, its reference-built part (oracle/_ref)."""
    odir = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-s", "-C", odir, "-B", "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    if os.path.isdir("/root/reference/ENCORE/src"):
        r = subprocess.run(["make", "-s", "-C", odir, "-B", "ref"], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("reference predicate build failed:\n" + r.stdout + r.stderr)
    if verbose:
        print("built oracle")


if __name__ == "__main__":
    build(verbose=True)
    if "--oracle" in sys.argv:
        build_oracle(verbose=True)
