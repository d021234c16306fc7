"""Build recipe of the native library (libcellgpu_amd.so) for gfx950.

Compiles every HIP kernel file and the C++ host layer with hipcc and links one
shared library next to this file. Parallel compilation, no incremental logic:
every call rebuilds (a few seconds per file).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcellgpu_amd.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CGPU_OFFLOAD_ARCH", "gfx950")
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Wno-unused-result",
]
DEVICE = [
    "--offload-arch=" + ARCH,
    # packed FP32 VALU forms are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    kern = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    host_dir = os.path.join(CSRC, "host")
    host = sorted(os.path.join(host_dir, f) for f in os.listdir(host_dir) if f.endswith(".cpp"))
    return kern, host


def _compile(args):
    src, obj, flags = args
    cmd = [HIPCC] + flags + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    kern, host = sources()
    jobs = []
    for s in kern:
        jobs.append((s, os.path.join(BUILD, os.path.basename(s) + ".o"), COMMON + DEVICE))
    for s in host:
        jobs.append((s, os.path.join(BUILD, os.path.basename(s) + ".o"), COMMON + DEVICE))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        objs = list(ex.map(_compile, jobs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, OUT)
    # drivers (reference voronoi.cpp command line) linked against the library
    ddir = os.path.join(CSRC, "drivers")
    bdir = os.path.join(HERE, "bin")
    os.makedirs(bdir, exist_ok=True)
    for f in sorted(os.listdir(ddir)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(bdir, f[:-4])
        cmd = [HIPCC] + COMMON + ["-o", exe, os.path.join(ddir, f), "-L" + HERE, "-lcellgpu_amd",
                                  "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("driver build failed: %s\n%s" % (r.stdout, r.stderr))
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
