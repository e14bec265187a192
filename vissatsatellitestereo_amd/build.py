"""Build libvissat_mvs.so (HIP, gfx950) in-tree.

    python -m vissatsatellitestereo_amd.build [--arch gfx950] [--jobs N]

Every translation unit is compiled by hipcc for gfx950 only, with
``-ffp-contract=off`` (explicit fmaf in kernels, bit-reproducible FP64 grid
indexing) and packed-FP32 instructions disabled.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB_NAME = "libvissat_mvs.so"
SOURCES = ["vsm_pm.hip", "vsm_fuse.hip", "vsm_api.cpp"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags(arch):
    return ["-x", "hip", f"--offload-arch={arch}", "-O3", "-fPIC", "-std=c++17",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-I", CSRC, "-I", INCLUDE]


def build(arch="gfx950", jobs=3, verbose=False):
    out_dir = os.path.join(HERE, "_build")
    os.makedirs(out_dir, exist_ok=True)
    cc = hipcc()

    def compile_one(src):
        obj = os.path.join(out_dir, os.path.splitext(src)[0] + ".o")
        cmd = [cc] + flags(arch) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    lib = os.path.join(HERE, LIB_NAME)
    tmp = lib + ".tmp"
    cmd = [cc, f"--offload-arch={arch}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    return lib


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default="gfx950")
    ap.add_argument("--jobs", type=int, default=3)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.arch, a.jobs, a.verbose))


if __name__ == "__main__":
    sys.exit(main())
