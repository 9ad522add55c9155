"""Build libgqf.so (HIP kernels + C ABI) for gfx950, in-tree.

Usage: python -m cqf_amd.build   (or call build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgqf.so")
SOURCES = ["gqf_api.hip", "gqf_sort.hip", "gqf_rebuild.hip", "gqf_query.hip", "gqf_image.hip", "gqf_point.hip", "gqf_group.hip", "gqf_enum.hip",
           "gqf_agg.hip", "gqf_server.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GQF_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # no packed-FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
]


def build(verbose=False, jobs=None):
    """Compiles every source (in parallel, up to `jobs` at once) and links
    libgqf.so."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    objs = [os.path.join(HERE, "_build", src.replace(".hip", ".o")) for src in SOURCES]

    def compile_one(k):
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, SOURCES[k]), "-o", objs[k]]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(compile_one, range(len(SOURCES))))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
