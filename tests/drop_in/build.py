"""Builds tests/drop_in/libdropin.so: a reference-style caller of libgqf
(the reference's wrapper calls and its point_bulk_insert launch, compiled
against include/gqf.h and include/gqf_device.h, linked to
cqf_amd/libgqf.so).  Test infrastructure; cross-compiles without a GPU."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "dropin.hip")
LIB = os.path.join(HERE, "libdropin.so")


def build():
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    arch = os.environ.get("GQF_ARCH", "gfx950")
    tmp = f"{LIB}.{os.getpid()}.tmp"
    subprocess.run([hipcc, f"--offload-arch={arch}", "-O3", "-std=c++17", "-fPIC", "-shared",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-I", os.path.join(ROOT, "include"), SRC,
                    "-L", os.path.join(ROOT, "cqf_amd"), "-lgqf", "-Wl,-rpath,$ORIGIN/../../cqf_amd",
                    "-o", tmp], check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
