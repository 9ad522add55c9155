"""Builds tests/point_host/libpointhost.so: the device point operations of
include/gqf_device.h compiled for the host (test infrastructure)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "point_host.hip")
LIB = os.path.join(HERE, "libpointhost.so")
DEPS = [SRC, os.path.join(ROOT, "include", "gqf_device.h"), os.path.join(ROOT, "include", "gqf.h")]


def build(force=False):
    if not force and os.path.exists(LIB) and all(os.path.getmtime(LIB) >= os.path.getmtime(d) for d in DEPS):
        return LIB
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    # a name of its own per process: parallel test workers may build at once
    tmp = f"{LIB}.{os.getpid()}.tmp"
    subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-O2", "-fPIC", "-shared", "-std=c++17",
                    "-I", os.path.join(ROOT, "include"), SRC, "-o", tmp], check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force=True))
