"""Builds tests/emu/libnova_emu.so: the device path stages compiled for the host CPU
(same flags as the HIP library, so no contraction and IEEE division)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libnova_emu.so")


def build():
    src = os.path.join(HERE, "host_emu.hip")
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950", "-ffp-contract=off",
           "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-fast-math", "-I", os.path.join(HERE, "..", "..", "include"), src, "-o", LIB + f".{os.getpid()}.tmp"]
    subprocess.check_call(cmd)
    os.replace(LIB + f".{os.getpid()}.tmp", LIB)  # per-process temporary: parallel test workers may build at once
    return LIB


if __name__ == "__main__":
    print(build())
