"""Builds tests/facade/libfacade_driver.so (test infrastructure): the C++ facade driver,
linked against axomae_amd/libnova_hip.so."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "libfacade_driver.so")


def build():
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "axomae_amd", "csrc"), os.path.join(HERE, "facade_driver.cpp"),
           "-L", os.path.join(ROOT, "axomae_amd"), "-lnova_hip", "-Wl,-rpath,$ORIGIN/../../axomae_amd", "-o", LIB + f".{os.getpid()}.tmp"]
    subprocess.check_call(cmd)
    os.replace(LIB + f".{os.getpid()}.tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build())
