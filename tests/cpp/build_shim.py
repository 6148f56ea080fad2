"""Build tests/cpp/_build/libformpt_shim.so: formpt_shim.cpp (the reference's formPT call
written against include/cqgpu_aointegrals.hpp) linked against libcqgpu.so.  Host code only
(g++); run by the package build and, when the library is missing, by the test that uses it."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "libformpt_shim.so")


def build() -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    lib_dir = os.path.join(ROOT, "chronusq_public_amd")
    cmd = ["g++", "-std=c++14", "-O2", "-fPIC", "-shared", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "formpt_shim.cpp"), "-o", OUT + ".tmp", "-L", lib_dir, "-lcqgpu",
           "-Wl,-rpath," + lib_dir]
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build())
