"""Build helpers: the gfx950 library (hipcc) and the CPU oracle (g++)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def build_library(jobs: int = 8) -> str:
    """Compiles every HIP source for gfx950 into rkhs_ba_amd/librkhs_ba_amd.so."""
    csrc = os.path.join(HERE, "csrc")
    # always rebuild: the build check must compile, not reuse an old object
    subprocess.run(["make", "-s", "-C", csrc, "clean"], check=True)
    subprocess.run(["make", "-s", "-C", csrc, f"-j{jobs}"], check=True)
    return os.path.join(HERE, "librkhs_ba_amd.so")


FACADE_CHECK = os.path.join(HERE, "facade_check.bin")


def build_facade_check() -> str:
    """Compiles tests/cpp/facade_check.cpp against the C++ facade
    (include/cvo/CvoGPU.hpp) and the library."""
    src = os.path.join(ROOT, "tests", "cpp", "facade_check.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), src, "-L" + HERE,
                    "-lrkhs_ba_amd", "-Wl,-rpath,$ORIGIN", "-o", FACADE_CHECK], check=True)
    return FACADE_CHECK


def build_oracle() -> str:
    """Compiles the CPU oracle (test infrastructure) into oracle/_build/."""
    od = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-s", "-C", od, "clean"], check=True)
    subprocess.run(["make", "-s", "-C", od], check=True)
    return os.path.join(od, "_build", "liboracle_cvo.so")


if __name__ == "__main__":
    print(build_library())
    print(build_facade_check())
    print(build_oracle())
