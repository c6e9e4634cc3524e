"""Build recipe for the CPU oracle (test infrastructure only): compiles oracle/obvhs_oracle.cpp with g++ into
oracle/liboracle.so.  Flags keep the arithmetic of the reference's default x86-64 target: SSE2 scalar maths, no FMA
contraction, no fast-math.

The reference itself is a Rust crate and no Rust toolchain exists in this image, so there is no oracle/_ref build."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "obvhs_oracle.cpp")
LIB = os.path.join(HERE, "liboracle.so")
FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", "-fno-fast-math",
         "-msse2", "-mfpmath=sse", "-mno-fma", "-Wall"]


def build():
    tmp = LIB + ".tmp"
    cmd = ["g++"] + FLAGS + [SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
    sys.exit(0)
