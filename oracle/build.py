"""Build recipe for the CPU oracle (test infrastructure only).

Compiles ``dense_oracle.c`` into ``oracle/liboracle.so`` with -ffp-contract=off
and no fast-math, so its f32/f64 arithmetic rounds like the Rust reference.

The reference itself (crates/stitch-core, Rust) cannot be built in this image:
there is no Rust toolchain and its crates (nalgebra, rayon, image, serde) are
not vendored. ``oracle/_ref/`` therefore stays empty.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

FLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
         "-Wall", "-Wextra", "-Wno-unused-parameter", "-pthread", "-std=c11"]


def build():
    tmp = LIB + ".tmp"
    cmd = ["gcc"] + FLAGS + [os.path.join(HERE, "dense_oracle.c"), "-o", tmp, "-lm"]
    print("+ " + " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
