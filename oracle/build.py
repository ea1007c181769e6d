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


def build_sanitized(out_dir, clang="/opt/rocm/llvm/bin/clang"):
    """The same library with AddressSanitizer + UBSan (clang, shared runtime;
    load clang's libclang_rt.asan first), into out_dir."""
    os.makedirs(out_dir, exist_ok=True)
    lib = os.path.join(out_dir, "liboracle.so")
    cmd = [clang] + FLAGS + ["-fsanitize=address", "-fsanitize=undefined",
                             "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g",
                             "-shared-libsan", os.path.join(HERE, "dense_oracle.c"), "-o", lib,
                             "-lm"]
    print("+ " + " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return lib


if __name__ == "__main__":
    build()
