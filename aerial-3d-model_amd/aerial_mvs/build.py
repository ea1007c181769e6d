"""Build recipes for the native libraries (run on the CPU; hipcc cross-compiles gfx950).

- ``libaerial_mvs.so``: the HIP engine behind ``include/aerial_mvs.h``.
- ``libsynth.so``: the seeded synthetic-survey generator (test/benchmark workload).
- ``libsynth_gpu.so``: the same generator on the GPU, for surveys too large to
  render on the host (test/benchmark workload; not part of the engine).

The oracle under ``oracle/`` is built by ``oracle/build.py``; this module does not
touch it.
"""
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
SRC_ROOT = os.path.dirname(PKG_DIR)
CSRC = os.path.join(SRC_ROOT, "csrc")
REPO = os.path.dirname(SRC_ROOT)

LIB_MVS = os.path.join(PKG_DIR, "libaerial_mvs.so")
LIB_SYNTH = os.path.join(PKG_DIR, "libsynth.so")
LIB_SYNTH_GPU = os.path.join(PKG_DIR, "libsynth_gpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the reference (Rust) never fuses a*b+c; the kernels match
# its rounding operation by operation. IEEE-correct f32 division/sqrt are the
# HIP defaults and are requested explicitly. Packed-FP32 instructions are
# disabled for the whole build.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-rdc",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]

C_FLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
           "-shared", "-Wall", "-pthread"]


def _run(cmd):
    print("+ " + " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


UNITS = ["mvs_api.hip", "mvs_kernels.hip", "mvs_pm_t5.hip", "mvs_pm_t7.hip", "mvs_pm_t11.hip",
         "mvs_pm_t5f.hip", "mvs_pm_t7f.hip", "mvs_pm_t11f.hip", "mvs_pm_lane.hip",
         "mvs_refine.hip", "mvs_refine_lane.hip", "mvs_fusion.hip", "mvs_prep.hip",
         "mvs_dsm.hip"]


def build_mvs(extra=None):
    """Compile the units in parallel into objects, then link the library."""
    objdir = os.path.join(SRC_ROOT, "_build")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in HIP_FLAGS if f != "-shared"] + (extra or [])
    procs = []
    for u in UNITS:
        obj = os.path.join(objdir, u.replace(".hip", ".o"))
        cmd = [HIPCC] + cflags + ["-c", os.path.join(CSRC, u), "-o", obj]
        print("+ " + " ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), obj))
    objs = []
    for p, obj in procs:
        if p.wait() != 0:
            raise RuntimeError(f"hipcc failed for {obj}")
        objs.append(obj)
    tmp = LIB_MVS + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-fno-gpu-rdc"] + objs +
         ["-o", tmp])
    os.replace(tmp, LIB_MVS)
    return LIB_MVS


def build_synth():
    tmp = LIB_SYNTH + ".tmp"
    _run(["gcc"] + C_FLAGS + [os.path.join(CSRC, "synth.c"), "-o", tmp, "-lm"])
    os.replace(tmp, LIB_SYNTH)
    return LIB_SYNTH


def build_synth_gpu():
    tmp = LIB_SYNTH_GPU + ".tmp"
    _run([HIPCC] + HIP_FLAGS + [os.path.join(CSRC, "synth_gpu.hip"), "-o", tmp])
    os.replace(tmp, LIB_SYNTH_GPU)
    return LIB_SYNTH_GPU


def build_all():
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found; the HIP engine cannot be built")
    build_synth()
    build_synth_gpu()
    build_mvs()


# ---- sanitizer builds of the host code (AddressSanitizer + UBSan) ----
#
# Host code only: every -fsanitize= goes to the host pass (-Xarch_host), the
# GPU code is built as usual. Any finding aborts (-fno-sanitize-recover).
# The libraries need clang's ASan runtime loaded first in the process
# (LD_PRELOAD=sanitizer_runtime()); select them with AERIAL_MVS_LIB and
# AERIAL_MVS_SYNTH_LIB.
SAN_DIR = os.path.join(SRC_ROOT, "_san")
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer",
            "-Xarch_host", "-g"]
SAN_C = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=all",
         "-fno-omit-frame-pointer", "-g"]
CLANG = os.environ.get("AERIAL_MVS_CLANG", "/opt/rocm/llvm/bin/clang")


def sanitizer_runtime():
    """Path of clang's shared ASan runtime (it carries the UBSan handlers)."""
    out = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], check=True,
                         capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(out) or not os.path.exists(out):
        raise RuntimeError("clang's ASan runtime (libclang_rt.asan-x86_64.so) not found")
    return out


def build_sanitized(out_dir=SAN_DIR, device=False):
    """libaerial_mvs.so and libsynth.so with instrumented host code in out_dir.
    device=False builds the host pass only (--offload-host-only: the library
    loads and runs its host-side entry points but has no GPU kernels), which
    is what a machine without a GPU can exercise; device=True adds the gfx950
    code for runs on a GPU."""
    os.makedirs(out_dir, exist_ok=True)
    objdir = os.path.join(out_dir, "obj")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in HIP_FLAGS if f != "-shared"] + SAN_HOST
    if not device:
        cflags.append("--offload-host-only")
    procs = []
    for u in UNITS:
        obj = os.path.join(objdir, u.replace(".hip", ".o"))
        cmd = [HIPCC] + cflags + ["-c", os.path.join(CSRC, u), "-o", obj]
        print("+ " + " ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), obj))
    objs = []
    for p, obj in procs:
        if p.wait() != 0:
            raise RuntimeError(f"hipcc failed for {obj}")
        objs.append(obj)
    lib_mvs = os.path.join(out_dir, "libaerial_mvs.so")
    link = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-fno-gpu-rdc"] + SAN_HOST
    if not device:
        link.append("--offload-host-only")
    _run(link + ["-shared-libsan"] + objs + ["-o", lib_mvs])
    lib_synth = os.path.join(out_dir, "libsynth.so")
    c_flags = C_FLAGS + SAN_C
    _run([CLANG] + c_flags + ["-shared-libsan", os.path.join(CSRC, "synth.c"), "-o", lib_synth,
                              "-lm"])
    return lib_mvs, lib_synth


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "mvs": build_mvs, "synth": build_synth,
         "synth_gpu": build_synth_gpu, "sanitized": build_sanitized,
         "sanitized_gpu": lambda: build_sanitized(device=True)}[w]()
