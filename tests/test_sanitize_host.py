"""The host code under AddressSanitizer and UBSan.

The engine's host side (argument checks, the `.depth` / `.view` / `.dsm` /
PLY readers and writers, neighbour selection, fusion voxel sizing, the
prepared-size arithmetic: everything `tests/` reaches without a GPU), the
synthetic-survey generator and the oracle are rebuilt with -fsanitize=address
and -fsanitize=undefined (clang, abort on the first finding), and the whole
CPU suite runs again on those builds in a child process with clang's ASan
runtime preloaded. The reference's safety here comes from Rust's ownership
rules (buffers indexed through checked slices, dense.rs:416-417); this is the
check that stands in for them on the C and C++ side.

The GPU-side host paths (batches, graphs, fusion passes) need a device and are
not exercised here. `python aerial-3d-model_amd/aerial_mvs/build.py
sanitized_gpu` builds the same instrumentation; a GPU run of it (ROCm 7.2,
ASan runtime preloaded ahead of the uninstrumented HIP runtime) aborted inside
the first device initialisation (mvs_create) without a sanitizer report, so
those paths stay unsanitized.
"""
import importlib.util
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.skipif(os.environ.get("AERIAL_MVS_SANITIZED") == "1",
                                reason="already running under the sanitizers")


def test_cpu_suite_under_asan_and_ubsan(tmp_path):
    from aerial_mvs import build as mvs_build
    spec = importlib.util.spec_from_file_location("oracle_build",
                                                  os.path.join(REPO, "oracle", "build.py"))
    oracle_build = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(oracle_build)

    out = str(tmp_path / "san")
    lib_mvs, lib_synth = mvs_build.build_sanitized(out, device=True)
    lib_oracle = oracle_build.build_sanitized(out)
    env = dict(os.environ)
    pre = mvs_build.sanitizer_runtime()
    env["LD_PRELOAD"] = pre + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    env.update(AERIAL_MVS_SANITIZED="1", AERIAL_MVS_LIB=lib_mvs, AERIAL_MVS_SYNTH_LIB=lib_synth,
               AERIAL_MVS_ORACLE_LIB=lib_oracle,
               # Python itself is not instrumented and keeps its arenas to the end
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    # the child loads the instrumented libraries, with the runtime first
    probe = ("import sys; sys.path[:0] = [%r, %r]; from aerial_mvs import capi, survey; "
             "import oracle; capi.lib(); survey._synth(); oracle.lib(); "
             "print(open('/proc/self/maps').read())"
             % (os.path.join(REPO, "aerial-3d-model_amd"), os.path.join(REPO, "oracle")))
    maps = subprocess.run([sys.executable, "-c", probe], cwd=REPO, env=env, capture_output=True,
                          text=True, timeout=300, check=True).stdout
    for lib in (pre, lib_mvs, lib_synth, lib_oracle):
        assert lib in maps, lib
    cmd = [sys.executable, "-m", "pytest", os.path.join(REPO, "tests"), "-x", "-q",
           "-m", "not gpu", "-p", "no:cacheprovider"]
    r = subprocess.run(cmd, cwd=REPO, env=env, capture_output=True, text=True, timeout=1500)
    log = r.stdout + r.stderr
    print(log[-3000:])
    assert "AddressSanitizer" not in log and "runtime error:" not in log
    assert r.returncode == 0
    assert " passed" in log
