"""aerial_mvs — MI355X (gfx950) dense multi-view-stereo depth engine.

Drop-in for the PatchMatch depth-map stage of stitch-core
(crates/stitch-core/src/recon/dense.rs). The engine is the HIP library
``libaerial_mvs.so`` behind the C ABI in ``include/aerial_mvs.h``; this package
holds its build recipe, a ctypes binding (``capi``) and the seeded synthetic
survey generator (``survey``) used by the tests and the benchmark.
"""
from . import build  # noqa: F401

__all__ = ["build", "capi", "survey"]
