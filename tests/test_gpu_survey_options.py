"""Survey residency options (mvs_survey_set_options), the layout that keeps a
survey of thousands of views in one GPU's memory: lean views (no f32 corner
plane) and depth-only maps give the same depth maps and consistency counts as
a full survey under the same schedule, and the calls they cannot serve fail
loudly instead of reading a missing plane.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from aerial_mvs import capi

pytestmark = pytest.mark.gpu

LEAN = capi.MVS_SURVEY_LEAN_VIEWS | capi.MVS_SURVEY_DEPTH_ONLY


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("split", [-1, -2])
def test_lean_depth_only_survey_matches_full(gpu_ctx, tiny_survey, split):
    s = tiny_survey
    spec = s.spec
    prm = capi.params(taps=spec.taps, stride=spec.stride, iterations=spec.iterations, fast=True,
                      refine_steps=2)
    jobs = [(i, list(s.neighbours[i][:4]), spec.depth_range) for i in range(len(s.views))]
    _say(f"lean + depth-only survey vs full survey, split {split}")
    gpu_ctx.set_schedule(split, 1, 16, False)
    full = capi.Survey(gpu_ctx, s.views)
    lean = capi.Survey(gpu_ctx, s.views, options=LEAN)
    try:
        full.depth_maps(jobs, prm)
        lean.depth_maps(jobs, prm)
        for (i, nb, _) in jobs:
            df, nf, _ = full.download(i)
            dl, nl, cl = lean.download(i)
            assert nl is None and cl is None
            assert np.array_equal(df, dl), i
            assert (dl > 0).mean() > 0.5
            af, nfc = full.consistency(i, nb, rule=capi.MVS_CONS_REPROJ)
            al, nlc = lean.consistency(i, nb, rule=capi.MVS_CONS_REPROJ)
            assert nfc == nlc and np.array_equal(af, al)
        # what a lean, depth-only survey cannot serve fails with MVS_ERR_ARG
        h, w = lean.shapes[0]
        d = np.zeros((h, w), np.float32)
        n = np.zeros((h, w, 3), np.float32)
        rc = capi.lib().mvs_survey_download(lean._h, 0, d.ctypes.data, n.ctypes.data, None)
        assert rc == capi.MVS_ERR_ARG
        with pytest.raises(capi.MvsError):
            lean.depth_maps(jobs[:1], capi.params(taps=spec.taps, stride=spec.stride,
                                                  iterations=spec.iterations))  # exact mode
        assert capi.lib().mvs_survey_set_options(lean._h, C.c_int32(8).value) == capi.MVS_ERR_ARG
    finally:
        gpu_ctx.set_schedule(0, 0, 16, True)
        full.close()
        lean.close()
