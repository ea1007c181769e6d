"""The allocation-free consistency filter (mvs_survey_consistency_noalloc)
against the allocating entry point it varies (mvs_survey_consistency),
through the C ABI: per-pixel agree counts and consistent-pixel counts
identical (integer work, bit-exact), on two calls in a row whose second view
is smaller than the first (the shared scratch keeps the first call's
contents), under both rules; the device's free memory is unchanged across the
calls; the stated limits are refused before any device work.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from aerial_mvs import capi
from conftest import survey_cached

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _survey_maps(ctx, s):
    spec = s.spec
    prm = capi.params(taps=spec.taps, stride=spec.stride, iterations=spec.iterations, fast=True)
    jobs = [(i, list(s.neighbours[i][:4]), spec.depth_range) for i in range(len(s.views))]
    sv = capi.Survey(ctx, s.views)
    sv.depth_maps(jobs, prm)
    return sv


def test_noalloc_consistency_matches_the_allocating_entry_point(gpu_ctx):
    small, tiny = survey_cached("small"), survey_cached("tiny")
    _say("no-alloc consistency: depth maps of the small and tiny surveys")
    sv_small, sv_tiny = _survey_maps(gpu_ctx, small), _survey_maps(gpu_ctx, tiny)
    try:
        cases = []
        for sv, s in ((sv_small, small), (sv_tiny, tiny)):
            ref = s.reference_index()
            nb = list(s.neighbours[ref][:5])
            cases.append((sv, ref, nb, s.views[ref].w * s.views[ref].h))
        assert cases[1][3] < cases[0][3]  # the second call's view is the smaller
        for rule in (capi.MVS_CONS_REPROJ, capi.MVS_CONS_REFERENCE):
            # the allocating entry point first (its buffers are then in place)
            want = [sv.consistency(ref, nb, rule=rule) for (sv, ref, nb, _) in cases]
            gpu_ctx.synchronize()
            free0 = _free_bytes()
            got = []
            for (sv, ref, nb, npx) in cases:
                _say(f"  rule {rule}: no-alloc call on a {npx}-pixel view")
                got.append(sv.consistency(ref, nb, rule=rule, noalloc=True))
            gpu_ctx.synchronize()
            free1 = _free_bytes()
            _say(f"  device free memory {free0} -> {free1}")
            assert free1 == free0
            for (ga, gn), (wa, wn) in zip(got, want):
                assert gn == wn and gn > 0
                assert np.array_equal(ga, wa)
            # counts only, no agree plane
            _, n_only = cases[1][0].consistency(cases[1][1], cases[1][2], rule=rule,
                                                download=False, noalloc=True)
            assert n_only == want[1][1]
    finally:
        sv_small.close()
        sv_tiny.close()


def test_noalloc_consistency_limits(gpu_ctx):
    tiny = survey_cached("tiny")
    sv = _survey_maps(gpu_ctx, tiny)
    try:
        ref = tiny.reference_index()
        nb = [j for j in range(len(tiny.views)) if j != ref]
        many = (nb * 3)[:capi.MVS_NOALLOC_MAX_NBRS + 1]
        _say("no-alloc consistency: too many neighbours")
        with pytest.raises(capi.MvsError) as e:
            sv.consistency(ref, many, noalloc=True)
        assert e.value.code == capi.MVS_ERR_ARG
        with pytest.raises(capi.MvsError) as e:
            sv.consistency(ref, nb[:2], rule=7, noalloc=True)
        assert e.value.code == capi.MVS_ERR_ARG
    finally:
        sv.close()
    # a view one row past the pixel limit: refused before any device work
    w = 4096
    h = capi.MVS_NOALLOC_MAX_PIXELS // w + 1
    big = tiny.views[0].__class__(np.zeros((h, w), np.uint8), 900.0, w / 2, h / 2, 0.0, 0.0,
                                  tiny.views[0].R, tiny.views[0].C)
    _say(f"no-alloc consistency: a {w}x{h} view")
    sv = capi.Survey(gpu_ctx, [big], options=capi.MVS_SURVEY_LEAN_VIEWS |
                     capi.MVS_SURVEY_DEPTH_ONLY)
    try:
        sv.set_depth(0, np.full((h, w), 50.0, np.float32))
        with pytest.raises(capi.MvsError) as e:
            sv.consistency(0, [0], noalloc=True)
        assert e.value.code == capi.MVS_ERR_ARG
    finally:
        sv.close()
