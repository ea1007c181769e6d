"""The GPU survey generator (libsynth_gpu.so), workload of the large-survey
benchmark: it renders the CPU generator's scene (libsynth.so) — same
heightfield, boxes, texture and sensor noise — with an exact ray/bilinear-
patch intersection instead of the CPU's conservative march; and the small
version of the large survey runs end to end through a lean, depth-only
survey with fast arithmetic and refinement.
"""
import sys

import numpy as np
import pytest

from aerial_mvs import capi, survey

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_gpu_render_matches_cpu_generator():
    # the views see only the inside of the grid: outside it the CPU generator
    # holds the border height, the GPU one has no surface
    args = (-30.0, -30.0, 300.0, 300.0, 0.1, 40.0, 6, 0.5, 12.0, 0xA3D0E001, 0.2, 5)
    boxes = np.array([[20.0, 15.0, 4.0, 6.0, 8.0], [35.0, 30.0, 5.0, 3.0, 15.0],
                      [10.0, 40.0, 3.0, 3.0, 4.0], [45.0, 8.0, 6.0, 5.0, 22.0]])
    L = survey._synth()
    cpu = L.synth_scene_create(*args, 8)
    L.synth_scene_add_boxes(cpu, np.ascontiguousarray(boxes.ravel()), len(boxes))
    gpu = survey.GpuScene(*args)
    gpu.add_boxes(boxes)
    try:
        ci, gi = np.zeros(6), gpu.info()
        L.synth_scene_info(cpu, ci)
        assert abs(ci[0] - gi[0]) < 1e-5 and abs(ci[1] - gi[1]) < 1e-5  # hmin, hmax
        w, h, f = 320, 240, 281.25
        for kind, (c, yaw, pitch) in (("nadir", ((25.0, 25.0, 70.0), 90.0, -90.0)),
                                      ("oblique", ((120.0, 0.0, 70.0), 10.0, -45.0))):
            R, Cc = survey.from_angles(c, yaw, pitch, 0.0)
            img_c = np.zeros((h, w), np.uint8)
            dep_c = np.zeros((h, w), np.float64)
            L.synth_render(cpu, w, h, f, w / 2, h / 2, 0.0, 0.0, np.ascontiguousarray(R.ravel()),
                           Cc, 77, 1.0, img_c.ravel(), dep_c.ctypes.data, 8)
            img_g, dep_g = gpu.render(w, h, f, w / 2, h / 2, R, Cc, 77, 1.0, True)
            hit = (dep_c > 0) & (dep_g > 0)
            rel = np.abs(dep_g[hit] - dep_c[hit]) / dep_c[hit]
            same_px = np.mean(img_g == img_c)
            _say(f"{kind}: hits agree {np.mean((dep_c > 0) == (dep_g > 0)):.5f}, depth rel "
                 f"median {np.median(rel):.2e} p99.9 {np.quantile(rel, 0.999):.2e}, "
                 f"grey equal {same_px:.5f}")
            assert np.all(dep_c > 0) and np.mean((dep_c > 0) == (dep_g > 0)) > 0.999
            assert np.median(rel) < 1e-7 and np.quantile(rel, 0.999) < 1e-4
            assert same_px > 0.99
            assert np.mean(np.abs(img_g.astype(int) - img_c) <= 1) > 0.999
    finally:
        L.synth_scene_destroy(cpu)
        gpu.close()


def test_large_survey_small_end_to_end(gpu_ctx):
    surv, per = survey.make_large_survey("large_small")
    spec = surv.extra["large_spec"]
    assert len(surv.views) == 24 + 16
    centres = surv.centres()
    jobs = []
    for i in range(len(surv.views)):
        rg = per[i]["range"]
        nb = capi.select_neighbours(centres, i, surv.neighbours[i], rg[0], spec.neighbours)
        if nb:
            jobs.append((i, nb, rg))
    assert len(jobs) >= 36
    prm = capi.params(taps=spec.taps, stride=spec.stride, iterations=spec.iterations,
                      refine_steps=3, fast=True)
    sv = capi.Survey(gpu_ctx, surv.views,
                     options=capi.MVS_SURVEY_LEAN_VIEWS | capi.MVS_SURVEY_DEPTH_ONLY)
    try:
        sv.depth_maps(jobs, prm)
        for kind in ("nadir", "oblique"):
            within, valid = [], []
            for (i, _, _) in jobs:
                if per[i]["kind"] != kind:
                    continue
                d, _, _ = sv.download(i)
                gt = surv.views[i].depth_gt
                both = (d > 0) & (gt > 0)
                valid.append(both.mean())
                within.append(np.mean(np.abs(d[both] - gt[both]) / gt[both] <= 0.005))
            _say(f"large_small {kind}: valid {np.mean(valid):.3f}, within 0.5% "
                 f"{np.mean(within):.4f} over {len(within)} views")
            # oblique views see far ground at grazing angles from only three
            # stations of the same heading: fewer pixels pass the cost threshold
            assert np.mean(valid) > (0.6 if kind == "nadir" else 0.3)
            assert np.mean(within) > (0.93 if kind == "nadir" else 0.85)
    finally:
        sv.close()
