"""Uneven batches: views of different sizes, neighbour counts, window sizes,
strides and depth ranges in one survey call (mvs_survey_depth_maps_ex).

Bars: with the exact kernels every view's map in the batch is identical to a
one-view call with the same parameters and to the oracle's map (bit-exact
work); with the fast kernels the batch equals one-view calls (deterministic)
and stays within 0.5% of the oracle on >= 99% of pixels valid in both.
"""
import sys

import numpy as np
import pytest

from aerial_mvs import capi, survey

import oracle as orc

pytestmark = pytest.mark.gpu

_CACHE = {}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _uneven():
    if "s" not in _CACHE:
        _CACHE["s"] = survey.make_uneven_survey("uneven_small")
    return _CACHE["s"]


def _jobs(s, per, fast, refine=0):
    jobs, prms = [], []
    for i, p in enumerate(per):
        jobs.append((i, list(s.neighbours[i][:p["k"]]), p["range"]))
        prms.append(capi.params(taps=p["taps"], stride=p["stride"], iterations=3,
                                refine_steps=refine, fast=fast))
    return jobs, prms


@pytest.mark.parametrize("fast", [False, True])
def test_uneven_batch_matches_single_views(gpu_ctx, fast):
    s, per = _uneven()
    jobs, prms = _jobs(s, per, fast, refine=2 if fast else 0)
    _say(f"uneven batch: {len(jobs)} views, windows {sorted(set(p['taps'] for p in per))}, "
         f"K {min(p['k'] for p in per)}..{max(p['k'] for p in per)}, fast={fast}")
    sv = capi.Survey(gpu_ctx, s.views)
    try:
        sv.depth_maps(jobs, prms)
        for (i, nb, rng), prm in zip(jobs, prms):
            db, nbm, cb = sv.download(i)
            d1, n1, c1 = gpu_ctx.depth_map(s.views[i], [s.views[j] for j in nb], rng[0], rng[1],
                                           prm)
            assert np.array_equal(db, d1), i
            assert np.array_equal(cb, c1), i
            # oblique views matched against nadir neighbours keep fewer pixels
            assert (db > 0).mean() > 0.1, i
    finally:
        sv.close()


def test_uneven_fast_batch_one_launch_per_window(gpu_ctx):
    """Fast jobs on undistorted views share one persistent launch per window
    and half-sweep, whatever their neighbour counts, sizes and strides."""
    s, per = _uneven()
    jobs, prms = _jobs(s, per, True)
    windows = sorted(set(p["taps"] for p in per))
    ks = sorted(set(p["k"] for p in per))
    sv = capi.Survey(gpu_ctx, s.views)
    try:
        sv.depth_maps(jobs, prms)  # tunes the schedules
        gpu_ctx.reset_stats()
        sv.depth_maps(jobs, prms)
        st = gpu_ctx.stats()
        half_sweeps = 2 * 3
        _say(f"fast batch: windows {windows}, K {ks}: {st['sweep_kernels']} sweep kernels "
             f"in {half_sweeps} half-sweeps, schedule {st['split']}")
        assert len(ks) > 2
        assert st["sweep_kernels"] == half_sweeps * len(windows)
        assert st["split"] in (-1, -2)
    finally:
        sv.close()


def test_uneven_batch_matches_oracle(gpu_ctx):
    s, per = _uneven()
    jobs, prms = _jobs(s, per, False)
    sv = capi.Survey(gpu_ctx, s.views)
    try:
        sv.depth_maps(jobs, prms)
        # one view of each window size, nadir and oblique
        picks = {}
        for i, p in enumerate(per):
            picks.setdefault((p["taps"], p["kind"]), i)
        for (taps, kind), i in sorted(picks.items()):
            p = per[i]
            _say(f"oracle view {i}: {kind} {s.views[i].w}x{s.views[i].h} {taps}x{taps}/"
                 f"{p['stride']} K={p['k']}")
            o = orc.OracleDepth(s.views[i], [s.views[j] for j in jobs[i][1]], *p["range"],
                                taps=taps, stride=p["stride"])
            od, _, _, _ = o.depth_map(iterations=3)
            db, _, _ = sv.download(i)
            assert np.array_equal(db, od), i
    finally:
        sv.close()


def test_uneven_fast_within_tolerance(gpu_ctx):
    s, per = _uneven()
    jobs, prms = _jobs(s, per, True)
    sv = capi.Survey(gpu_ctx, s.views)
    try:
        sv.depth_maps(jobs, prms)
        for i in (0, 3, 15, 16):
            p = per[i]
            o = orc.OracleDepth(s.views[i], [s.views[j] for j in jobs[i][1]], *p["range"],
                                taps=p["taps"], stride=p["stride"])
            od, _, _, _ = o.depth_map(iterations=3)
            db, _, _ = sv.download(i)
            both = (db > 0) & (od > 0)
            within = np.mean(np.abs(db[both] - od[both]) / od[both] <= 0.005)
            _say(f"fast view {i} ({p['kind']}, {p['taps']}x{p['taps']}): within 0.5% {within:.5f}")
            assert within >= 0.99
    finally:
        sv.close()
