"""Benchmark: dense-MVS depth maps on one MI355X.

Workload (the survey the metric is quoted on): 77 nadir views at 2000x1500 u8
grey, seeded synthetic heightfield survey; each view is a reference view with
its 8 nearest-GPS neighbours, 7x7 NCC window, 6 PatchMatch iterations, then the
geometric-consistency filter (reprojection <= 1 px in >= 3 views). A "step" is
one reference view's depth map (initialisation + 12 half-sweeps + Gauss-Newton
refinement + threshold).

Three paths:
  - batched+refined (the headline `value`): all K views as one batch - every
    half-sweep of all views is one persistent launch, run as a cached graph -
    with the fast-arithmetic kernels (costs within 2e-5 of the exact ones) and
    5 Gauss-Newton steps, plus the consistency filter, images resident in HBM;
  - plain: one view per call with the exact kernels and no refinement - the
    bit-identical drop-in of the reference's algorithm - on a subset;
  - CPU: the oracle restatement of recon/dense.rs on a bounded crop.
Depth errors are against the FP64 ray-cast ground truth of the scene.

Prints one JSON line (rank 0 of one GPU).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(REPO, "aerial-3d-model_amd")]

METRIC = ("depth Mpix/s (7x7 NCC, 8 nbrs) batched+refined GPU vs plain GPU vs CPU; depth err")


def metric_for(taps, k):
    """The metric string of a survey run with a taps x taps window and k
    neighbours (the 7x7, 8-neighbour survey's is METRIC)."""
    return (f"depth Mpix/s ({taps}x{taps} NCC, {k:g} nbrs) batched+refined GPU vs plain GPU vs "
            "CPU; depth err")


def host_cpus():
    """The host's CPUs as this process sees them: the logical CPUs of the
    machine, the CPUs this process may run on, the thread budget the
    environment sets (OMP_NUM_THREADS), and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    return dict(nproc=os.cpu_count(), affinity=affinity,
                omp_num_threads=int(omp) if omp and omp.isdigit() else None, model=model)


def default_threads():
    """Every CPU this process may run on, capped by the environment's thread
    budget (OMP_NUM_THREADS) where one is set: on a shared host that budget
    is this job's share of the cores."""
    h = host_cpus()
    n = h["affinity"]
    if h["omp_num_threads"]:
        n = min(n, h["omp_num_threads"])
    return max(1, n)


PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# HBM-side traffic of the half-sweep kernel per cost evaluation, from rocprofv3
# FETCH_SIZE / WRITE_SIZE over the 12 half-sweeps of an 8-view batch of this
# workload, scaled by a 1 GiB device copy (read counter x2.000, write x1.000);
# profiles/k_pml_sweep_fast_8x2000x1500_k8_7x7_counters.md (fast arithmetic,
# one lane per pixel) and profiles/k_pm_sweep_exact_8x2000x1500_k8_7x7_counters.md
# hold the per-launch figures.
BYTES_PER_EVAL = {"exact": (442.3, 6.55), "fast": (165.6, 8.06)}


# The roofline object's "bound" names the compute roofline ("mfma": the FP32
# peak of 157.3 TFLOP/s, MI355X's figure for vector and matrix FP32 alike);
# the half-sweep kernel itself computes on the vector ALUs, and what limits it
# is the texture-address unit of its sample gathers.
LIMITER = "texture-address unit (gathers)"
BOUND_NOTE = ("compute roofline at MI355X's FP32 peak (157.3 TFLOP/s); the kernel issues no "
              "matrix instructions: it computes on the vector ALUs (0.55-0.88 wave-instructions "
              "per CU per clock), and its sample gathers keep the texture-address unit 87-97% "
              "busy, a third to a half of it stalled on the L1 cache (the limiter); achieved "
              "counts the evaluations performed, not the propagation candidates skipped as "
              "scored before")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def flops_per_eval(taps, k):
    # declared algorithmic count of one (pixel, hypothesis) cost evaluation:
    # per window sample 12 flops of plane/reference work, and per sample and
    # source 48 (3x3 apply 15, divide 2, projection 13, bilinear 13, sums 5)
    n = taps * taps
    return n * (12 + 48 * k)


def depth_errors(depth, gt):
    valid = (depth > 0) & (gt > 0)
    if not valid.any():
        return dict(valid=0.0)
    rel = np.abs(depth[valid] - gt[valid]) / gt[valid]
    return dict(valid=float(valid.mean()), within_0p5=float(np.mean(rel <= 0.005)),
                within_0p2=float(np.mean(rel <= 0.002)), median_rel=float(np.median(rel)))


def mean_errors(errs):
    return {k: float(np.mean([e.get(k, 0.0) for e in errs])) for k in errs[0]}


def within(a, b, tol=0.005):
    both = (a > 0) & (b > 0)
    return float(np.mean(np.abs(a[both] - b[both]) / b[both] <= tol)) if both.any() else 0.0


def cpu_baseline(surv, jobs, spec, nthreads, crop_w, crop_h, gpu_ctx, prm_exact, prm_fast):
    """The CPU oracle on a bounded sample: a centred crop of the first
    reference view (its intrinsics shifted accordingly) against the same 8
    full-size sources, I and window as the GPU run; the GPU's maps of the same
    crop (exact and fast kernels, no refinement) are compared with it."""
    import copy
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as orc
    ref, srcs, rng = jobs[0]
    v = surv.views[ref]
    x0, y0 = (v.w - crop_w) // 2, (v.h - crop_h) // 2
    cv = copy.copy(v)
    cv.gray = np.ascontiguousarray(v.gray[y0:y0 + crop_h, x0:x0 + crop_w])
    cv.cx, cv.cy = v.cx - x0, v.cy - y0
    sv = [surv.views[j] for j in srcs]
    o = orc.OracleDepth(cv, sv, rng[0], rng[1], taps=spec.taps, stride=spec.stride)
    t = time.perf_counter()
    od, _, _, ev = o.depth_map(iterations=spec.iterations, nthreads=nthreads)
    dt = time.perf_counter() - t
    ge, _, _ = gpu_ctx.depth_map(cv, sv, rng[0], rng[1], prm_exact)
    gf, _, _ = gpu_ctx.depth_map(cv, sv, rng[0], rng[1], prm_fast)
    mpix = crop_w * crop_h / 1e6
    return dict(value=mpix / dt, unit="depth Mpix/s", cores=nthreads, kind="port",
                host=dict(host_cpus(), threads_used=nthreads,
                          note="threads = the CPUs this process may run on, capped by "
                               "OMP_NUM_THREADS (the job's share of a shared host); rows of "
                               "each half-sweep split over them (the rayon model, "
                               "dense.rs:417)"),
                sample=(f"oracle C restatement of recon/dense.rs, {crop_w}x{crop_h} centre crop "
                        f"of view {ref} vs its {len(srcs)} full-size neighbours, "
                        f"{spec.taps}x{spec.taps} window, {spec.iterations} iterations: "
                        f"{dt:.1f} s, {ev} cost evaluations"),
                seconds=dt,
                sample_exact_identical=float(np.mean(ge == od)),
                sample_fast_within_0p5=within(gf, od),
                sample_fast_valid_agree=float(np.mean((gf > 0) == (od > 0))))


def view_prep_stage(ctx, spec, photos=8):
    """prepare_views' per-photo work (dense.rs:100-112) on the GPU for photos
    twice the survey's working size (4000x3000 RGB8 for c2), to the working
    size (box by 2) and to 0.7 of it (box, then bilinear resize): input Mpix/s
    end to end (upload, kernels, download). Returns the figures, the photo and
    the GPU's outputs (for the CPU-baseline leg's check)."""
    h, w = 2 * spec.height, 2 * spec.width
    rng = np.random.default_rng(11)
    base = rng.integers(0, 256, (h // 8 + 1, w // 8 + 1, 3), dtype=np.uint8)
    photo = np.repeat(np.repeat(base, 8, 0), 8, 1)[:h, :w]
    photo = np.clip(photo.astype(np.int16) + rng.integers(-20, 21, photo.shape), 0, 255)
    photo = np.ascontiguousarray(photo.astype(np.uint8))
    out, outputs = {}, {}
    for tag, m in (("box", max(spec.width, spec.height)),
                   ("box_resize", int(0.7 * max(spec.width, spec.height)))):
        ctx.prepare_view(photo, m)  # warm
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(photos):
            rgb, grey = ctx.prepare_view(photo, m)
        dt = (time.perf_counter() - t) / photos
        out[tag] = dict(max_side=m, out=f"{grey.shape[1]}x{grey.shape[0]}",
                        gpu_mpix_s=w * h / dt / 1e6, gpu_ms_per_photo=dt * 1e3)
        outputs[tag] = (m, rgb, grey)
        log(f"view prep {tag}: {out[tag]}")
    return (dict(photo=f"{w}x{h} RGB8 (synthetic)", photos=photos,
                 path="mvs_prepare_view: upload, box downsample, bilinear resize, grey, download",
                 **out), photo, outputs)


def dsm_stage(ctx, xyz, voxel, reps=5):
    """build_dsm (recon/dsm.rs:128-290) on the GPU from the fused cloud, at the
    Standard profile's cell budget: upload, extent, cell binning, filters,
    download. Returns the figures and the GPU's grid (for the CPU leg's check)."""
    ctx.build_dsm(xyz, voxel, 1)  # warm (buffers)
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        info, h = ctx.build_dsm(xyz, voxel, 1)
    dt = (time.perf_counter() - t) / reps
    out = dict(points=int(len(xyz)), grid=f"{info['nx']}x{info['ny']}", cell_m=info["cell"],
               filled=float(np.isfinite(h).mean()), gpu_ms=dt * 1e3,
               gpu_mpoints_s=len(xyz) / dt / 1e6,
               path="mvs_build_dsm: upload, extent, binning, despike, islands, holes, "
                    "median, download")
    log(f"dsm: {out}")
    return out, (info, h)


def cpu_baseline_dsm(xyz, voxel, got):
    """CPU-baseline leg of the surface-model stage: the oracle on the same
    cloud (one thread), and the GPU's grid checked identical to it."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as orc
    tc = time.perf_counter()
    info, h = orc.build_dsm(xyz, voxel, orc.DSM_BUDGET[1])
    dc = time.perf_counter() - tc
    ginfo, gh = got
    return dict(cpu_ms=dc * 1e3, cpu_mpoints_s=len(xyz) / dc / 1e6, cores=1, kind="port",
                identical=bool(info == ginfo and np.array_equal(h, gh, equal_nan=True)))


def gradient_check(ctx, surv, jobs, spec, sv, n_items, seed=0xA3D0C4):
    """CPU-baseline leg: the refinement's backward kernel at full size. At
    n_items random valid pixels of the refined maps (spread over the views),
    the refined plane perturbed by up to 1% in depth and ~0.05 rad in normal
    (at the refined plane itself the gradient is near zero: it is a
    Gauss-Newton minimum, where a relative error says nothing), the
    Gauss-Newton terms the refinement kernel itself forms (mvs_refine_terms:
    the one-lane-per-pixel kernel on these 8-bit views) at the pair it
    selects: the pair cost against the oracle's f64 pair cost, the gradient
    w.r.t. (depth, normal) against f64 central finite differences (1e-4 m in
    depth, 1e-5 per normal component) and J^T J against the oracle's f64
    J^T J. Items whose bilinear cells change within a difference step are
    skipped (the cost is not differentiable there)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as orc
    from aerial_mvs import capi
    rng = np.random.default_rng(seed)
    prm = capi.params(taps=spec.taps, stride=spec.stride)
    per = max(1, n_items // len(jobs))
    hs = (1e-4, 1e-5, 1e-5, 1e-5)
    rels, dcost, jrels = [], [], []
    tried = 0
    kernels = set()
    tc = time.perf_counter()
    for (i, nb, rg) in jobs:
        v = surv.views[i]
        srcs = [surv.views[j] for j in nb]
        d, n, _ = sv.download(i)
        m = 12
        yy, xx = np.nonzero(d[m:-m, m:-m] > 0)
        if len(xx) == 0:
            continue
        pick = rng.choice(len(xx), min(per, len(xx)), replace=False)
        xs = (xx[pick] + m).astype(np.int32)
        ys = (yy[pick] + m).astype(np.int32)
        nrm = n[ys, xs] + rng.normal(0, 0.05, (len(xs), 3))
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        dep = d[ys, xs] * rng.uniform(0.99, 1.01, len(xs))
        hy = np.concatenate([dep[:, None], nrm], axis=1).astype(np.float32)
        cost, pq, grad, jtj = ctx.refine_terms(v, srcs, rg[0], rg[1], prm, xs, ys, hy)
        kernels.add("k_refine_lane (one lane per pixel)" if ctx.stats()["split"] == -1
                    else "k_refine (one lane per source)")
        o = orc.OracleDepth(v, srcs, rg[0], rg[1], taps=spec.taps, stride=spec.stride)
        for k in range(len(xs)):
            if cost[k] >= 2.0:
                continue
            tried += 1
            h64 = hy[k].astype(np.float64)
            s0, s1 = int(pq[k][0]), int(pq[k][1])
            base = o.pair_terms64(xs[k], ys[k], h64, s0, s1)
            if base is None:
                continue
            dcost.append(abs(float(cost[k]) - base[0]))
            # J^T J against the oracle's f64 matrix, entries weighed like the
            # gradient (normal rows and columns scaled by 1/depth)
            sc = np.array([1.0, 1 / h64[0], 1 / h64[0], 1 / h64[0]])
            S = np.outer(sc, sc)
            jrels.append(np.abs((jtj[k].astype(np.float64) - base[2]) * S).max()
                         / max(np.abs(base[2] * S).max(), 1e-30))
            fd = np.zeros(4)
            ok = True
            for a in range(4):
                hp, hm = h64.copy(), h64.copy()
                hp[a] += hs[a]
                hm[a] -= hs[a]
                p_ = o.pair_cost64(xs[k], ys[k], hp, s0, s1)
                m_ = o.pair_cost64(xs[k], ys[k], hm, s0, s1)
                if p_ is None or m_ is None or p_[2] != base[3] or m_[2] != base[3]:
                    ok = False
                    break
                fd[a] = (p_[0] - m_[0]) / (2 * hs[a])
            if not ok:
                continue
            # error relative to the gradient's size, normal components scaled
            # by 1/depth so both parts weigh alike
            scale = np.array([1.0, 1 / h64[0], 1 / h64[0], 1 / h64[0]])
            e = np.abs(grad[k].astype(np.float64) - fd) / scale
            r = np.abs(fd) / scale
            rels.append(e.max() / max(r.max(), 1e-12))
        del o
    rels = np.array(rels)
    dcost = np.array(dcost)
    out = dict(items=int(tried), checked=int(len(rels)), seconds=time.perf_counter() - tc,
               hypotheses="refined plane, depth x U(0.99, 1.01), normal + N(0, 0.05)",
               fd_steps="1e-4 m depth, 1e-5 per normal component",
               kernel=", ".join(sorted(kernels)),
               bars="gradient and J^T J median < 2e-4, p99 < 3e-3; cost p99 < 3e-5")
    if len(rels) and len(dcost):
        out.update(grad_rel_err_median=float(np.median(rels)),
                   grad_rel_err_p99=float(np.quantile(rels, 0.99)),
                   grad_rel_err_max=float(rels.max()),
                   cost_abs_diff_median=float(np.median(dcost)),
                   cost_abs_diff_p99=float(np.quantile(dcost, 0.99)),
                   cost_abs_diff_max=float(dcost.max()),
                   jtj_rel_err_median=float(np.median(jrels)),
                   jtj_rel_err_p99=float(np.quantile(jrels, 0.99)))
        out["pass"] = bool(out["grad_rel_err_median"] < 2e-4 and out["grad_rel_err_p99"] < 3e-3
                           and out["jtj_rel_err_median"] < 2e-4
                           and out["jtj_rel_err_p99"] < 3e-3
                           and out["cost_abs_diff_p99"] < 3e-5)
    return out


def cpu_baseline_view_prep(photo, outputs):
    """CPU-baseline leg of the view-preparation stage: the oracle on the same
    photo (one thread), and the GPU's outputs checked identical to it."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as orc
    h, w = photo.shape[:2]
    res = {}
    for tag, (m, rgb, grey) in outputs.items():
        tc = time.perf_counter()
        want_rgb, want_grey = orc.prepare_view(photo, m)
        dc = time.perf_counter() - tc
        res[tag] = dict(cpu_mpix_s=w * h / dc / 1e6, cores=1, kind="port",
                        identical=bool(np.array_equal(rgb, want_rgb)
                                       and np.array_equal(grey, want_grey)))
    return res


def reference_grey(views, seed=0xF32):
    """The views as the reference's `.view` grey (dense.rs:104-111): an RGB
    photo made from each 8-bit image (per-channel offsets and noise) through
    the grey formula of dense.rs:108 in f32, so the values are off the byte
    lattice and only the f32 kernels can take them."""
    import copy
    rng = np.random.default_rng(seed)
    out = []
    for v in views:
        g = v.gray.astype(np.int16)
        rgb = [np.clip(g + rng.integers(-12, 13, g.shape, dtype=np.int16) + off, 0,
                       255).astype(np.float32) for off in (6, -4, 9)]
        grey = ((np.float32(0.299) * rgb[0] + np.float32(0.587) * rgb[1])
                + np.float32(0.114) * rgb[2]) / np.float32(255.0)
        v2 = copy.copy(v)
        v2.gray_f32 = grey
        out.append(v2)
    return out


def f32_input_leg(ctx, surv, timed, prm):
    """The batched refined path on f32 grey input (MVS_GRAY_F32, the `.view`
    cache's own grey): same views, neighbours and parameters as the headline
    run; one untimed batch of the first four views tunes the schedules (as
    the headline run's warm-up does: tuned on a single view, the schedules
    suit a one-view launch rather than the batch)."""
    from aerial_mvs import capi
    t = time.perf_counter()
    views = reference_grey(surv.views)
    sv = capi.Survey(ctx, views)
    ctx.synchronize()
    setup = time.perf_counter() - t
    sv.depth_maps(timed[:4], prm)
    ctx.synchronize()
    ctx.reset_stats()
    t0 = time.perf_counter()
    sv.depth_maps(timed, prm)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    st = ctx.stats()
    pixels = sum(surv.views[j[0]].w * surv.views[j[0]].h for j in timed)
    errs = [depth_errors(sv.download(i)[0], surv.views[i].depth_gt) for (i, _, _) in timed]
    sv.close()
    out = dict(value=pixels / dt / 1e6, unit="depth Mpix/s", views=len(timed), seconds=dt,
               split=st["split"], sweep_ms=st["sweep_ms"], refine_ms=st["refine_ms"],
               input="f32 grey from RGB by dense.rs:108 (off the byte lattice), uploaded "
                     f"as MVS_GRAY_F32 ({setup:.1f} s to make and upload)",
               path="batched GPU, fast arithmetic, one lane per pixel on the f32 "
                    "corner-difference plane or the grey plane (per iteration, same costs), "
                    "refinement one lane per pixel",
               schedule=[dict(iteration=t["iteration"], split=t["split"], gpc=t["gpc"],
                              per_cu=t["per_cu"])
                         for t in ctx.tuned_schedules() if t["mode"] == 2],
               depth_error=mean_errors(errs))
    log(f"f32 input: {out['value']:.2f} depth Mpix/s over {len(timed)} views")
    return out


def run_uneven(args):
    """The batched uneven workload: every view of an uneven survey (nadir and
    oblique views of two sizes, per-view neighbour count 3..12 and window
    5x5/7x7/11x11) in one survey call; one untimed pass tunes the schedules,
    the second is timed."""
    from aerial_mvs import capi, survey
    spec = survey.UNEVEN[args.config]
    t = time.perf_counter()
    surv, per = survey.make_uneven_survey(spec, nthreads=args.threads)
    log(f"uneven survey {spec.name}: {len(surv.views)} views rendered in "
        f"{time.perf_counter() - t:.1f} s")
    n = len(surv.views) if args.steps <= 0 else min(args.steps, len(surv.views))
    jobs, prms = [], []
    for i in range(n):
        p = per[i]
        jobs.append((i, list(surv.neighbours[i][:p["k"]]), p["range"]))
        prms.append(capi.params(taps=p["taps"], stride=p["stride"], iterations=spec.iterations,
                                refine_steps=args.refine_steps, fast=args.mode == "fast"))
    ctx = capi.Context(0)
    sv = capi.Survey(ctx, surv.views)
    sv.depth_maps(jobs, prms)  # warm-up: one untimed pass tunes the schedules
    ctx.synchronize()
    ctx.reset_stats()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    sv.depth_maps(jobs, prms)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    st = ctx.stats()
    pixels = sum(surv.views[j[0]].w * surv.views[j[0]].h for j in jobs)
    errs = []
    for (i, _, _) in jobs:
        d, _, _ = sv.download(i)
        errs.append(depth_errors(d, surv.views[i].depth_gt))
    kinds = {}
    for p in per[:n]:
        key = f"{p['kind']} {p['taps']}x{p['taps']}"
        kinds[key] = kinds.get(key, 0) + 1
    line = dict(
        metric="depth Mpix/s, batched uneven survey (one call, per-view K and window)",
        value=pixels / elapsed / 1e6, unit="depth Mpix/s", n_gpus=1, steps=n,
        warmup=n, ms_per_step=elapsed * 1e3 / n, higher_is_better=True,
        scaling="weak", vs_baseline=None, dtype="f32", data="synthetic",
        config=dict(workload=(f"{spec.nadir_lines * spec.nadir_per_line} nadir "
                              f"{spec.nadir_size[0]}x{spec.nadir_size[1]} + {spec.oblique} "
                              f"oblique {spec.oblique_size[0]}x{spec.oblique_size[1]} u8 views, "
                              f"K in {spec.k_range}, windows {spec.windows}, "
                              f"{spec.iterations} iterations"),
                    views=n, mode=args.mode, refine_steps=args.refine_steps, classes=kinds),
        stats=dict(sweep_ms=st["sweep_ms"], sweep_launches=st["sweep_launches"],
                   sweep_kernels=st["sweep_kernels"],
                   init_ms=st["init_ms"], refine_ms=st["refine_ms"], evals=st["evals"],
                   graph_replays=st["graph_replays"]),
        depth_error=mean_errors(errs),
    )
    print(json.dumps(line), flush=True)
    sv.close()
    ctx.close()


def run_large(args):
    """The large survey (c5): 2000 views 2400x1800 (1400 nadir, 600 oblique)
    rendered on the GPU, all resident in HBM as lean views (f32 grey, 8-bit
    and f16 corner planes: 16 B per pixel) with depth-only maps; a few views
    of each kind tune the schedules, then every view's depth map (fast
    arithmetic, refinement) and the reprojection consistency filter are
    timed in one survey call, which batches the views by memory."""
    from aerial_mvs import capi, survey
    if args.mode != "fast":
        print("the large survey runs in fast mode only: its lean views keep no f32 plane",
              file=sys.stderr)
        sys.exit(2)
    spec = survey.LARGE[args.config]
    t = time.perf_counter()
    surv, per = survey.make_large_survey(
        spec, progress=lambda k, n: log(f"rendered {k}/{n} views"))
    n_views = len(surv.views)
    log(f"large survey {spec.name}: {n_views} views {spec.size[0]}x{spec.size[1]} rendered on "
        f"the GPU in {time.perf_counter() - t:.1f} s")
    centres = surv.centres()
    jobs = []
    for i in range(n_views):
        rg = per[i]["range"]
        nb = capi.select_neighbours(centres, i, surv.neighbours[i], rg[0], spec.neighbours)
        if nb:
            jobs.append((i, nb, rg))
    if args.steps > 0:
        jobs = jobs[:args.steps]
    prm = capi.params(taps=spec.taps, stride=spec.stride, iterations=spec.iterations,
                      refine_steps=args.refine_steps, fast=True)
    ctx = capi.Context(0)
    t = time.perf_counter()
    sv = capi.Survey(ctx, surv.views,
                     options=capi.MVS_SURVEY_LEAN_VIEWS | capi.MVS_SURVEY_DEPTH_ONLY)
    ctx.synchronize()
    gpix = sum(v.w * v.h for v in surv.views) / 1e9
    log(f"{n_views} views resident in HBM ({gpix:.2f} Gpix, {16 * gpix:.0f} GB of planes) in "
        f"{time.perf_counter() - t:.1f} s")
    # tune the schedules on a few views of each kind (untimed)
    warm = ([j for j in jobs if per[j[0]]["kind"] == "nadir"][:8] +
            [j for j in jobs if per[j[0]]["kind"] == "oblique"][:8])
    sv.depth_maps(warm, prm)
    ctx.synchronize()
    ctx.reset_stats()
    ctx.set_timing(True)
    done_views = [0]

    def progress(user, stage, k, n):
        if k - done_views[0] >= 100 or k == n:
            done_views[0] = k
            log(f"  depth maps {k}/{n} ({time.perf_counter() - t0:.1f} s)")
        return 0

    done = set(j[0] for j in jobs)
    t0 = time.perf_counter()
    sv.depth_maps(jobs, prm, progress=progress)
    n_cons = 0
    for (i, _, _) in jobs:
        nbrs = [j for j in surv.neighbours[i][:spec.neighbours + 1] if j in done]
        if nbrs:
            n_cons += sv.consistency(i, nbrs, rule=capi.MVS_CONS_REPROJ, download=False)[1]
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    st = ctx.stats()
    pixels = sum(surv.views[j[0]].w * surv.views[j[0]].h for j in jobs)
    value = pixels / elapsed / 1e6
    log(f"timed {len(jobs)} views: {elapsed:.1f} s, {value:.2f} depth Mpix/s; stats {st}")
    errs = {"nadir": [], "oblique": []}
    errs_c = {"nadir": [], "oblique": []}
    for (i, _, _) in jobs:
        if surv.views[i].depth_gt is not None:
            d, _, _ = sv.download(i)
            errs[per[i]["kind"]].append(depth_errors(d, surv.views[i].depth_gt))
            nbrs = [j for j in surv.neighbours[i][:spec.neighbours + 1] if j in done]
            if nbrs:
                agree, _ = sv.consistency(i, nbrs, rule=capi.MVS_CONS_REPROJ)
                errs_c[per[i]["kind"]].append(depth_errors(np.where(agree >= 3, d, 0.0),
                                                           surv.views[i].depth_gt))
    err = {k: mean_errors(v) for k, v in errs.items() if v}
    for k, v in errs_c.items():
        if v:
            err[k + "_consistent"] = mean_errors(v)
    err["consistency_rule"] = "reprojection <= 1 px in >= 3 of the K+1 nearest views"
    err["views_checked"] = {k: len(v) for k, v in errs.items()}
    err["reference"] = "FP64 ray-cast ground truth (stored as f32) of every 25th view"
    err["consistent_pixels"] = int(n_cons)
    # roofline of the half-sweeps: flops per evaluation at the pixel-weighted mean K
    kmean = sum(len(j[1]) * surv.views[j[0]].w * surv.views[j[0]].h for j in jobs) / pixels
    fpe = spec.taps * spec.taps * (12 + 48 * kmean)
    avg_ms = st["sweep_ms"] / max(1, st["sweep_launches"])
    evals_per_launch = st["sweep_evals"] / max(1, st["sweep_launches"])
    achieved = fpe * evals_per_launch / (avg_ms * 1e-3) / 1e12
    roofline = dict(bound="mfma", achieved=achieved, peak=157.3, unit="TFLOP/s",
                    frac=achieved / 157.3, traffic=None, avg_launch_ms=avg_ms,
                    launches=st["sweep_launches"], flops_per_eval=fpe,
                    skipped_per_launch=st["sweep_skipped"] / max(1, st["sweep_launches"]),
                    limiter=LIMITER, bound_note=BOUND_NOTE)
    cpu = None
    if not args.no_cpu:
        cw, ch = (int(a) for a in args.cpu_crop.split("x"))
        log("CPU baseline (oracle) ...")
        cpu = cpu_baseline(surv, jobs, spec, args.threads, cw, ch, ctx,
                           capi.params(taps=spec.taps, stride=spec.stride,
                                       iterations=spec.iterations),
                           capi.params(taps=spec.taps, stride=spec.stride,
                                       iterations=spec.iterations, fast=True))
        log(f"CPU {cpu['value']:.4f} Mpix/s on {cpu['cores']} threads")
    n_nadir = sum(1 for j in jobs if per[j[0]]["kind"] == "nadir")
    line = dict(
        metric=metric_for(spec.taps, spec.neighbours), value=value, unit="depth Mpix/s",
        n_gpus=1, steps=len(jobs),
        warmup=len(warm), ms_per_step=elapsed * 1e3 / len(jobs), higher_is_better=True,
        scaling="weak", vs_baseline=None, dtype="f32", data="synthetic",
        config=dict(workload=(f"{n_views} views {spec.size[0]}x{spec.size[1]} u8 grey "
                              f"({n_nadir} nadir, {len(jobs) - n_nadir} oblique timed), "
                              f"{spec.block[0] / 1000:g}x{spec.block[1] / 1000:g} km fractal "
                              f"heightfield with {spec.boxes} boxes 3-30 m, "
                              f"{spec.neighbours} neighbours, {spec.taps}x{spec.taps} NCC, "
                              f"{spec.iterations} iterations, {args.refine_steps} refinement "
                              "steps, reprojection consistency filter"),
                    views=n_views, resident_gpix=gpix, layout="lean views + depth-only maps",
                    mean_k=kmean,
                    schedule=[dict(iteration=t_["iteration"], mode=t_["mode"], split=t_["split"],
                                   gpc=t_["gpc"], per_cu=t_["per_cu"])
                              for t_ in ctx.tuned_schedules()]),
        roofline=roofline, cpu_baseline=cpu, depth_error=err,
        gpu_vs_cpu=(value / cpu["value"]) if cpu else None,
        stats=dict(sweep_ms=st["sweep_ms"], sweep_launches=st["sweep_launches"],
                   init_ms=st["init_ms"], refine_ms=st["refine_ms"], evals=st["evals"],
                   graph_replays=st["graph_replays"]),
    )
    print(json.dumps(line), flush=True)
    sv.close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="reference views timed (default 0: every view of the survey)")
    ap.add_argument("--warmup", type=int, default=4,
                    help="untimed warm-up views, run as one batch (it tunes the schedules)")
    ap.add_argument("--config", default="c2")
    ap.add_argument("--mode", choices=("fast", "exact"), default="fast",
                    help="arithmetic of the batched path's PatchMatch kernels")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fusion", action="store_true")
    ap.add_argument("--no-dsm", action="store_true")
    ap.add_argument("--no-f32", action="store_true",
                    help="skip the leg on f32 grey input (the `.view` cache's format)")
    ap.add_argument("--grad-items", type=int, default=-1,
                    help="pixels of the full-size gradient check in the CPU leg "
                         "(default 10000 for c4, 2000 otherwise; 0 skips it)")
    ap.add_argument("--no-prep", action="store_true",
                    help="skip the view-preparation stage (RGB photo -> .view contents)")
    ap.add_argument("--cpu-crop", default="1280x960")
    ap.add_argument("--threads", type=int, default=default_threads(),
                    help="CPU threads of the oracle leg (default: the CPUs this process may "
                         "run on, capped by OMP_NUM_THREADS)")
    ap.add_argument("--plain-views", type=int, default=8)
    ap.add_argument("--refine-steps", type=int, default=5)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy of this many MiB, 5 times (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this depth engine runs on a single MI355X "
              "(multi-GPU sharding by reference view is designed but not built); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        n = args.copy_probe << 18  # float32 elements
        a = torch.ones(n, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {args.copy_probe} MiB 5 times")
        return

    from aerial_mvs import capi, survey

    if args.config in survey.UNEVEN:
        return run_uneven(args)
    if args.config in survey.LARGE:
        return run_large(args)
    spec = survey.SPECS[args.config]
    t = time.perf_counter()
    surv = survey.make_survey(spec, nthreads=args.threads)
    log(f"survey {spec.name}: {len(surv.views)} views {spec.width}x{spec.height} "
        f"rendered in {time.perf_counter() - t:.1f} s")
    centres = surv.centres()
    jobs = []
    for i in range(len(surv.views)):
        nb = capi.select_neighbours(centres, i, surv.neighbours[i], spec.depth_range[0],
                                    spec.neighbours)
        jobs.append((i, nb, spec.depth_range))
    steps = len(jobs) if args.steps <= 0 else max(1, min(args.steps, len(jobs)))
    geom = dict(taps=spec.taps, stride=spec.stride, iterations=spec.iterations)
    prm = capi.params(**geom, refine_steps=args.refine_steps, fast=args.mode == "fast")
    prm_exact = capi.params(**geom)                    # the reference's algorithm
    prm_fast0 = capi.params(**geom, fast=True)         # fast, no refinement

    ctx = capi.Context(0)
    tu = time.perf_counter()
    sv = capi.Survey(ctx, surv.views)
    ctx.synchronize()
    upload_s = time.perf_counter() - tu
    log(f"views resident in HBM (upload and planes: {upload_s:.2f} s)")
    # warm-up: the first W views as one batch (the autotuner times its
    # candidate schedules on up to four of them, per iteration), and one view
    # of the plain path
    if args.warmup > 0:
        sv.depth_maps(jobs[:min(args.warmup, len(jobs))], prm)
        j = jobs[0]
        sv.depth_map(j[0], j[1], j[2][0], j[2][1], prm_exact)
    ctx.synchronize()
    ctx.reset_stats()
    ctx.set_timing(True)
    timed = jobs[:steps]
    done = set(j[0] for j in timed)

    # plain path: one reference view per call, exact kernels, no refinement
    plain_n = max(0, min(args.plain_views, len(jobs)))
    plain = None
    plain_maps = {}
    if plain_n:
        ctx.synchronize()
        tp = time.perf_counter()
        for (i, nb, rng) in jobs[:plain_n]:
            sv.depth_map(i, nb, rng[0], rng[1], prm_exact)
        ctx.synchronize()
        dtp = time.perf_counter() - tp
        plain = dict(value=sum(surv.views[j[0]].w * surv.views[j[0]].h
                               for j in jobs[:plain_n]) / dtp / 1e6,
                     unit="depth Mpix/s", views=plain_n,
                     path="one view per call, exact kernels (bit-identical to the oracle), "
                          "no refinement")
        plain_maps = {i: sv.download(i)[0] for (i, _, _) in jobs[:plain_n]}
        plain["depth_error"] = mean_errors([depth_errors(plain_maps[i], surv.views[i].depth_gt)
                                            for i in plain_maps])
        log(f"plain GPU path: {plain['value']:.2f} depth Mpix/s over {plain_n} views")
    ctx.reset_stats()

    ctx.synchronize()
    t0 = time.perf_counter()
    sv.depth_maps(timed, prm)
    n_cons = 0
    cons_nbrs = {}
    for (i, _, _) in timed:
        nbrs = [j for j in surv.neighbours[i][:spec.neighbours + 1] if j in done]
        cons_nbrs[i] = nbrs
        if nbrs:
            n_cons += sv.consistency(i, nbrs, rule=capi.MVS_CONS_REPROJ, download=False)[1]
    ctx.synchronize()
    t1 = time.perf_counter()

    st = ctx.stats()
    pixels = sum(surv.views[j[0]].w * surv.views[j[0]].h for j in timed)
    elapsed = t1 - t0
    value = pixels / elapsed / 1e6
    log(f"timed {steps} views: {elapsed:.2f} s, {value:.2f} depth Mpix/s; stats {st}")

    # roofline of the dominant kernel (the half-sweep)
    fpe = flops_per_eval(spec.taps, spec.neighbours)
    avg_ms = st["sweep_ms"] / max(1, st["sweep_launches"])
    evals_per_launch = st["sweep_evals"] / max(1, st["sweep_launches"])
    achieved = fpe * evals_per_launch / (avg_ms * 1e-3) / 1e12
    # bilinear samples gathered per evaluation: taps x sources
    samples_per_eval = spec.taps * spec.taps * spec.neighbours
    rd, wr = BYTES_PER_EVAL[args.mode]
    traffic = (rd + wr) * evals_per_launch if rd is not None else None
    sched = [t for t in ctx.tuned_schedules()
             if t["mode"] == (1 if args.mode == "fast" else 0) and t["taps"] == spec.taps]
    splits = sorted({t["split"] for t in sched})
    if args.mode == "fast" and splits and max(splits) < 0:
        kname = (f"k_pml<{spec.taps},sweep,persistent> (one lane per pixel; corner plane per "
                 "iteration: " + ", ".join(f"{t['iteration']}:{'f16' if t['split'] == -1 else '8-bit'}"
                                          for t in sorted(sched, key=lambda t: t['iteration']))
                 + ")")
    else:
        kname = (f"k_pm<{spec.taps},8,split {'/'.join(map(str, splits))},pinhole,sweep,persistent,"
                 f"{'fast' if args.mode == 'fast' else 'exact'}>")
    roofline = dict(bound="mfma", achieved=achieved, peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                    frac=achieved / PEAK_FP32_TFLOPS, traffic=traffic,
                    limiter=LIMITER, bound_note=BOUND_NOTE,
                    traffic_note=("HBM-side bytes per launch: rocprofv3 FETCH_SIZE/WRITE_SIZE per "
                                  "evaluation (copy-calibrated) x evaluations per launch"),
                    hbm_gbs=(traffic / (avg_ms * 1e-3) / 1e9) if traffic else None,
                    kernel=kname,
                    avg_launch_ms=avg_ms, evals_per_launch=evals_per_launch, flops_per_eval=fpe,
                    skipped_per_launch=st["sweep_skipped"] / max(1, st["sweep_launches"]),
                    launches=st["sweep_launches"],
                    samples_per_s=samples_per_eval * evals_per_launch / (avg_ms * 1e-3),
                    samples_note=("bilinear samples gathered per second by the half-sweeps "
                                  "(taps x sources per evaluation), the quantity that bounds them"))

    # depth error against the FP64 ground truth, all timed views: every valid
    # pixel, and the pixels the consistency filter keeps
    errs, errs_c = [], []
    maps = {}
    for (i, _, _) in timed:
        d, _, _ = sv.download(i)
        maps[i] = d
        errs.append(depth_errors(d, surv.views[i].depth_gt))
        if cons_nbrs[i]:
            agree, _ = sv.consistency(i, cons_nbrs[i], rule=capi.MVS_CONS_REPROJ)
            errs_c.append(depth_errors(np.where(agree >= 3, d, 0.0), surv.views[i].depth_gt))
    err = mean_errors(errs)
    if errs_c:
        err["consistent"] = mean_errors(errs_c)
        # the refined-depth bar (within 0.2% of the FP64 truth on >= 99% of
        # valid pixels) is declared on the pixels the consistency filter
        # keeps: the set fusion takes into the cloud. Raw maps keep every
        # pixel under the cost threshold (dense.rs:468-477), including the
        # PatchMatch mismatches at occlusions and image borders that the
        # reference's algorithm makes too (the plain exact maps, identical to
        # the oracle, carry the same share).
        c = err["consistent"]
        err["refined_bar"] = dict(
            valid_set="pixels kept by the reprojection consistency filter (>= 3 of the K+1 "
                      "nearest views agree within 1 px)",
            valid_fraction=c["valid"], within_0p2=c["within_0p2"],
            bar="within 0.2% of the FP64 ground truth on >= 99% of valid pixels",
            passed=bool(c["within_0p2"] >= 0.99),
            raw_maps_within_0p2=err["within_0p2"], raw_maps_valid_fraction=err["valid"])
    err["consistent_pixels"] = int(n_cons)
    err["consistency_rule"] = "reprojection <= 1 px in >= 3 of the K+1 nearest views"
    err["reference"] = "FP64 ray-cast ground truth of the synthetic scene"
    common = [i for i in plain_maps if i in maps]
    if common:
        # the batched refined maps against the plain exact ones, same views
        err["vs_plain_exact_within_0p5"] = float(np.mean([within(maps[i], plain_maps[i])
                                                          for i in common]))

    # fusion of the timed views into the voxel cloud (dense.rs:542-651), timed
    # beside the depth metric: voxel edge from a sample of ground-truth points
    # standing in for the sparse cloud, Standard-profile point budget
    fusion = None
    if not args.no_fusion:
        pts = []
        for (i, _, _) in timed[::max(1, len(timed) // 16)]:
            v = surv.views[i]
            yy, xx = np.mgrid[0:v.h:97, 0:v.w:97]
            d = v.depth_gt[yy, xx]
            q = np.stack([(xx + 0.5 - v.cx) / v.f * d, (yy + 0.5 - v.cy) / v.f * d, d], -1)
            pts.append((q.reshape(-1, 3) @ np.asarray(v.R)) + v.C)
        pts = np.concatenate(pts)
        voxel = capi.fusion_voxel(pts[:, 0], pts[:, 1], [j[2] for j in jobs], spec.focal,
                                  max(spec.width, spec.height), spec.width, spec.height,
                                  5_000_000)
        refs = [i for (i, _, _) in timed]
        ctx.synchronize()
        tf = time.perf_counter()
        xyz, rgb = sv.fuse(refs, [cons_nbrs[i] for i in refs], voxel, rule=capi.MVS_CONS_REFERENCE)
        dtf = time.perf_counter() - tf
        fusion = dict(voxel_m=voxel, points=int(len(xyz)), seconds=dtf,
                      rule="dense.rs:585-612 (relative depth < 1% in >= 2 of the K+1 nearest)",
                      ply_bytes=int(len(xyz)) * 15)
        log(f"fusion: {len(xyz)} voxels of {voxel:.4f} m in {dtf:.2f} s")

    # the same batch on the reference's own input format (f32 grey)
    f32_leg = None
    if not args.no_f32 and args.mode == "fast":
        f32_leg = f32_input_leg(ctx, surv, timed, prm)

    # the surface model from the fused cloud (dsm.rs:128-290)
    dsm = dsm_got = None
    if fusion is not None and not args.no_dsm and len(xyz) >= 100:
        dsm, dsm_got = dsm_stage(ctx, xyz, voxel)

    prep = prep_photo = prep_out = None
    if not args.no_prep:
        prep, prep_photo, prep_out = view_prep_stage(ctx, spec)

    cpu = None
    if not args.no_cpu:
        cw, ch = (int(a) for a in args.cpu_crop.split("x"))
        log("CPU baseline (oracle) ...")
        cpu = cpu_baseline(surv, jobs, spec, args.threads, cw, ch, ctx, prm_exact, prm_fast0)
        log(f"CPU {cpu['value']:.4f} Mpix/s on {cpu['cores']} threads")
        if prep_out:
            cpu["view_prep"] = cpu_baseline_view_prep(prep_photo, prep_out)
            log(f"CPU view prep: {cpu['view_prep']}")
        if dsm_got:
            cpu["dsm"] = cpu_baseline_dsm(xyz, voxel, dsm_got)
            log(f"CPU dsm: {cpu['dsm']}")
        n_grad = args.grad_items if args.grad_items >= 0 else (10_000 if args.config == "c4"
                                                               else 2_000)
        if n_grad and args.refine_steps > 0:
            log(f"gradient check on {n_grad} pixels ...")
            cpu["gradient_check"] = gradient_check(ctx, surv, timed, spec, sv, n_grad)
            log(f"gradient check: {cpu['gradient_check']}")

    metric = (METRIC if (spec.taps, spec.neighbours) == (7, 8)
              else metric_for(spec.taps, spec.neighbours))
    line = dict(
        metric=metric, value=value, unit="depth Mpix/s", n_gpus=1, steps=steps,
        warmup=args.warmup, ms_per_step=elapsed * 1e3 / steps, higher_is_better=True,
        scaling="weak", vs_baseline=None, dtype="f32", data="synthetic",
        config=dict(workload=(f"{len(surv.views)} nadir views {spec.width}x{spec.height} u8 grey, "
                              f"{spec.neighbours} nearest-GPS neighbours, {spec.taps}x{spec.taps} "
                              f"NCC window, {spec.iterations} PatchMatch iterations, "
                              "geometric-consistency filter"),
                    views=len(surv.views), width=spec.width, height=spec.height,
                    neighbours=spec.neighbours, window=f"{spec.taps}x{spec.taps}/stride{spec.stride}",
                    iterations=spec.iterations, timed_views=steps,
                    refine_steps=args.refine_steps, mode=args.mode,
                    path=("batched GPU (one persistent launch per half-sweep for all views, "
                          f"cached graphs), {args.mode} arithmetic, "
                          f"{args.refine_steps} Gauss-Newton refinement steps"),
                    schedule=[dict(iteration=t["iteration"], mode=t["mode"], split=t["split"],
                                   gpc=t["gpc"], per_cu=t["per_cu"])
                              for t in ctx.tuned_schedules()]),
        roofline=roofline, cpu_baseline=cpu, depth_error=err, plain_gpu=plain,
        gpu_vs_cpu=(value / cpu["value"]) if cpu else None,
        evals_per_pixel=st["evals"] / max(1, pixels),
        refine=dict(ms=st["refine_ms"], launches=st["refine_launches"],
                    pixels=st["refine_pixels"], accepted_steps=st["refine_accepted"]),
        graphs=dict(captures=st["graph_captures"], replays=st["graph_replays"]),
        fusion=fusion, view_prep=prep, dsm=dsm, f32_input=f32_leg,
        end_to_end=dict(upload_s=upload_s,
                        upload_gb=sum(v.w * v.h for v in surv.views) / 1e9,
                        mpix_s_with_upload=pixels / (elapsed + upload_s) / 1e6,
                        note="host-to-device upload of all views (8-bit, pageable) and their "
                             "planes, before the timed region; not part of value"),
    )
    print(json.dumps(line), flush=True)
    sv.close()
    ctx.close()


if __name__ == "__main__":
    main()
