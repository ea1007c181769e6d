// C-ABI layer of the dense-MVS depth engine: contexts, device buffers, the
// resident survey, batching of reference views, schedule choice, and the host
// logic of the reference's depth stage (crates/stitch-core/src/recon/dense.rs)
// around the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/aerial_mvs.h"
#include "mvs_internal.h"

using namespace amvs;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= n) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            n = 0;
        }
        hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <class T>
    T *as() const {
        return (T *)p;
    }
};

// A device-resident view: grey plane plus camera.
struct ViewDev {
    int32_t w = 0, h = 0;
    DevBuf gray;
    DevBuf quad;   // corner-packed 8-bit plane (8-bit inputs)
    DevBuf quadh;  // corner-packed f16 plane with column differences (8-bit inputs)
    DevBuf quadf;  // corner-packed grey plane
    bool has_quad = false;
    DevBuf quadd;  // corner-difference plane (f32 grey inputs)
    bool has_quadd = false;
    DevBuf rgb;    // w*h*3 colour for the fused cloud (optional)
    bool has_rgb = false;
    double f = 0, cx = 0, cy = 0, k1 = 0, k2 = 0;
    double R[9] = {0}, C[3] = {0};
    bool set = false;
};

// Outputs of one reference view kept on the device.
struct MapDev {
    int32_t w = 0, h = 0;
    DevBuf depth, normal, cost;
    bool keep_normal = true, keep_cost = true;  // false: the buffer is not kept
    bool valid = false;
};

// Work buffers of one job slot of a batch.
struct JobWork {
    DevBuf hyp, cost, textured, changed, pair, list, counts, srcs;
};

// One reference view to process.
struct JobSpec {
    const ViewDev *ref = nullptr;
    std::vector<const ViewDev *> srcs;
    double range0 = 0, range1 = 0;
    mvs_params prm{};       // window, iterations, thresholds, refinement, mode
    MapDev *out = nullptr;  // thresholded outputs (nullable)
    bool depth_only = false;  // the output keeps the depth only
};

typedef std::tuple<double, double, double, double, double, int32_t, int32_t> RayKey;
// taps, source lanes, pinhole, log2 pixels, iteration, fast arithmetic
typedef std::tuple<int32_t, int32_t, int32_t, int32_t, int32_t, int32_t> TuneKey;

struct Schedule {
    int32_t split = 1;
    int32_t gpc = 8;
    int32_t per_cu = 16;  // workgroups per CU in the grid; 0: one per chunk
};

struct TimedLaunch {
    hipEvent_t a, b;
    int kind;  // 0 sweep, 1 init+eval, 2 refinement
};

}  // namespace

// ---- scratch of the allocation-free entry point ----
// mvs_survey_consistency_noalloc keeps its neighbour-camera table, the
// per-pixel agree counts and the consistent-pixel counter in this fixed
// device array instead of allocating: [counter | camera table | agree plane].
// Each GPU holds its own copy (at its own address) until the process ends.
constexpr size_t kNaHead = 256;
constexpr size_t kNaCams = (sizeof(amvs::NbrCam) * MVS_NOALLOC_MAX_NBRS + 255) / 256 * 256;
constexpr size_t kNaBytes = kNaHead + kNaCams + sizeof(int32_t) * (size_t)MVS_NOALLOC_MAX_PIXELS;
static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};

// This device's address of `scratch`, looked up once per device with the
// device current (the array's host-side name is not a device address).
hipError_t scratch_for_device(int32_t device, unsigned char **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = (unsigned char *)p;
    }
    *out = g_scratch[device];
    return hipSuccess;
}
}  // namespace

struct mvs_ctx {
    int32_t device = 0;
    int32_t num_cu = 256;
    hipStream_t st = nullptr;
    // this device's copy of the allocation-free entry point's scratch array
    // (hipGetSymbolAddress at creation; null with scratch_err set on failure)
    unsigned char *scratch = nullptr;
    std::string scratch_err;
    std::string err;
    std::vector<JobWork> work;  // job slots
    DevBuf jobs_dev, prefix, counters, evals, ok_count;
    DevBuf up_u8;                    // staging of u8 uploads
    std::vector<ViewDev> tmp_views;  // host-view calls
    MapDev tmp_map;
    DevBuf px, py, hyp_in, cost_out, nbr_cams, agree, save_hyp, save_cost, save_changed,
        save_pair;
    DevBuf prep_box, prep_rgb, prep_grey;  // view preparation
    DevBuf dsm_xyz, dsm_h, dsm_a, dsm_b, dsm_c, dsm_d, dsm_snap, dsm_parent, dsm_la, dsm_lb,
        dsm_ctr, dsm_sel;  // surface model
    int64_t dsm_cells = 0;  // cells of the last grid built (kept in dsm_h)
    DevBuf pairs, grad_io, gout, jtj_out;  // pair-cost lists
    std::map<RayKey, DevBuf *> rays;
    // host copies that must outlive the asynchronous uploads of a call
    std::vector<RefJob> host_jobs;
    std::vector<std::vector<SrcCam>> host_tables;
    // schedule: forced (split != 0) or tuned per class
    Schedule forced{0, 0, 0};
    bool autotune = true;
    std::map<TuneKey, Schedule> tuned;
    // half-sweep graphs, keyed by the bytes of their launches
    bool graphs = true;
    std::map<std::string, hipGraphExec_t> graph_cache;
    // timing
    bool timing = false;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> free_events;
    mvs_stats stats{};
};

struct mvs_survey {
    mvs_ctx *ctx = nullptr;
    int32_t options = 0;  // MVS_SURVEY_* flags
    std::vector<ViewDev> views;
    std::vector<MapDev> maps;
    // voxel table of the last fusion
    DevBuf fkeys, fsum, fcol, fcnt, fctr;
    int64_t fcap = 0;
    double fvoxel = 0.0;
    int64_t fvox = 0;          // voxels of the last fusion
    bool fpts_ready = false;   // its points are in fo_xyz / fo_rgb
    DevBuf fo_kin, fo_kout, fo_sin, fo_sout, fo_tmp, fo_xyz, fo_rgb, fo_ctr;
    bool fused = false;
};

namespace {

int fail(mvs_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(mvs_ctx *c, hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(c, MVS_ERR_DEVICE, m);
}

#define HIP_TRY(c, expr)                                     \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return hip_fail(c, _e, #expr); \
    } while (0)

#define TRY(expr)            \
    do {                     \
        int _rc = (expr);    \
        if (_rc) return _rc; \
    } while (0)

hipEvent_t take_event(mvs_ctx *c) {
    if (!c->free_events.empty()) {
        hipEvent_t e = c->free_events.back();
        c->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Resolve recorded kernel timings (after a stream synchronisation).
void collect_timing(mvs_ctx *c) {
    for (auto &t : c->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            if (t.kind == 0) {
                c->stats.sweep_ms += ms;
                c->stats.sweep_launches += 1;
            } else if (t.kind == 2) {
                c->stats.refine_ms += ms;
                c->stats.refine_launches += 1;
            } else {
                c->stats.init_ms += ms;
                c->stats.init_launches += 1;
            }
        }
        c->free_events.push_back(t.a);
        c->free_events.push_back(t.b);
    }
    c->pending.clear();
}

// Device counters (c->evals): [0] initial-cost evaluations, [1] half-sweep
// evaluations, [2] accepted refinement steps, [3] refined pixels, [4]
// half-sweep candidates skipped as scored and not taken before.
constexpr int kEvalSlots = 5;

int sync_ctx(mvs_ctx *c) {
    HIP_TRY(c, hipStreamSynchronize(c->st));
    collect_timing(c);
    unsigned long long ev[kEvalSlots] = {0, 0, 0, 0, 0};
    if (c->evals.p) {
        HIP_TRY(c, hipMemcpy(ev, c->evals.p, sizeof ev, hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemset(c->evals.p, 0, sizeof ev));
    }
    c->stats.init_evals += (int64_t)ev[0];
    c->stats.sweep_evals += (int64_t)ev[1];
    c->stats.evals += (int64_t)(ev[0] + ev[1]);
    c->stats.refine_accepted += (int64_t)ev[2];
    c->stats.refine_pixels += (int64_t)ev[3];
    c->stats.sweep_skipped += (int64_t)ev[4];
    c->host_jobs.clear();
    c->host_tables.clear();
    return MVS_OK;
}

int check_view(mvs_ctx *c, const mvs_view *v, const char *what) {
    if (!v) return fail(c, MVS_ERR_ARG, std::string(what) + ": null view");
    if (v->w < 2 || v->h < 2 || v->stride < v->w || !v->pixels)
        return fail(c, MVS_ERR_ARG, std::string(what) + ": bad size, stride or pixels");
    if ((int64_t)v->w * v->h >= (1ll << 31))
        return fail(c, MVS_ERR_ARG, std::string(what) + ": more than 2^31 pixels");
    if (v->format != MVS_GRAY_U8 && v->format != MVS_GRAY_F32)
        return fail(c, MVS_ERR_FORMAT, std::string(what) + ": unknown pixel format");
    if (!(v->f > 0.0)) return fail(c, MVS_ERR_ARG, std::string(what) + ": focal length <= 0");
    return MVS_OK;
}

int check_params(mvs_ctx *c, const mvs_params *p) {
    if (!p) return fail(c, MVS_ERR_ARG, "null params");
    if (!taps_supported(p->taps))
        return fail(c, MVS_ERR_ARG,
                    "window taps must be 5, 7 or 11 (got " + std::to_string(p->taps) + ")");
    if (p->stride < 1 || p->stride > 8) return fail(c, MVS_ERR_ARG, "tap stride must be 1..8");
    if (p->iterations < 0 || p->iterations > 64)
        return fail(c, MVS_ERR_ARG, "iterations out of range");
    if (p->refine_steps < 0 || p->refine_steps > 64)
        return fail(c, MVS_ERR_ARG, "refine steps out of range (0..64)");
    if (p->mode != MVS_MODE_EXACT && p->mode != MVS_MODE_FAST)
        return fail(c, MVS_ERR_ARG, "unknown arithmetic mode");
    return MVS_OK;
}

// Upload a host view into a ViewDev (grey converted to f32 on the device).
// lean: an 8-bit view gets no f32 corner plane.
int upload_view(mvs_ctx *c, ViewDev &dv, const mvs_view *v, bool lean = false) {
    TRY(check_view(c, v, "view"));
    const size_t n = (size_t)v->w * v->h;
    const size_t pad = (size_t)grey_pad(v->w);
    HIP_TRY(c, dv.gray.ensure((n + pad) * sizeof(float)));
    HIP_TRY(c, hipMemsetAsync(dv.gray.as<float>() + n, 0, pad * sizeof(float), c->st));
    if (v->format == MVS_GRAY_U8) {
        const size_t bytes = (size_t)v->stride * v->h;
        HIP_TRY(c, c->up_u8.ensure(bytes));
        HIP_TRY(c, hipMemcpyAsync(c->up_u8.p, v->pixels, bytes, hipMemcpyHostToDevice, c->st));
        launch_gray_from_u8(c->up_u8.as<uint8_t>(), v->w, v->h, v->stride, dv.gray.as<float>(),
                            c->st);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, dv.quad.ensure(n * sizeof(uint32_t)));
        launch_quad_from_u8(c->up_u8.as<uint8_t>(), v->w, v->h, v->stride,
                            dv.quad.as<uint32_t>(), c->st);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, dv.quadh.ensure(n * sizeof(uint2)));
        launch_quadh_from_u8(c->up_u8.as<uint8_t>(), v->w, v->h, v->stride,
                             dv.quadh.as<uint2>(), c->st);
        HIP_TRY(c, hipGetLastError());
        dv.has_quad = true;
        dv.quadd.release();
        dv.has_quadd = false;
    } else {
        dv.has_quad = false;
        HIP_TRY(c, hipMemcpy2DAsync(dv.gray.p, (size_t)v->w * sizeof(float), v->pixels,
                                    (size_t)v->stride * sizeof(float),
                                    (size_t)v->w * sizeof(float), v->h, hipMemcpyHostToDevice,
                                    c->st));
        // the plane the fast one-lane-per-pixel kernels sample on f32 views
        HIP_TRY(c, dv.quadd.ensure(n * sizeof(float4)));
        launch_quadd(dv.gray.as<float>(), v->w, v->h, dv.quadd.as<float4>(), c->st);
        HIP_TRY(c, hipGetLastError());
        dv.has_quadd = true;
    }
    if (lean && dv.has_quad) {
        dv.quadf.release();
    } else {
        HIP_TRY(c, dv.quadf.ensure(n * sizeof(float4)));
        launch_quadf(dv.gray.as<float>(), v->w, v->h, dv.quadf.as<float4>(), c->st);
        HIP_TRY(c, hipGetLastError());
    }
    // the staging buffer and the caller's pixels are reused after return
    HIP_TRY(c, hipStreamSynchronize(c->st));
    dv.w = v->w;
    dv.h = v->h;
    dv.f = v->f;
    dv.cx = v->cx;
    dv.cy = v->cy;
    dv.k1 = v->k1;
    dv.k2 = v->k2;
    memcpy(dv.R, v->R, sizeof dv.R);
    memcpy(dv.C, v->C, sizeof dv.C);
    dv.set = true;
    return MVS_OK;
}

// Ray table of a camera (dense.rs:343-349), cached per intrinsics and size.
int get_rays(mvs_ctx *c, const ViewDev &v, const float2 **out) {
    RayKey k(v.f, v.cx, v.cy, v.k1, v.k2, v.w, v.h);
    auto it = c->rays.find(k);
    if (it != c->rays.end()) {
        *out = it->second->as<float2>();
        return MVS_OK;
    }
    DevBuf *b = new DevBuf();
    hipError_t e = b->ensure((size_t)v.w * v.h * sizeof(float2));
    if (e != hipSuccess) {
        delete b;
        return hip_fail(c, e, "ray table");
    }
    launch_rays(v.f, v.cx, v.cy, v.k1, v.k2, v.w, v.h, b->as<float2>(), c->st);
    HIP_TRY(c, hipGetLastError());
    c->rays[k] = b;
    *out = b->as<float2>();
    return MVS_OK;
}

// dense.rs:299-309 — relative pose in f64, cast to f32.
void relative_pose(const double Rr[9], const double Cr[3], const double Rs[9], const double Cs[3],
                   float R[9], float t[3]) {
    const double dc[3] = {Cr[0] - Cs[0], Cr[1] - Cs[1], Cr[2] - Cs[2]};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            double acc = Rs[i * 3 + 0] * Rr[j * 3 + 0];
            acc = acc + Rs[i * 3 + 1] * Rr[j * 3 + 1];
            acc = acc + Rs[i * 3 + 2] * Rr[j * 3 + 2];
            R[i * 3 + j] = (float)acc;
        }
        double a = Rs[i * 3 + 0] * dc[0];
        a = a + Rs[i * 3 + 1] * dc[1];
        a = a + Rs[i * 3 + 2] * dc[2];
        t[i] = (float)a;
    }
}

void split_f64(double v, float &hi, float &lo) {
    hi = (float)v;
    lo = (float)(v - (double)hi);
}

SrcCam make_src(const ViewDev &ref, const ViewDev &s) {
    SrcCam c;
    memset(&c, 0, sizeof c);
    relative_pose(ref.R, ref.C, s.R, s.C, c.R, c.t);
    split_f64(s.f, c.f, c.f_lo);
    split_f64(s.cx, c.cx, c.cx_lo);
    split_f64(s.cy, c.cy, c.cy_lo);
    split_f64(s.k1, c.k1, c.k1_lo);
    split_f64(s.k2, c.k2, c.k2_lo);
    c.gray = s.gray.as<float>();
    c.quad = s.has_quad ? s.quad.as<uint32_t>() : nullptr;
    c.quadh = s.has_quad ? s.quadh.as<uint2>() : nullptr;
    c.quadf = s.quadf.as<float4>();
    c.quadd = s.has_quadd ? s.quadd.as<float4>() : nullptr;
    c.w = s.w;
    c.h = s.h;
    c.pinhole_exact = (s.k1 == 0.0 && s.k2 == 0.0 && c.f_lo == 0.0f && c.cx_lo == 0.0f &&
                       c.cy_lo == 0.0f)
                          ? 1
                          : 0;
    return c;
}

int ensure_map(mvs_ctx *c, MapDev &m, int32_t w, int32_t h, bool depth_only = false) {
    const size_t n = (size_t)w * h;
    HIP_TRY(c, m.depth.ensure(n * sizeof(float)));
    m.keep_normal = m.keep_cost = !depth_only;
    if (depth_only) {
        m.normal.release();
        m.cost.release();
    } else {
        HIP_TRY(c, m.normal.ensure(3 * n * sizeof(float)));
        HIP_TRY(c, m.cost.ensure(n * sizeof(float)));
    }
    m.w = w;
    m.h = h;
    return MVS_OK;
}

// A batch prepared on the device: jobs of equal kernel class are contiguous
// in the device job table.
struct KClass {
    int32_t srcl = 0, pinhole = 1, first = 0, count = 0;
    int32_t fast = 0;                     // MVS_MODE_FAST kernels
    int32_t taps = 7;                     // window taps per side
    int32_t steps = 0;                    // half-sweeps of its jobs
    int64_t cap_sweep = 0, cap_eval = 0;  // upper bounds of the list lengths
    std::vector<int64_t> caps;            // pixels of each job
    int32_t size = 0;                     // log2 pixels of the first job
    // the first half-sweep of each colour scores the initial planes itself
    // (one-lane-per-pixel kernels in iteration 0) instead of a separate pass
    bool own_init = false;
    Schedule sch;
};

int32_t size_bucket(int64_t px) {
    int32_t b = 0;
    while ((1ll << (b + 1)) <= px) ++b;
    return b;
}

PmArgs base_args(mvs_ctx *c, const KClass &k) {
    PmArgs a;
    memset(&a, 0, sizeof a);
    a.jobs = c->jobs_dev.as<RefJob>() + k.first;
    a.njobs = k.count;
    a.prefix = c->prefix.as<int32_t>();
    a.counters = c->counters.as<int32_t>();
    a.gpc = k.sch.gpc;
    a.range = 1.0f;
    return a;
}

// Grid of a launch. Persistent: resident workgroups, at most one per chunk.
// Static: one workgroup per position of the longest possible XCD sequence
// (each job's chunks split into 8 bands), times 8. items_of(j) bounds the
// items of job j in this launch.
template <class F>
int32_t grid_for(mvs_ctx *c, const KClass &k, bool persist, F items_of) {
    const int64_t pix = k.sch.split < 0 ? 64 : 64 / (k.srcl * k.sch.split),
                  chunk = pix * k.sch.gpc;
    int64_t chunks = 0, seq = 0;
    for (int32_t j = 0; j < k.count; ++j) {
        const int64_t n = (items_of(j) + chunk - 1) / chunk;
        chunks += n;
        seq += (n + 7) / 8;
    }
    const int64_t g = persist ? std::min<int64_t>(chunks, (int64_t)c->num_cu * k.sch.per_cu)
                              : 8 * seq;
    return (int32_t)std::max<int64_t>(1, std::min<int64_t>(g, (1ll << 31) - 1));
}

int pm_launch(mvs_ctx *c, int mode, int32_t taps, const KClass &k, const PmArgs &a,
              int64_t n_items = 0) {
    const bool persist = mode == PM_SWEEP && k.sch.per_cu > 0;
    int32_t grid;
    if (mode == PM_COST_LIST)
        grid = grid_for(c, k, false, [&](int32_t) { return n_items; });
    else if (mode == PM_SWEEP)
        grid = grid_for(c, k, persist, [&](int32_t j) { return (k.caps[(size_t)j] + 1) / 2; });
    else
        grid = grid_for(c, k, persist, [&](int32_t j) { return k.caps[(size_t)j]; });
    if (launch_pm(mode, taps, k.srcl, k.sch.split, k.pinhole, persist, k.fast != 0, a, grid,
                  c->st))
        return fail(c, MVS_ERR_ARG, "kernel class not instantiated");
    HIP_TRY(c, hipGetLastError());
    return MVS_OK;
}

// Candidate schedules of a class: tap split x chunk size, and for the fast
// arithmetic on undistorted cameras the one-lane-per-pixel kernels (split -1
// on the f16 corner plane, -2 on the 8-bit one). A class without source lanes
// (srcl 0: jobs of mixed neighbour counts) has only the latter, and so has a
// half-sweep that scores the initial planes itself (pixel_lanes).
std::vector<Schedule> candidates(const KClass &k, bool pixel_lanes = false) {
    if (k.fast == 2) {
        // f32 views: only the one-lane-per-pixel kernels, on the f32
        // corner-difference plane or the grey plane (identical values)
        return {Schedule{-3, 1, 16}, Schedule{-3, 1, 8}, Schedule{-4, 1, 16},
                Schedule{-4, 1, 8}};
    }
    std::vector<Schedule> cand;
    if (k.srcl > 0 && !pixel_lanes) {
        int32_t sp[4];
        const int ns = splits_for(k.taps, sp);
        for (int i = 0; i < ns; ++i) {
            if (k.srcl * sp[i] > 64) continue;
            cand.push_back(Schedule{sp[i], 1, 0});   // static grid
            cand.push_back(Schedule{sp[i], 1, 16});  // persistent
            cand.push_back(Schedule{sp[i], 4, 16});
        }
    }
    if (k.fast && k.pinhole == 2) {
        cand.push_back(Schedule{-1, 1, 16});  // f16 corner plane
        cand.push_back(Schedule{-1, 1, 8});
        cand.push_back(Schedule{-5, 1, 16});  // f16 corner plane staged in LDS tiles
        cand.push_back(Schedule{-5, 1, 8});
        cand.push_back(Schedule{-2, 1, 16});  // 8-bit corner plane
        cand.push_back(Schedule{-2, 1, 8});
    }
    return cand;
}

// Jobs of a class the autotuner times its candidates on, and the runs per
// candidate (the fastest counts): a few views keep the launch in the
// many-chunks regime of a batch, and a second run discounts a first run's
// cold caches.
constexpr int32_t kTuneJobs = 4, kTuneRuns = 2;

// Time the half-sweep (it, parity) of the first kTuneJobs jobs of a class per
// candidate schedule on copies of their states, keep the fastest. With the
// exact arithmetic all schedules give bit-identical results; with the fast
// arithmetic they sum in different orders, so costs differ within the fast
// tolerance. Speed differs between the first iterations (random planes,
// scattered gathers) and the converged ones, so the choice is made per
// iteration.

int autotune_class(mvs_ctx *c, int32_t taps, KClass &k, const RefJob *j0, int32_t it,
                   int32_t parity, bool skip_unchanged, bool eval_own) {
    const std::vector<Schedule> cand = candidates(k, eval_own);
    if (cand.empty()) return fail(c, MVS_ERR_ARG, "no schedule for this class");
    if (cand.size() == 1) {
        k.sch = cand[0];
        return MVS_OK;
    }
    const int32_t nt = std::min(k.count, kTuneJobs);
    std::vector<size_t> off((size_t)nt + 1, 0);
    for (int32_t j = 0; j < nt; ++j) off[(size_t)j + 1] = off[(size_t)j] + (size_t)j0[j].w * j0[j].h;
    const size_t n = off[(size_t)nt];
    HIP_TRY(c, c->save_hyp.ensure(n * sizeof(float4)));
    HIP_TRY(c, c->save_cost.ensure(n * sizeof(float)));
    HIP_TRY(c, c->save_changed.ensure(n));
    HIP_TRY(c, c->save_pair.ensure(n));
    // the tuning jobs' states (hypotheses, costs, change flags, pairs), saved
    // before and restored before every run
    auto copy_states = [&](bool save) -> int {
        for (int32_t j = 0; j < nt; ++j) {
            const size_t m = off[(size_t)j + 1] - off[(size_t)j];
            void *sv[4] = {c->save_hyp.as<float4>() + off[(size_t)j],
                           c->save_cost.as<float>() + off[(size_t)j],
                           c->save_changed.as<uint8_t>() + off[(size_t)j],
                           c->save_pair.as<uint8_t>() + off[(size_t)j]};
            void *jv[4] = {j0[j].hyp, j0[j].cost, j0[j].changed, j0[j].pair};
            const size_t bytes[4] = {m * sizeof(float4), m * sizeof(float), m, m};
            for (int q = 0; q < 4; ++q)
                HIP_TRY(c, hipMemcpyAsync(save ? sv[q] : jv[q], save ? jv[q] : sv[q], bytes[q],
                                          hipMemcpyDeviceToDevice, c->st));
        }
        return MVS_OK;
    };
    TRY(copy_states(true));
    hipEvent_t ea = take_event(c), eb = take_event(c);
    float best_ms = 1e30f;
    Schedule best = cand[0];
    KClass one = k;
    one.count = nt;
    one.caps.assign(k.caps.begin(), k.caps.begin() + nt);
    for (const Schedule &s : cand) {
        one.sch = s;
        PmArgs a = base_args(c, one);
        a.it = it;
        a.parity = parity;
        a.range = ldexpf(1.0f, -it);
        a.skip_unchanged = skip_unchanged ? 1 : 0;
        a.eval_own = eval_own ? 1 : 0;
        for (int32_t r = 0; r < kTuneRuns; ++r) {
            TRY(copy_states(false));
            HIP_TRY(c, hipEventRecord(ea, c->st));
            TRY(pm_launch(c, PM_SWEEP, taps, one, a));
            HIP_TRY(c, hipEventRecord(eb, c->st));
            HIP_TRY(c, hipEventSynchronize(eb));
            float ms = 0.0f;
            HIP_TRY(c, hipEventElapsedTime(&ms, ea, eb));
            if (ms < best_ms) {
                best_ms = ms;
                best = s;
            }
        }
    }
    TRY(copy_states(false));
    c->free_events.push_back(ea);
    c->free_events.push_back(eb);
    k.sch = best;
    return MVS_OK;
}

// Schedule of the initial-cost evaluation (one cost per pixel, scattered
// random planes): the one-lane-per-pixel kernels where they apply, whose
// gathers each read one source image. 8-bit views: the 8-bit corner plane,
// one 4-byte gather per sample (105 ms on 77 views 2000x1500, 143 ms on the
// f16 plane); f32 views: the f32 coefficient plane, one 16-byte gather per
// sample (140 ms; 214 ms on the grey plane's two 8-byte row gathers). Else the
// first source-lane candidate.
Schedule eval_schedule(const KClass &k) {
    if (k.fast == 2) return Schedule{-3, 1, 0};
    if (k.fast && k.pinhole == 2) return Schedule{-2, 1, 0};
    return candidates(k)[0];
}

Schedule forced_schedule(mvs_ctx *c, const KClass &k) {
    const std::vector<Schedule> cand = candidates(k);
    Schedule s = c->forced;
    if (s.gpc <= 0) s.gpc = 8;
    bool ok = false;
    for (const Schedule &x : cand) ok = ok || x.split == s.split;
    if (!ok) s.split = cand.empty() ? 1 : cand[0].split;
    return s;
}

// Enqueue PatchMatch for a batch of reference views. first_step = -1 runs the
// initialisation (dense.rs:365-402) then n_steps half-sweeps from 0;
// first_step >= 0 keeps the state already in the job slots and runs half-sweeps
// first_step .. first_step + n_steps - 1.
// One half-sweep of every class of a batch, launched directly or as a cached
// hipGraph. A graph is keyed by the bytes of its launches (arguments, grids,
// kernel selectors), so a replay repeats exactly what a capture recorded; the
// job tables it reads are device memory rewritten by each call. The kernels
// clear their own counters (no memset nodes).
struct SweepLaunch {
    PmArgs a;
    int32_t grid, srcl, split, pinhole, persist, taps, fast;
};

int sweep_launches(mvs_ctx *c, const std::vector<const KClass *> &classes, int32_t it,
                   int32_t parity, bool skip_unchanged, bool first_pair,
                   unsigned long long *evals, unsigned long long *skipped,
                   std::vector<SweepLaunch> &out) {
    out.clear();
    for (const KClass *kp : classes) {
        const KClass &k = *kp;
        SweepLaunch L;
        memset(&L, 0, sizeof L);
        L.a = base_args(c, k);
        L.a.it = it;
        L.a.parity = parity;
        L.a.range = ldexpf(1.0f, -it);  // 0.5f32.powi(it), dense.rs:414
        L.a.evals = evals;
        L.a.skip_unchanged = skip_unchanged ? 1 : 0;
        L.a.skipped = skipped;
        L.a.eval_own = k.own_init && first_pair ? 1 : 0;
        L.persist = k.sch.per_cu > 0;
        L.grid = grid_for(c, k, L.persist != 0,
                          [&](int32_t j) { return (k.caps[(size_t)j] + 1) / 2; });
        L.srcl = k.srcl;
        L.split = k.sch.split;
        L.pinhole = k.pinhole;
        L.taps = k.taps;
        L.fast = k.fast;
        out.push_back(L);
    }
    return MVS_OK;
}

int enqueue_sweeps(mvs_ctx *c, const std::vector<SweepLaunch> &ls) {
    for (const SweepLaunch &L : ls) {
        if (launch_pm(PM_SWEEP, L.taps, L.srcl, L.split, L.pinhole, L.persist != 0,
                      L.fast != 0, L.a,
                      L.grid, c->st))
            return fail(c, MVS_ERR_ARG, "kernel class not instantiated");
        HIP_TRY(c, hipGetLastError());
    }
    return MVS_OK;
}

int run_sweeps_graph(mvs_ctx *c, const std::vector<SweepLaunch> &ls) {
    std::string key;
    auto put = [&key](const void *p, size_t n) { key.append((const char *)p, n); };
    for (const SweepLaunch &L : ls) {
        put(&L.a.jobs, sizeof L.a.jobs);
        put(&L.a.njobs, sizeof L.a.njobs);
        put(&L.a.prefix, sizeof L.a.prefix);
        put(&L.a.counters, sizeof L.a.counters);
        put(&L.a.it, sizeof L.a.it);
        put(&L.a.parity, sizeof L.a.parity);
        put(&L.a.range, sizeof L.a.range);
        put(&L.a.gpc, sizeof L.a.gpc);
        put(&L.a.evals, sizeof L.a.evals);
        put(&L.a.skip_unchanged, sizeof L.a.skip_unchanged);
        put(&L.a.skipped, sizeof L.a.skipped);
        put(&L.a.eval_own, sizeof L.a.eval_own);
        const int32_t sel[7] = {L.grid, L.srcl, L.split, L.pinhole, L.persist, L.taps, L.fast};
        put(sel, sizeof sel);
    }
    auto f = c->graph_cache.find(key);
    if (f == c->graph_cache.end()) {
        if (c->graph_cache.size() >= 512) {
            for (auto &kv : c->graph_cache) (void)hipGraphExecDestroy(kv.second);
            c->graph_cache.clear();
        }
        HIP_TRY(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_sweeps(c, ls);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        HIP_TRY(c, e);
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_TRY(c, ei);
        f = c->graph_cache.emplace(key, x).first;
        c->stats.graph_captures += 1;
    }
    HIP_TRY(c, hipGraphLaunch(f->second, c->st));
    c->stats.graph_replays += 1;
    return MVS_OK;
}

// Run PatchMatch for a batch of jobs, each with its own parameters. Jobs are
// grouped into kernel classes (window, half-sweep count, source lanes,
// projection, arithmetic); every half-sweep is one launch per active class.
// first_step < 0: initialise first (from step 0); otherwise continue from
// first_step on the jobs' current state. n_steps < 0: all 2 x iterations
// half-sweeps of each job, then refinement and thresholded outputs for jobs
// that have an output map.
// step_report (nullable) is called before each half-sweep is enqueued with
// (half-sweeps enqueued, half-sweeps in all) under stage "patchmatch", where
// cancel is polled; a nonzero return cancels.
struct StepReport {
    mvs_progress_fn fn;
    void *user;
};

int run_batch(mvs_ctx *c, const std::vector<JobSpec> &jobs, int32_t first_step,
              int32_t n_steps, const volatile int32_t *cancel,
              std::vector<KClass> *classes_out, const StepReport *step_report = nullptr) {
    const int32_t nj = (int32_t)jobs.size();
    if (nj == 0) return MVS_OK;
    for (const JobSpec &j : jobs) {
        const int32_t K = (int32_t)j.srcs.size();
        const int32_t half = (j.prm.taps / 2) * j.prm.stride;
        if (K < 1) return fail(c, MVS_NO_NEIGHBOURS, "no neighbours");
        if (lanes_for(K) < 0)
            return fail(c, MVS_ERR_ARG, "at most 16 source views per reference view");
        if (j.ref->w <= 2 * half || j.ref->h <= 2 * half)
            return fail(c, MVS_ERR_ARG, "image smaller than the window");
    }
    if ((int32_t)c->work.size() < nj) c->work.resize((size_t)nj);
    if (!c->evals.p) {
        HIP_TRY(c, c->evals.ensure(kEvalSlots * sizeof(unsigned long long)));
        HIP_TRY(c, hipMemset(c->evals.p, 0, kEvalSlots * sizeof(unsigned long long)));
    }
    HIP_TRY(c, c->prefix.ensure(sizeof(int32_t) * 9 * ((size_t)nj + 1)));
    HIP_TRY(c, c->counters.ensure(sizeof(int32_t) * 8));

    // host job descriptions and their class keys
    typedef std::tuple<int32_t, int32_t, int32_t, int32_t, int32_t> ClassKey;
    std::vector<RefJob> hj((size_t)nj);
    std::vector<ClassKey> ckey((size_t)nj);
    for (int32_t i = 0; i < nj; ++i) {
        const JobSpec &js = jobs[(size_t)i];
        const mvs_params &prm = js.prm;
        JobWork &wk = c->work[(size_t)i];
        const ViewDev &ref = *js.ref;
        const int64_t n = (int64_t)ref.w * ref.h, cap = (n + 1) / 2;
        const int32_t K = (int32_t)js.srcs.size();
        HIP_TRY(c, wk.hyp.ensure((size_t)n * sizeof(float4)));
        HIP_TRY(c, wk.cost.ensure((size_t)n * sizeof(float)));
        HIP_TRY(c, wk.textured.ensure((size_t)n));
        HIP_TRY(c, wk.changed.ensure((size_t)n));
        HIP_TRY(c, wk.pair.ensure((size_t)n));
        HIP_TRY(c, wk.list.ensure((size_t)(2 * cap) * sizeof(int32_t)));
        HIP_TRY(c, wk.counts.ensure((4 + 2 * (size_t)list_waves(ref.w, ref.h)) * sizeof(int32_t)));
        HIP_TRY(c, wk.srcs.ensure(sizeof(SrcCam) * (size_t)K));
        c->host_tables.emplace_back();
        std::vector<SrcCam> &tab = c->host_tables.back();
        bool pinhole = true, quads = true, quadds = ref.has_quadd, lane_size = true;
        for (const ViewDev *sv : js.srcs) {
            tab.push_back(make_src(ref, *sv));
            pinhole = pinhole && tab.back().pinhole_exact;
            quads = quads && sv->has_quad;
            quadds = quadds && sv->has_quadd;
            // the one-lane-per-pixel kernels address a source with 32-bit byte
            // offsets into its 8-byte-per-pixel plane
            lane_size = lane_size && (int64_t)sv->w * sv->h <= (1ll << 28);
        }
        HIP_TRY(c, hipMemcpyAsync(wk.srcs.p, tab.data(), sizeof(SrcCam) * (size_t)K,
                                  hipMemcpyHostToDevice, c->st));
        const float2 *rays = nullptr;
        TRY(get_rays(c, ref, &rays));
        RefJob &job = hj[(size_t)i];
        memset(&job, 0, sizeof job);
        job.gray = ref.gray.as<float>();
        job.rays = rays;
        job.hyp = wk.hyp.as<float4>();
        job.cost = wk.cost.as<float>();
        job.textured = wk.textured.as<uint8_t>();
        job.changed = wk.changed.as<uint8_t>();
        job.pair = wk.pair.as<uint8_t>();
        job.list = wk.list.as<int32_t>();
        job.counts = wk.counts.as<int32_t>();
        job.srcs = wk.srcs.as<SrcCam>();
        job.w = ref.w;
        job.h = ref.h;
        job.nsrc = K;
        job.list_cap = (int32_t)cap;
        // dense.rs:362-363
        job.dmin = (float)(js.range0 > 0.5 ? js.range0 : 0.5);
        job.dmax = (float)js.range1;
        job.imin = 1.0f / job.dmax;
        job.imax = 1.0f / job.dmin;
        // dense.rs:367-370: R_ref * (0, 0, 1)
        job.up0 = (float)ref.R[2];
        job.up1 = (float)ref.R[5];
        job.up2 = (float)ref.R[8];
        job.texture_min = prm.texture_min;
        job.cost_max = prm.cost_max;
        job.stride = prm.stride;
        job.half = (prm.taps / 2) * prm.stride;
        job.pinhole = pinhole ? 1 : 0;
        job.ray_step = (float)((double)prm.stride / ref.f);
        const int32_t steps = n_steps < 0 ? 2 * prm.iterations : n_steps;
        // 2: the reference is undistorted as well (its window rays are affine)
        // and 8-bit (its grey is a byte over 255)
        const bool ref_affine = ref.k1 == 0.0 && ref.k2 == 0.0 && lane_size;
        int32_t ph = pinhole ? (ref_affine && ref.has_quad ? 2 : 1) : 0;
        job.lane_ok = ph == 2 && quads && K <= 16 ? 1 : 0;
        // the fast kernels read the corner-packed 8-bit planes (fast = 1);
        // views all given as f32 grey on undistorted cameras run the
        // one-lane-per-pixel kernels on the f32 corner-difference plane
        // (fast = 2); other f32 views run the exact kernels
        int32_t fast = 0;
        if (prm.mode == MVS_MODE_FAST && quads) {
            fast = 1;
        } else if (prm.mode == MVS_MODE_FAST && quadds && pinhole && ref_affine && K <= 16) {
            fast = 2;
            ph = 2;
        }
        // f32 views on undistorted cameras: the one-lane-per-pixel
        // refinement on the f32 corner-difference plane, in either mode
        if (!job.lane_ok && quadds && pinhole && ref_affine && K <= 16) job.lane_ok = 2;
        // a lean view (8-bit, without the f32 corner plane) serves only the
        // fast kernels on undistorted views, which read the 8-bit and f16 planes
        bool lean = !ref.quadf.p;
        for (const ViewDev *sv : js.srcs) lean = lean || !sv->quadf.p;
        if (lean && !(fast && job.lane_ok))
            return fail(c, MVS_ERR_ARG,
                        "a lean view (MVS_SURVEY_LEAN_VIEWS) serves only fast-mode jobs on "
                        "undistorted 8-bit views");
        // The one-lane-per-pixel kernels loop over each job's own sources, so
        // fast jobs on undistorted views of any neighbour count share one class
        // (source lanes 0) unless a source-lane schedule is forced.
        const bool lanes_only = fast && ph == 2 && c->forced.split <= 0;
        ckey[(size_t)i] = ClassKey(prm.taps, steps, lanes_only ? 0 : lanes_for(K), ph, fast);
    }

    // group jobs by kernel class; jobs of a class are contiguous in the table
    std::map<ClassKey, std::vector<int32_t>> groups;
    for (int32_t i = 0; i < nj; ++i) groups[ckey[(size_t)i]].push_back(i);
    std::vector<KClass> classes;
    std::vector<RefJob> table;
    std::vector<int32_t> order;  // table position -> job index
    for (const auto &g : groups) {
        KClass k;
        k.taps = std::get<0>(g.first);
        k.steps = std::get<1>(g.first);
        k.srcl = std::get<2>(g.first);
        k.pinhole = std::get<3>(g.first);
        k.fast = std::get<4>(g.first);
        k.first = (int32_t)table.size();
        for (int32_t i : g.second) {
            table.push_back(hj[(size_t)i]);
            order.push_back(i);
            const int64_t n = (int64_t)hj[(size_t)i].w * hj[(size_t)i].h;
            k.cap_sweep += (n + 1) / 2;
            k.cap_eval += n;
            k.caps.push_back(n);
            if (k.count == 0) k.size = size_bucket(n);
            ++k.count;
        }
        classes.push_back(k);
    }
    const size_t tbl0 = c->host_jobs.size();
    c->host_jobs.insert(c->host_jobs.end(), table.begin(), table.end());
    HIP_TRY(c, c->jobs_dev.ensure(sizeof(RefJob) * table.size()));
    HIP_TRY(c, hipMemcpyAsync(c->jobs_dev.p, c->host_jobs.data() + tbl0,
                              sizeof(RefJob) * table.size(), hipMemcpyHostToDevice, c->st));

    unsigned long long *evals = c->evals.as<unsigned long long>();
    // Classes on the one-lane-per-pixel kernels that make at least a
    // half-sweep of each colour score the initial planes in those half-sweeps
    // (each pixel's own plane first, ahead of its candidates: the same costs,
    // and the pixel's window set up once); the others in a pass of their own.
    for (KClass &k : classes)
        k.own_init = first_step < 0 && k.fast && k.pinhole == 2 && k.steps >= 2 &&
                     c->forced.split <= 0;
    // initialisation: hypotheses, texture mask, parity lists, then the cost of
    // every textured pixel's initial hypothesis
    {
        hipEvent_t ea = nullptr, eb = nullptr;
        if (c->timing) {
            ea = take_event(c);
            eb = take_event(c);
            HIP_TRY(c, hipEventRecord(ea, c->st));
        }
        for (int32_t i = 0; i < nj; ++i) {
            launch_init(hj[(size_t)i], jobs[(size_t)i].prm.taps, first_step < 0 ? 1 : 0, c->st);
            HIP_TRY(c, hipGetLastError());
        }
        if (first_step < 0) {
            for (KClass &k : classes) {
                if (k.own_init) continue;
                k.sch = eval_schedule(k);
                PmArgs a = base_args(c, k);
                a.evals = evals;
                TRY(pm_launch(c, PM_EVAL_CURRENT, k.taps, k, a));
            }
        }
        if (c->timing) {
            HIP_TRY(c, hipEventRecord(eb, c->st));
            c->pending.push_back({ea, eb, 1});
        }
    }
    // half-sweeps (dense.rs:413-467); the schedule of each class is forced,
    // cached, or tuned per iteration on the class's first job. A half-sweep
    // whose schedules are all known runs as a cached graph. Classes with fewer
    // half-sweeps drop out when theirs are done.
    const int32_t s0 = first_step < 0 ? 0 : first_step;
    int32_t max_steps = 0;
    for (const KClass &k : classes) max_steps = std::max(max_steps, k.steps);
    std::vector<SweepLaunch> ls;
    std::vector<const KClass *> active;
    for (int32_t s = s0; s < s0 + max_steps; ++s) {
        if (cancel && *cancel) return fail(c, MVS_ERR_CANCELLED, "cancelled");
        if (step_report && step_report->fn &&
            step_report->fn(step_report->user, "patchmatch", s - s0, max_steps))
            return fail(c, MVS_ERR_CANCELLED, "cancelled");
        const int32_t it = s / 2, parity = s % 2;
        // the first half-sweep of each colour of this call
        const bool first_pair = s < s0 + 2;
        bool known = true;
        active.clear();
        for (KClass &k : classes) {
            if (s - s0 >= k.steps) continue;
            active.push_back(&k);
            if (c->forced.split != 0) {
                k.sch = forced_schedule(c, k);
            } else {
                TuneKey key(k.taps, k.srcl, k.pinhole, k.size, it, k.fast);
                auto f = c->tuned.find(key);
                if (f != c->tuned.end()) {
                    k.sch = f->second;
                } else if (c->autotune) {
                    TRY(autotune_class(c, k.taps, k, table.data() + k.first, it, parity,
                                       s >= s0 + 2, k.own_init && first_pair));
                    c->tuned[key] = k.sch;
                    known = false;
                } else {
                    k.sch = candidates(k, k.own_init && first_pair)[0];
                }
                // (a cached choice of a call that scored the initial planes
                // apart may be a source-lane schedule)
                if (k.own_init && first_pair && k.sch.split > 0)
                    k.sch = candidates(k, true)[0];
            }
        }
        hipEvent_t ea = nullptr, eb = nullptr;
        if (c->timing) {
            ea = take_event(c);
            eb = take_event(c);
            HIP_TRY(c, hipEventRecord(ea, c->st));
        }
        // both colours have had a half-sweep of this call from s0 + 2 on: the
        // candidates scored two half-sweeps before are known from then
        const bool skip_unchanged = s >= s0 + 2;
        TRY(sweep_launches(c, active, it, parity, skip_unchanged, first_pair, evals + 1,
                           evals + 4, ls));
        c->stats.sweep_kernels += (int64_t)ls.size();
        if (c->graphs && known)
            TRY(run_sweeps_graph(c, ls));
        else
            TRY(enqueue_sweeps(c, ls));
        if (c->timing) {
            HIP_TRY(c, hipEventRecord(eb, c->st));
            c->pending.push_back({ea, eb, 0});
        }
    }
    for (const JobSpec &js : jobs) c->stats.pixels += (int64_t)js.ref->w * js.ref->h;
    // Gauss-Newton refinement of the accepted pixels (no reference
    // counterpart), before the outputs are thresholded
    if (first_step < 0) {
        hipEvent_t ea = nullptr, eb = nullptr;
        bool any = false;
        for (int32_t i = 0; i < nj; ++i) {
            const JobSpec &js = jobs[(size_t)i];
            if (!js.out || js.prm.refine_steps <= 0) continue;
            if (!any && c->timing) {
                ea = take_event(c);
                eb = take_event(c);
                HIP_TRY(c, hipEventRecord(ea, c->st));
            }
            any = true;
            if (launch_refine(js.prm.taps, hj[(size_t)i], js.prm.refine_steps, evals + 2, c->st))
                return fail(c, MVS_ERR_ARG, "no refinement kernel for this window");
            HIP_TRY(c, hipGetLastError());
        }
        if (any && c->timing) {
            HIP_TRY(c, hipEventRecord(eb, c->st));
            c->pending.push_back({ea, eb, 2});
        }
    }
    // thresholded outputs (dense.rs:468-477)
    for (int32_t i = 0; i < nj; ++i) {
        MapDev *m = jobs[(size_t)i].out;
        if (!m) continue;
        TRY(ensure_map(c, *m, hj[(size_t)i].w, hj[(size_t)i].h, jobs[(size_t)i].depth_only));
        launch_threshold(hj[(size_t)i], m->depth.as<float>(),
                         m->keep_normal ? m->normal.as<float>() : nullptr, c->st);
        HIP_TRY(c, hipGetLastError());
        if (m->keep_cost)
            HIP_TRY(c, hipMemcpyAsync(m->cost.p, hj[(size_t)i].cost,
                                      sizeof(float) * (size_t)hj[(size_t)i].w * hj[(size_t)i].h,
                                      hipMemcpyDeviceToDevice, c->st));
        m->valid = true;
    }
    if (!classes.empty()) {
        c->stats.split = classes[0].sch.split;
        c->stats.gpc = classes[0].sch.gpc;
    }
    if (classes_out) *classes_out = classes;
    return MVS_OK;
}

int download_map(mvs_ctx *c, const MapDev &m, float *depth, float *normal, float *cost) {
    const size_t n = (size_t)m.w * m.h;
    if ((normal && !m.keep_normal) || (cost && !m.keep_cost))
        return fail(c, MVS_ERR_ARG, "this map keeps the depth only (MVS_SURVEY_DEPTH_ONLY)");
    if (depth)
        HIP_TRY(c, hipMemcpyAsync(depth, m.depth.p, n * sizeof(float), hipMemcpyDeviceToHost,
                                  c->st));
    if (normal)
        HIP_TRY(c, hipMemcpyAsync(normal, m.normal.p, 3 * n * sizeof(float),
                                  hipMemcpyDeviceToHost, c->st));
    if (cost)
        HIP_TRY(c, hipMemcpyAsync(cost, m.cost.p, n * sizeof(float), hipMemcpyDeviceToHost,
                                  c->st));
    return sync_ctx(c);
}

int upload_host_views(mvs_ctx *c, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                      std::vector<const ViewDev *> &out) {
    if (nsrc < 0) return fail(c, MVS_ERR_ARG, "negative source count");
    if (nsrc > 0 && !srcs) return fail(c, MVS_ERR_ARG, "null source views");
    if ((int32_t)c->tmp_views.size() < nsrc + 1) c->tmp_views.resize((size_t)nsrc + 1);
    TRY(upload_view(c, c->tmp_views[0], ref));
    out.clear();
    for (int32_t i = 0; i < nsrc; ++i) {
        TRY(upload_view(c, c->tmp_views[(size_t)i + 1], &srcs[i]));
        out.push_back(&c->tmp_views[(size_t)i + 1]);
    }
    return MVS_OK;
}

int abort_call(mvs_ctx *c, int rc) {
    (void)hipStreamSynchronize(c->st);
    collect_timing(c);
    c->host_jobs.clear();
    c->host_tables.clear();
    return rc;
}

// Work-buffer bytes of one job slot (hyp, cost, mask, lists, list counts,
// sources).
int64_t slot_bytes(const ViewDev &v) {
    const int64_t n = (int64_t)v.w * v.h;
    return n * (16 + 4 + 1 + 4) + 8 * list_waves(v.w, v.h) + 16 * (int64_t)sizeof(SrcCam) + 64;
}

}  // namespace

extern "C" {

void mvs_default_params(mvs_params *p) {
    if (!p) return;
    memset(p, 0, sizeof *p);
    p->taps = 5;
    p->stride = 2;
    p->iterations = 4;
    p->texture_min = 0.008f;
    p->cost_max = 0.45f;
}

int mvs_dense_params_for_profile(int32_t kind, int64_t out[4]) {
    static const int64_t tab[3][4] = {
        {640, 3, 3, 2000000}, {960, 4, 4, 5000000}, {1400, 5, 5, 12000000}};
    if (kind < 0 || kind > 2 || !out) return MVS_ERR_ARG;
    for (int i = 0; i < 4; ++i) out[i] = tab[kind][i];
    return MVS_OK;
}

int32_t mvs_select_neighbours(const double *centres, int32_t ncams, int32_t ci,
                              const int32_t *neighbours, int32_t nn, double range0, int32_t take,
                              int32_t *out) {
    if (!centres || ci < 0 || ci >= ncams || nn < 0 || (nn > 0 && !neighbours)) return 0;
    int32_t k = 0;
    const double lim = 0.02 * (range0 > 1.0 ? range0 : 1.0);
    const double *cc = centres + 3 * (size_t)ci;
    for (int32_t a = 0; a < nn && k < take; ++a) {
        const int32_t j = neighbours[a];
        if (j < 0 || j >= ncams) continue;
        const double *cj = centres + 3 * (size_t)j;
        const double dx = cj[0] - cc[0], dy = cj[1] - cc[1], dz = cj[2] - cc[2];
        const double base = sqrt(dx * dx + dy * dy + dz * dz);
        if (base > lim) out[k++] = j;
    }
    return k;
}

int mvs_create(int32_t device, mvs_ctx **out, char *err, size_t errlen) {
    auto report = [&](int code, const std::string &m) {
        if (err && errlen) snprintf(err, errlen, "%s", m.c_str());
        return code;
    };
    if (!out) return report(MVS_ERR_ARG, "null output pointer");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return report(MVS_ERR_DEVICE, std::string("no HIP device: ") + hipGetErrorString(e));
    if (device < 0 || device >= n) return report(MVS_ERR_ARG, "device index out of range");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return report(MVS_ERR_DEVICE, hipGetErrorString(e));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return report(MVS_ERR_DEVICE,
                      std::string("built for gfx950, device is ") + prop.gcnArchName);
    e = hipSetDevice(device);
    if (e != hipSuccess) return report(MVS_ERR_DEVICE, hipGetErrorString(e));
    mvs_ctx *c = new mvs_ctx();
    c->device = device;
    c->num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return report(MVS_ERR_DEVICE, hipGetErrorString(e));
    }
    // the allocation-free entry point's scratch on this device (its failure
    // only disables that entry point)
    e = scratch_for_device(device, &c->scratch);
    if (e != hipSuccess) {
        c->scratch = nullptr;
        c->scratch_err = std::string("scratch array address: ") + hipGetErrorString(e);
    }
    *out = c;
    return MVS_OK;
}

void mvs_destroy(mvs_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->st);
    for (auto &t : c->pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (auto e : c->free_events) (void)hipEventDestroy(e);
    for (auto &kv : c->graph_cache) (void)hipGraphExecDestroy(kv.second);
    for (auto &kv : c->rays) {
        kv.second->release();
        delete kv.second;
    }
    for (auto &wk : c->work) {
        DevBuf *b[] = {&wk.hyp, &wk.cost, &wk.textured, &wk.changed, &wk.pair, &wk.list,
                       &wk.counts, &wk.srcs};
        for (DevBuf *x : b) x->release();
    }
    DevBuf *bufs[] = {&c->jobs_dev, &c->prefix, &c->counters, &c->evals, &c->ok_count,
                      &c->up_u8,    &c->px,     &c->py,       &c->hyp_in, &c->cost_out,
                      &c->nbr_cams, &c->agree,  &c->save_hyp, &c->save_cost, &c->save_changed,
                      &c->save_pair, &c->pairs,    &c->grad_io, &c->gout,
                      &c->tmp_map.depth, &c->tmp_map.normal, &c->tmp_map.cost,
                      &c->prep_box, &c->prep_rgb, &c->prep_grey,
                      &c->dsm_xyz, &c->dsm_h, &c->dsm_a, &c->dsm_b, &c->dsm_c, &c->dsm_d,
                      &c->dsm_snap, &c->dsm_parent, &c->dsm_la, &c->dsm_lb, &c->dsm_ctr, &c->dsm_sel};
    for (DevBuf *b : bufs) b->release();
    for (auto &v : c->tmp_views) {
        v.gray.release();
        v.quadf.release();
        v.quad.release();
        v.quadh.release();
    }
    (void)hipStreamDestroy(c->st);
    delete c;
}

const char *mvs_last_error(const mvs_ctx *c) { return c ? c->err.c_str() : "null context"; }

int mvs_synchronize(mvs_ctx *c) {
    if (!c) return MVS_ERR_ARG;
    return sync_ctx(c);
}

int mvs_set_schedule(mvs_ctx *c, int32_t split, int32_t gpc, int32_t per_cu,
                     int32_t autotune) {
    if (!c) return MVS_ERR_ARG;
    if (split < -5 || gpc < 0 || gpc > 1024 || per_cu < 0 || per_cu > 1024)
        return fail(c, MVS_ERR_ARG, "schedule out of range (split >= -5, gpc and per_cu 0..1024)");
    c->forced = Schedule{split, gpc, per_cu};
    c->autotune = autotune != 0;
    return MVS_OK;
}

int mvs_set_graphs(mvs_ctx *c, int32_t enable) {
    if (!c) return MVS_ERR_ARG;
    if (!enable) {
        (void)hipStreamSynchronize(c->st);
        for (auto &kv : c->graph_cache) (void)hipGraphExecDestroy(kv.second);
        c->graph_cache.clear();
    }
    c->graphs = enable != 0;
    return MVS_OK;
}

int32_t mvs_tuned_schedules(mvs_ctx *c, int32_t *out, int32_t cap) {
    if (!c) return 0;
    int32_t k = 0;
    for (const auto &kv : c->tuned) {
        if (out && k < cap) {
            int32_t *r = out + 9 * k;
            r[0] = std::get<0>(kv.first);
            r[1] = std::get<1>(kv.first);
            r[2] = std::get<2>(kv.first);
            r[3] = std::get<3>(kv.first);
            r[4] = std::get<4>(kv.first);
            r[5] = std::get<5>(kv.first);
            r[6] = kv.second.split;
            r[7] = kv.second.gpc;
            r[8] = kv.second.per_cu;
        }
        ++k;
    }
    return k;
}

int mvs_depth_map(mvs_ctx *c, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                  double range0, double range1, const mvs_params *prm, float *out_depth,
                  float *out_normal, float *out_cost, const volatile int32_t *cancel,
                  mvs_progress_fn progress, void *user) {
    if (!c) return MVS_ERR_ARG;
    TRY(check_params(c, prm));
    if (nsrc < 1) return fail(c, MVS_NO_NEIGHBOURS, "no neighbours");
    if (!out_depth) return fail(c, MVS_ERR_ARG, "null depth output");
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<const ViewDev *> sv;
    TRY(upload_host_views(c, ref, srcs, nsrc, sv));
    std::vector<JobSpec> jobs(1);
    jobs[0].ref = &c->tmp_views[0];
    jobs[0].srcs = sv;
    jobs[0].range0 = range0;
    jobs[0].range1 = range1;
    jobs[0].out = &c->tmp_map;
    jobs[0].prm = *prm;
    const StepReport rep{progress, user};
    int rc = run_batch(c, jobs, -1, -1, cancel, nullptr, &rep);
    if (rc) return abort_call(c, rc);
    TRY(download_map(c, c->tmp_map, out_depth, out_normal, out_cost));
    const int32_t total = 2 * prm->iterations;
    if (progress && progress(user, "patchmatch", total, total))
        return fail(c, MVS_ERR_CANCELLED, "cancelled");
    return MVS_OK;
}

int mvs_ncc_cost(mvs_ctx *c, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                 double range0, double range1, const mvs_params *prm, int64_t n,
                 const int32_t *px, const int32_t *py, const float *hyp4, float *out_cost) {
    if (!c) return MVS_ERR_ARG;
    TRY(check_params(c, prm));
    if (nsrc < 1) return fail(c, MVS_NO_NEIGHBOURS, "no neighbours");
    if (n < 0 || n >= (1ll << 31) || (n > 0 && (!px || !py || !hyp4 || !out_cost)))
        return fail(c, MVS_ERR_ARG, "bad cost list");
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<const ViewDev *> sv;
    TRY(upload_host_views(c, ref, srcs, nsrc, sv));
    const ViewDev &rf = c->tmp_views[0];
    for (int64_t i = 0; i < n; ++i)
        if (px[i] < 0 || py[i] < 0 || px[i] >= rf.w || py[i] >= rf.h)
            return fail(c, MVS_ERR_ARG, "pixel outside the reference view");
    if (n == 0) return MVS_OK;
    HIP_TRY(c, c->px.ensure((size_t)n * 4));
    HIP_TRY(c, c->py.ensure((size_t)n * 4));
    HIP_TRY(c, c->hyp_in.ensure((size_t)n * 16));
    HIP_TRY(c, c->cost_out.ensure((size_t)n * 4));
    HIP_TRY(c, hipMemcpyAsync(c->px.p, px, (size_t)n * 4, hipMemcpyHostToDevice, c->st));
    HIP_TRY(c, hipMemcpyAsync(c->py.p, py, (size_t)n * 4, hipMemcpyHostToDevice, c->st));
    HIP_TRY(c, hipMemcpyAsync(c->hyp_in.p, hyp4, (size_t)n * 16, hipMemcpyHostToDevice, c->st));
    // Build the job (rays, sources, mask) without running PatchMatch.
    std::vector<JobSpec> jobs(1);
    jobs[0].ref = &rf;
    jobs[0].srcs = sv;
    jobs[0].range0 = range0;
    jobs[0].range1 = range1;
    jobs[0].prm = *prm;
    std::vector<KClass> classes;
    int rc = run_batch(c, jobs, 0, 0, nullptr, &classes);
    if (rc) return abort_call(c, rc);
    KClass k = classes[0];
    k.sch = c->forced.split != 0 ? forced_schedule(c, k) : candidates(k)[0];
    PmArgs a = base_args(c, k);
    a.px = c->px.as<int32_t>();
    a.py = c->py.as<int32_t>();
    a.hyp_in = c->hyp_in.as<float4>();
    a.out = c->cost_out.as<float>();
    a.n_items = n;
    rc = pm_launch(c, PM_COST_LIST, prm->taps, k, a, n);
    if (rc) return abort_call(c, rc);
    c->stats.split = k.sch.split;  // the layout that scored the list
    c->stats.gpc = k.sch.gpc;
    HIP_TRY(c, hipMemcpyAsync(out_cost, c->cost_out.p, (size_t)n * 4, hipMemcpyDeviceToHost,
                              c->st));
    return sync_ctx(c);
}

}  // extern "C"

namespace {

// The explicit (pixel, plane, pair) list calls: mvs_ncc_pair (terms = false:
// the autograd-style pair kernel) and mvs_refine_terms (terms = true: the
// refinement kernels' own Gauss-Newton terms).
int pair_list_call(mvs_ctx *c, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                   double range0, double range1, const mvs_params *prm, int64_t n,
                   const int32_t *px, const int32_t *py, const float *hyp4, int32_t *pairs,
                   const float *grad_out, float *out_cost, float *grad4, float *jtj10,
                   bool terms) {
    if (!c) return MVS_ERR_ARG;
    TRY(check_params(c, prm));
    if (nsrc < 2) return fail(c, MVS_NO_NEIGHBOURS, "a source pair needs two neighbours");
    if (n < 0 || n >= (1ll << 31) || (n > 0 && (!px || !py || !hyp4 || !pairs || !out_cost)))
        return fail(c, MVS_ERR_ARG, "bad pair list");
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<const ViewDev *> sv;
    TRY(upload_host_views(c, ref, srcs, nsrc, sv));
    const ViewDev &rf = c->tmp_views[0];
    for (int64_t i = 0; i < n; ++i) {
        if (px[i] < 0 || py[i] < 0 || px[i] >= rf.w || py[i] >= rf.h)
            return fail(c, MVS_ERR_ARG, "pixel outside the reference view");
        const int32_t a = pairs[2 * i], b = pairs[2 * i + 1];
        const bool choose = a == -1 && b == -1;
        if (!choose && (a < 0 || b < 0 || a >= nsrc || b >= nsrc || a == b))
            return fail(c, MVS_ERR_ARG, "source pair must be two distinct sources or (-1, -1)");
    }
    if (n == 0) return MVS_OK;
    HIP_TRY(c, c->px.ensure((size_t)n * 4));
    HIP_TRY(c, c->py.ensure((size_t)n * 4));
    HIP_TRY(c, c->hyp_in.ensure((size_t)n * 16));
    HIP_TRY(c, c->cost_out.ensure((size_t)n * 4));
    HIP_TRY(c, c->pairs.ensure((size_t)n * 8));
    HIP_TRY(c, c->grad_io.ensure((size_t)n * 16));
    HIP_TRY(c, c->gout.ensure((size_t)n * 4));
    if (jtj10) HIP_TRY(c, c->jtj_out.ensure((size_t)n * 40));
    HIP_TRY(c, hipMemcpyAsync(c->px.p, px, (size_t)n * 4, hipMemcpyHostToDevice, c->st));
    HIP_TRY(c, hipMemcpyAsync(c->py.p, py, (size_t)n * 4, hipMemcpyHostToDevice, c->st));
    HIP_TRY(c, hipMemcpyAsync(c->hyp_in.p, hyp4, (size_t)n * 16, hipMemcpyHostToDevice, c->st));
    HIP_TRY(c, hipMemcpyAsync(c->pairs.p, pairs, (size_t)n * 8, hipMemcpyHostToDevice, c->st));
    if (grad_out)
        HIP_TRY(c, hipMemcpyAsync(c->gout.p, grad_out, (size_t)n * 4, hipMemcpyHostToDevice,
                                  c->st));
    // Build the job (rays, sources, mask) without running PatchMatch.
    std::vector<JobSpec> jobs(1);
    jobs[0].ref = &rf;
    jobs[0].srcs = sv;
    jobs[0].range0 = range0;
    jobs[0].range1 = range1;
    jobs[0].prm = *prm;
    int rc = run_batch(c, jobs, 0, 0, nullptr, nullptr);
    if (rc) return abort_call(c, rc);
    const RefJob &job = c->host_jobs.back();
    if (terms) {
        rc = launch_refine_terms(prm->taps, job, n, c->px.as<int32_t>(), c->py.as<int32_t>(),
                                 c->hyp_in.as<float4>(), c->pairs.as<int32_t>(),
                                 c->cost_out.as<float>(),
                                 grad4 ? c->grad_io.as<float4>() : nullptr,
                                 jtj10 ? c->jtj_out.as<float>() : nullptr, c->st);
        // the layout that scored the list: one lane per pixel (-1) or one
        // lane per source of the pair (1)
        c->stats.split = job.lane_ok ? -1 : 1;
        c->stats.gpc = 1;
    } else {
        rc = launch_pair_list(prm->taps, job, n, c->px.as<int32_t>(), c->py.as<int32_t>(),
                              c->hyp_in.as<float4>(), c->pairs.as<int32_t>(),
                              grad_out ? c->gout.as<float>() : nullptr, c->cost_out.as<float>(),
                              grad4 ? c->grad_io.as<float4>() : nullptr, c->st);
    }
    if (rc) return abort_call(c, fail(c, MVS_ERR_ARG, "no pair kernel for this window"));
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(out_cost, c->cost_out.p, (size_t)n * 4, hipMemcpyDeviceToHost,
                              c->st));
    HIP_TRY(c, hipMemcpyAsync(pairs, c->pairs.p, (size_t)n * 8, hipMemcpyDeviceToHost, c->st));
    if (grad4)
        HIP_TRY(c, hipMemcpyAsync(grad4, c->grad_io.p, (size_t)n * 16, hipMemcpyDeviceToHost,
                                  c->st));
    if (jtj10)
        HIP_TRY(c, hipMemcpyAsync(jtj10, c->jtj_out.p, (size_t)n * 40, hipMemcpyDeviceToHost,
                                  c->st));
    return sync_ctx(c);
}

}  // namespace

extern "C" {

int mvs_ncc_pair(mvs_ctx *c, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                 double range0, double range1, const mvs_params *prm, int64_t n,
                 const int32_t *px, const int32_t *py, const float *hyp4, int32_t *pairs,
                 const float *grad_out, float *out_cost, float *grad4) {
    return pair_list_call(c, ref, srcs, nsrc, range0, range1, prm, n, px, py, hyp4, pairs,
                          grad_out, out_cost, grad4, nullptr, false);
}

int mvs_refine_terms(mvs_ctx *c, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                     double range0, double range1, const mvs_params *prm, int64_t n,
                     const int32_t *px, const int32_t *py, const float *hyp4, int32_t *pairs,
                     float *out_cost, float *grad4, float *jtj10) {
    return pair_list_call(c, ref, srcs, nsrc, range0, range1, prm, n, px, py, hyp4, pairs,
                          nullptr, out_cost, grad4, jtj10, true);
}

int mvs_patchmatch_steps(mvs_ctx *c, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                         double range0, double range1, const mvs_params *prm,
                         int32_t first_step, int32_t n_steps, float *state, uint8_t *textured,
                         int64_t *evals) {
    if (!c) return MVS_ERR_ARG;
    TRY(check_params(c, prm));
    if (nsrc < 1) return fail(c, MVS_NO_NEIGHBOURS, "no neighbours");
    if (!state || n_steps < 0 || first_step < -1)
        return fail(c, MVS_ERR_ARG, "bad state or steps");
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<const ViewDev *> sv;
    TRY(upload_host_views(c, ref, srcs, nsrc, sv));
    const ViewDev &rf = c->tmp_views[0];
    const size_t n = (size_t)rf.w * rf.h;
    TRY(sync_ctx(c));  // flush earlier evaluations into the stats
    const int64_t evals0 = c->stats.evals;
    std::vector<float> hyp4(4 * n), cst(n);
    if (c->work.empty()) c->work.resize(1);
    JobWork &wk = c->work[0];
    if (first_step >= 0) {
        HIP_TRY(c, wk.hyp.ensure(n * sizeof(float4)));
        HIP_TRY(c, wk.cost.ensure(n * sizeof(float)));
        for (size_t i = 0; i < n; ++i) {
            for (int k = 0; k < 4; ++k) hyp4[4 * i + k] = state[5 * i + k];
            cst[i] = state[5 * i + 4];
        }
        HIP_TRY(c, hipMemcpyAsync(wk.hyp.p, hyp4.data(), n * 16, hipMemcpyHostToDevice, c->st));
        HIP_TRY(c, hipMemcpyAsync(wk.cost.p, cst.data(), n * 4, hipMemcpyHostToDevice, c->st));
    }
    std::vector<JobSpec> jobs(1);
    jobs[0].ref = &rf;
    jobs[0].srcs = sv;
    jobs[0].range0 = range0;
    jobs[0].range1 = range1;
    jobs[0].prm = *prm;
    int rc = run_batch(c, jobs, first_step, n_steps, nullptr, nullptr);
    if (rc) return abort_call(c, rc);
    HIP_TRY(c, hipMemcpyAsync(hyp4.data(), wk.hyp.p, n * 16, hipMemcpyDeviceToHost, c->st));
    HIP_TRY(c, hipMemcpyAsync(cst.data(), wk.cost.p, n * 4, hipMemcpyDeviceToHost, c->st));
    if (textured)
        HIP_TRY(c, hipMemcpyAsync(textured, wk.textured.p, n, hipMemcpyDeviceToHost, c->st));
    TRY(sync_ctx(c));
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 4; ++k) state[5 * i + k] = hyp4[4 * i + k];
        state[5 * i + 4] = cst[i];
    }
    if (evals) *evals = c->stats.evals - evals0;
    return MVS_OK;
}

int mvs_survey_create(mvs_ctx *c, int32_t nviews, mvs_survey **out) {
    if (!c || !out || nviews < 1) return MVS_ERR_ARG;
    mvs_survey *s = new mvs_survey();
    s->ctx = c;
    s->views.resize((size_t)nviews);
    s->maps.resize((size_t)nviews);
    *out = s;
    return MVS_OK;
}

void mvs_survey_destroy(mvs_survey *s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->st);
    DevBuf *fb[] = {&s->fkeys,  &s->fsum,   &s->fcol,  &s->fcnt,   &s->fctr,
                    &s->fo_kin, &s->fo_kout, &s->fo_sin, &s->fo_sout, &s->fo_tmp,
                    &s->fo_xyz, &s->fo_rgb, &s->fo_ctr};
    for (DevBuf *b : fb) b->release();
    for (auto &v : s->views) {
        v.gray.release();
        v.rgb.release();
        v.quadf.release();
        v.quad.release();
        v.quadh.release();
    }
    for (auto &m : s->maps) {
        m.depth.release();
        m.normal.release();
        m.cost.release();
    }
    delete s;
}

int mvs_survey_set_view(mvs_survey *s, int32_t i, const mvs_view *v) {
    if (!s) return MVS_ERR_ARG;
    mvs_ctx *c = s->ctx;
    if (i < 0 || i >= (int32_t)s->views.size())
        return fail(c, MVS_ERR_ARG, "view index out of range");
    HIP_TRY(c, hipSetDevice(c->device));
    return upload_view(c, s->views[(size_t)i], v, (s->options & MVS_SURVEY_LEAN_VIEWS) != 0);
}

int mvs_survey_set_options(mvs_survey *s, int32_t options) {
    if (!s) return MVS_ERR_ARG;
    if (options & ~(MVS_SURVEY_LEAN_VIEWS | MVS_SURVEY_DEPTH_ONLY))
        return fail(s->ctx, MVS_ERR_ARG, "unknown survey option");
    s->options = options;
    return MVS_OK;
}

}  // extern "C"

namespace {

// The survey loop of dense.rs:528-540 over resident views: jobs in CSR form,
// params per job (prm_step 1) or shared (prm_step 0), batched by memory budget.
int survey_depth_maps(mvs_survey *s, int32_t njobs, const int32_t *refs, const int32_t *src_off,
                      const int32_t *srcs, const double *ranges, const mvs_params *prms,
                      int32_t prm_step, const volatile int32_t *cancel, mvs_progress_fn progress,
                      void *user) {
    mvs_ctx *c = s->ctx;
    if (njobs < 0 || (njobs > 0 && (!refs || !src_off || !ranges || !prms)))
        return fail(c, MVS_ERR_ARG, "bad job arrays");
    for (int32_t j = 0; j < (prm_step ? njobs : 1); ++j) TRY(check_params(c, &prms[j]));
    HIP_TRY(c, hipSetDevice(c->device));
    const int32_t nv = (int32_t)s->views.size();
    // validate; views without neighbours get no map (the reference's Ok(None))
    std::vector<JobSpec> all;
    for (int32_t j = 0; j < njobs; ++j) {
        const int32_t ref = refs[j];
        if (ref < 0 || ref >= nv || !s->views[(size_t)ref].set)
            return fail(c, MVS_ERR_ARG, "reference view missing");
        JobSpec js;
        js.ref = &s->views[(size_t)ref];
        for (int32_t k = src_off[j]; k < src_off[j + 1]; ++k) {
            const int32_t q = srcs[k];
            if (q < 0 || q >= nv || !s->views[(size_t)q].set || q == ref)
                return fail(c, MVS_ERR_ARG, "bad source view index");
            js.srcs.push_back(&s->views[(size_t)q]);
        }
        js.range0 = ranges[2 * j];
        js.range1 = ranges[2 * j + 1];
        js.prm = prms[prm_step ? j : 0];
        js.out = &s->maps[(size_t)ref];
        js.depth_only = (s->options & MVS_SURVEY_DEPTH_ONLY) != 0;
        if (js.srcs.empty()) {
            s->maps[(size_t)ref].valid = false;
            continue;
        }
        all.push_back(js);
    }
    // batches bounded by a work-buffer budget (a quarter of free memory)
    size_t free_b = 0, total_b = 0;
    HIP_TRY(c, hipMemGetInfo(&free_b, &total_b));
    const int64_t budget = std::max<int64_t>((int64_t)(free_b / 4), 1ll << 30);
    size_t i = 0;
    int32_t done = 0;
    while (i < all.size()) {
        std::vector<JobSpec> batch;
        int64_t bytes = 0;
        while (i < all.size() && (batch.empty() || bytes + slot_bytes(*all[i].ref) <= budget)) {
            bytes += slot_bytes(*all[i].ref);
            batch.push_back(all[i++]);
        }
        int rc = run_batch(c, batch, -1, -1, cancel, nullptr);
        if (rc) return abort_call(c, rc);
        if (progress) {
            TRY(sync_ctx(c));
            // Ctx::report("depth", k, n) per finished view (dense.rs:539)
            for (size_t b = 0; b < batch.size(); ++b)
                if (progress(user, "depth", ++done, (int32_t)all.size()))
                    return fail(c, MVS_ERR_CANCELLED, "cancelled");
        } else {
            done += (int32_t)batch.size();
        }
    }
    return sync_ctx(c);
}

}  // namespace

extern "C" {

int mvs_survey_depth_maps(mvs_survey *s, int32_t njobs, const int32_t *refs,
                          const int32_t *src_off, const int32_t *srcs, const double *ranges,
                          const mvs_params *prm, const volatile int32_t *cancel,
                          mvs_progress_fn progress, void *user) {
    if (!s) return MVS_ERR_ARG;
    return survey_depth_maps(s, njobs, refs, src_off, srcs, ranges, prm, 0, cancel, progress,
                             user);
}

int mvs_survey_depth_maps_ex(mvs_survey *s, int32_t njobs, const int32_t *refs,
                             const int32_t *src_off, const int32_t *srcs, const double *ranges,
                             const mvs_params *prms, const volatile int32_t *cancel,
                             mvs_progress_fn progress, void *user) {
    if (!s) return MVS_ERR_ARG;
    return survey_depth_maps(s, njobs, refs, src_off, srcs, ranges, prms, 1, cancel, progress,
                             user);
}

int mvs_survey_depth_map(mvs_survey *s, int32_t ref, const int32_t *srcs, int32_t nsrc,
                         double range0, double range1, const mvs_params *prm) {
    if (!s) return MVS_ERR_ARG;
    if (nsrc < 1) return fail(s->ctx, MVS_NO_NEIGHBOURS, "no neighbours");
    const int32_t off[2] = {0, nsrc};
    const double rng[2] = {range0, range1};
    return mvs_survey_depth_maps(s, 1, &ref, off, srcs, rng, prm, nullptr, nullptr, nullptr);
}

int mvs_survey_download(mvs_survey *s, int32_t view, float *depth, float *normal, float *cost) {
    if (!s) return MVS_ERR_ARG;
    mvs_ctx *c = s->ctx;
    if (view < 0 || view >= (int32_t)s->maps.size() || !s->maps[(size_t)view].valid)
        return fail(c, MVS_ERR_ARG, "no depth map for that view");
    HIP_TRY(c, hipSetDevice(c->device));
    return download_map(c, s->maps[(size_t)view], depth, normal, cost);
}

int mvs_survey_set_depth(mvs_survey *s, int32_t view, const float *depth) {
    if (!s) return MVS_ERR_ARG;
    mvs_ctx *c = s->ctx;
    if (view < 0 || view >= (int32_t)s->views.size() || !s->views[(size_t)view].set)
        return fail(c, MVS_ERR_ARG, "no such view");
    if (!depth) return fail(c, MVS_ERR_ARG, "null depth map");
    HIP_TRY(c, hipSetDevice(c->device));
    const ViewDev &v = s->views[(size_t)view];
    MapDev &m = s->maps[(size_t)view];
    TRY(ensure_map(c, m, v.w, v.h, (s->options & MVS_SURVEY_DEPTH_ONLY) != 0));
    const size_t n = (size_t)v.w * v.h;
    HIP_TRY(c, hipMemcpy(m.depth.p, depth, n * sizeof(float), hipMemcpyHostToDevice));
    if (m.keep_normal) HIP_TRY(c, hipMemset(m.normal.p, 0, 3 * n * sizeof(float)));
    if (m.keep_cost) HIP_TRY(c, hipMemset(m.cost.p, 0, n * sizeof(float)));
    m.valid = true;
    return MVS_OK;
}

}  // extern "C"

namespace {

RefCam64 ref_cam(const ViewDev &rv);

// Arguments of a consistency call: checks them, fills cams[0..nn) (capacity
// cap) with the neighbours' cameras and maps. No allocation.
int consistency_args(mvs_survey *s, int32_t ref, const int32_t *nbrs, int32_t nn, int32_t rule,
                     NbrCam *cams, int32_t cap) {
    mvs_ctx *c = s->ctx;
    if (rule != MVS_CONS_REFERENCE && rule != MVS_CONS_REPROJ)
        return fail(c, MVS_ERR_ARG, "unknown consistency rule");
    const int32_t nv = (int32_t)s->views.size();
    if (ref < 0 || ref >= nv || !s->maps[(size_t)ref].valid)
        return fail(c, MVS_ERR_ARG, "reference view has no depth map");
    if (nn < 0 || (nn > 0 && !nbrs)) return fail(c, MVS_ERR_ARG, "bad neighbour list");
    if (nn > cap) return fail(c, MVS_ERR_ARG, "too many neighbours");
    for (int32_t k = 0; k < nn; ++k) {
        const int32_t j = nbrs[k];
        if (j < 0 || j >= nv || !s->maps[(size_t)j].valid)
            return fail(c, MVS_ERR_ARG, "neighbour without a depth map");
        const ViewDev &v = s->views[(size_t)j];
        NbrCam &nc = cams[k];
        memcpy(nc.R, v.R, sizeof nc.R);
        memcpy(nc.C, v.C, sizeof nc.C);
        nc.f = v.f;
        nc.cx = v.cx;
        nc.cy = v.cy;
        nc.k1 = v.k1;
        nc.k2 = v.k2;
        nc.depth = s->maps[(size_t)j].depth.as<float>();
        nc.w = v.w;
        nc.h = v.h;
    }
    return MVS_OK;
}

}  // namespace

extern "C" {

int mvs_survey_consistency(mvs_survey *s, int32_t ref, const int32_t *nbrs, int32_t nn,
                           int32_t rule, int32_t *agree, int64_t *n_consistent) {
    if (!s) return MVS_ERR_ARG;
    mvs_ctx *c = s->ctx;
    std::vector<NbrCam> cams((size_t)std::max(nn, 0));
    TRY(consistency_args(s, ref, nbrs, nn, rule, cams.data(), std::max(nn, 0)));
    HIP_TRY(c, hipSetDevice(c->device));
    const ViewDev &rv = s->views[(size_t)ref];
    const RefCam64 rc64 = ref_cam(rv);
    const size_t n = (size_t)rv.w * rv.h;
    HIP_TRY(c, c->nbr_cams.ensure(sizeof(NbrCam) * (cams.size() + 1)));
    HIP_TRY(c, c->agree.ensure(n * sizeof(int32_t)));
    HIP_TRY(c, c->ok_count.ensure(sizeof(unsigned long long)));
    if (!cams.empty())
        HIP_TRY(c, hipMemcpyAsync(c->nbr_cams.p, cams.data(), sizeof(NbrCam) * cams.size(),
                                  hipMemcpyHostToDevice, c->st));
    HIP_TRY(c, hipMemsetAsync(c->ok_count.p, 0, sizeof(unsigned long long), c->st));
    // dense.rs:585 for the reference rule; three views for the reprojection rule
    const int32_t need = rule == MVS_CONS_REPROJ ? 3 : (nn >= 3 ? 2 : 1);
    launch_consistency(rc64, s->maps[(size_t)ref].depth.as<float>(), rv.w, rv.h,
                       c->nbr_cams.as<NbrCam>(), nn, need, rule, c->agree.as<int32_t>(),
                       c->ok_count.as<unsigned long long>(), c->st);
    HIP_TRY(c, hipGetLastError());
    unsigned long long ok = 0;
    if (agree)
        HIP_TRY(c, hipMemcpyAsync(agree, c->agree.p, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                                  c->st));
    HIP_TRY(c, hipMemcpyAsync(&ok, c->ok_count.p, sizeof ok, hipMemcpyDeviceToHost, c->st));
    TRY(sync_ctx(c));
    if (n_consistent) *n_consistent = (int64_t)ok;
    return MVS_OK;
}

int mvs_survey_consistency_noalloc(mvs_survey *s, int32_t ref, const int32_t *nbrs, int32_t nn,
                                   int32_t rule, int32_t *agree, int64_t *n_consistent) {
    if (!s) return MVS_ERR_ARG;
    mvs_ctx *c = s->ctx;
    if (!c->scratch) return fail(c, MVS_ERR_DEVICE, c->scratch_err);
    NbrCam cams[MVS_NOALLOC_MAX_NBRS];
    TRY(consistency_args(s, ref, nbrs, nn, rule, cams, MVS_NOALLOC_MAX_NBRS));
    const ViewDev &rv = s->views[(size_t)ref];
    const int64_t n = (int64_t)rv.w * rv.h;
    if (n > (int64_t)MVS_NOALLOC_MAX_PIXELS)
        return fail(c, MVS_ERR_ARG, "view larger than MVS_NOALLOC_MAX_PIXELS");
    HIP_TRY(c, hipSetDevice(c->device));
    // this device's scratch: [counter | camera table | agree plane]
    unsigned long long *ctr = (unsigned long long *)c->scratch;
    NbrCam *dcams = (NbrCam *)(c->scratch + kNaHead);
    int32_t *dagree = (int32_t *)(c->scratch + kNaHead + kNaCams);
    if (nn > 0)
        HIP_TRY(c, hipMemcpyAsync(dcams, cams, sizeof(NbrCam) * (size_t)nn, hipMemcpyHostToDevice,
                                  c->st));
    launch_zero_u64(ctr, 1, c->st);  // the last call's count is still there
    const int32_t need = rule == MVS_CONS_REPROJ ? 3 : (nn >= 3 ? 2 : 1);
    launch_consistency(ref_cam(rv), s->maps[(size_t)ref].depth.as<float>(), rv.w, rv.h, dcams, nn,
                       need, rule, dagree, ctr, c->st);
    HIP_TRY(c, hipGetLastError());
    unsigned long long ok = 0;
    if (agree)
        HIP_TRY(c, hipMemcpyAsync(agree, dagree, (size_t)n * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, c->st));
    HIP_TRY(c, hipMemcpyAsync(&ok, ctr, sizeof ok, hipMemcpyDeviceToHost, c->st));
    TRY(sync_ctx(c));
    if (n_consistent) *n_consistent = (int64_t)ok;
    return MVS_OK;
}

int mvs_selftest(mvs_ctx *c, int32_t which, int64_t n, uint64_t seed, int64_t *mismatches) {
    if (!c || !mismatches || n < 0 || n > (1ll << 34)) return MVS_ERR_ARG;
    if (which != MVS_SELFTEST_DIVISION) return fail(c, MVS_ERR_ARG, "unknown self-test");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, c->ok_count.ensure(sizeof(unsigned long long)));
    HIP_TRY(c, hipMemsetAsync(c->ok_count.p, 0, sizeof(unsigned long long), c->st));
    const int64_t chunk = 1ll << 28;
    for (int64_t done = 0; done < n; done += chunk) {
        const int64_t m = n - done < chunk ? n - done : chunk;
        launch_selftest_div(m, seed + (uint64_t)done * 0x9E3779B97F4A7C15ull,
                            c->ok_count.as<unsigned long long>(), c->st);
        HIP_TRY(c, hipGetLastError());
    }
    unsigned long long bad = 0;
    HIP_TRY(c, hipMemcpyAsync(&bad, c->ok_count.p, sizeof bad, hipMemcpyDeviceToHost, c->st));
    TRY(sync_ctx(c));
    *mismatches = (int64_t)bad;
    return MVS_OK;
}

int mvs_set_timing(mvs_ctx *c, int32_t enable) {
    if (!c) return MVS_ERR_ARG;
    c->timing = enable != 0;
    return MVS_OK;
}

int mvs_get_stats(mvs_ctx *c, mvs_stats *out) {
    if (!c || !out) return MVS_ERR_ARG;
    TRY(sync_ctx(c));
    *out = c->stats;
    return MVS_OK;
}

int mvs_reset_stats(mvs_ctx *c) {
    if (!c) return MVS_ERR_ARG;
    TRY(sync_ctx(c));
    memset(&c->stats, 0, sizeof c->stats);
    return MVS_OK;
}

int mvs_survey_set_colour(mvs_survey *s, int32_t view, const uint8_t *rgb) {
    if (!s) return MVS_ERR_ARG;
    mvs_ctx *c = s->ctx;
    if (view < 0 || view >= (int32_t)s->views.size() || !s->views[(size_t)view].set)
        return fail(c, MVS_ERR_ARG, "view missing");
    if (!rgb) return fail(c, MVS_ERR_ARG, "null colour plane");
    HIP_TRY(c, hipSetDevice(c->device));
    ViewDev &v = s->views[(size_t)view];
    const size_t n = (size_t)v.w * v.h * 3;
    HIP_TRY(c, v.rgb.ensure(n));
    HIP_TRY(c, hipMemcpyAsync(v.rgb.p, rgb, n, hipMemcpyHostToDevice, c->st));
    HIP_TRY(c, hipStreamSynchronize(c->st));
    v.has_rgb = true;
    return MVS_OK;
}

}  // extern "C"

namespace {

// Neighbour cameras of a fusion/consistency step (dense.rs:575-584): those of
// the given list that have a depth map.
void nbr_cams(mvs_survey *s, const int32_t *nbrs, int32_t nn, std::vector<NbrCam> &cams) {
    cams.clear();
    for (int32_t k = 0; k < nn; ++k) {
        const int32_t j = nbrs[k];
        if (j < 0 || j >= (int32_t)s->views.size() || !s->maps[(size_t)j].valid) continue;
        const ViewDev &v = s->views[(size_t)j];
        NbrCam nc;
        memcpy(nc.R, v.R, sizeof nc.R);
        memcpy(nc.C, v.C, sizeof nc.C);
        nc.f = v.f;
        nc.cx = v.cx;
        nc.cy = v.cy;
        nc.k1 = v.k1;
        nc.k2 = v.k2;
        nc.depth = s->maps[(size_t)j].depth.as<float>();
        nc.w = v.w;
        nc.h = v.h;
        cams.push_back(nc);
    }
}

RefCam64 ref_cam(const ViewDev &rv) {
    RefCam64 r;
    memcpy(r.R, rv.R, sizeof r.R);
    memcpy(r.C, rv.C, sizeof r.C);
    r.f = rv.f;
    r.cx = rv.cx;
    r.cy = rv.cy;
    r.k1 = rv.k1;
    r.k2 = rv.k2;
    return r;
}

FuseTable fuse_table(mvs_survey *s) {
    FuseTable t;
    t.keys = s->fkeys.as<unsigned long long>();
    t.sum = s->fsum.as<unsigned long long>();
    t.col = s->fcol.as<unsigned int>();
    t.cnt = s->fcnt.as<unsigned int>();
    t.counters = s->fctr.as<unsigned long long>();
    t.cap = s->fcap;
    t.limit = s->fcap / 10 * 7;
    return t;
}

}  // namespace

extern "C" {

int mvs_survey_fuse(mvs_survey *s, int32_t nrefs, const int32_t *refs, const int32_t *nbr_off,
                    const int32_t *nbrs, int32_t rule, double voxel, int64_t *n_voxels) {
    if (!s) return MVS_ERR_ARG;
    mvs_ctx *c = s->ctx;
    if (rule != MVS_CONS_REFERENCE && rule != MVS_CONS_REPROJ)
        return fail(c, MVS_ERR_ARG, "unknown consistency rule");
    if (!(voxel > 0.0)) return fail(c, MVS_ERR_ARG, "voxel size must be positive");
    if (nrefs < 0 || (nrefs > 0 && (!refs || !nbr_off))) return fail(c, MVS_ERR_ARG, "bad refs");
    const int32_t nv = (int32_t)s->views.size();
    int64_t pixels = 0;
    for (int32_t r = 0; r < nrefs; ++r) {
        if (refs[r] < 0 || refs[r] >= nv) return fail(c, MVS_ERR_ARG, "bad reference index");
        if (s->maps[(size_t)refs[r]].valid)
            pixels += (int64_t)s->views[(size_t)refs[r]].w * s->views[(size_t)refs[r]].h;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    // a table of about one slot per 32 pixels (a voxel gathers tens of
    // points); a pass that fills it past 70% stops inserting, and the pass is
    // repeated with a table sized from the voxels per point it saw
    constexpr int64_t kMaxCap = 1ll << 30;
    int64_t cap = 1 << 16;
    while (cap < pixels / 32 && cap < kMaxCap) cap <<= 1;
    std::vector<NbrCam> cams;
    for (;;) {  // ends: every repeat at least quadruples the table, up to 2^30 slots
        c->stats.fuse_passes += 1;
        HIP_TRY(c, s->fkeys.ensure((size_t)cap * 8));
        HIP_TRY(c, s->fsum.ensure((size_t)cap * 24));
        HIP_TRY(c, s->fcol.ensure((size_t)cap * 12));
        HIP_TRY(c, s->fcnt.ensure((size_t)cap * 4));
        HIP_TRY(c, s->fctr.ensure(4 * 8));
        s->fcap = cap;
        const FuseTable t = fuse_table(s);
        launch_fuse_clear(t, c->st);
        HIP_TRY(c, hipGetLastError());
        // every view's neighbour cameras in one upload, then one launch per
        // view with no host round trip; the counters are read once at the end
        // (an insert into a full table is counted, never lost silently)
        std::vector<NbrCam> all;
        std::vector<int64_t> first((size_t)nrefs + 1, 0);
        for (int32_t r = 0; r < nrefs; ++r) {
            nbr_cams(s, nbrs + nbr_off[r], nbr_off[r + 1] - nbr_off[r], cams);
            all.insert(all.end(), cams.begin(), cams.end());
            first[(size_t)r + 1] = (int64_t)all.size();
        }
        HIP_TRY(c, c->nbr_cams.ensure(sizeof(NbrCam) * (all.size() + 1)));
        if (!all.empty())
            HIP_TRY(c, hipMemcpyAsync(c->nbr_cams.p, all.data(), sizeof(NbrCam) * all.size(),
                                      hipMemcpyHostToDevice, c->st));
        for (int32_t r = 0; r < nrefs; ++r) {
            const int32_t ref = refs[r];
            if (!s->maps[(size_t)ref].valid) continue;  // dense.rs:571-573
            const int32_t nn = (int32_t)(first[(size_t)r + 1] - first[(size_t)r]);
            // dense.rs:585 for the reference rule; three views for reprojection
            const int32_t need = rule == MVS_CONS_REPROJ ? 3 : (nn >= 3 ? 2 : 1);
            const ViewDev &rv = s->views[(size_t)ref];
            launch_fuse(ref_cam(rv), s->maps[(size_t)ref].depth.as<float>(), rv.w, rv.h,
                        c->nbr_cams.as<NbrCam>() + first[(size_t)r], nn, rule, need, 1.0 / voxel,
                        rv.has_rgb ? rv.rgb.as<uint8_t>() : nullptr,
                        rv.has_quad ? rv.quad.as<uint32_t>() : nullptr, rv.gray.as<float>(), t,
                        c->st);
            HIP_TRY(c, hipGetLastError());
        }
        // the host vector must outlive the asynchronous upload
        HIP_TRY(c, hipStreamSynchronize(c->st));
        unsigned long long ctr[4];
        HIP_TRY(c, hipMemcpy(ctr, s->fctr.p, sizeof ctr, hipMemcpyDeviceToHost));
        const bool grow = (ctr[2] > 0 || (double)ctr[0] > 0.7 * (double)cap) && cap < kMaxCap;
        if (ctr[3] > 0)
            return fail(c, MVS_ERR_ARG, "voxel index beyond +-2^20: voxel too small for the scene");
        if (grow) {
            // every confirmed point was counted, stored or not: scale the
            // voxels seen by the points still missing (an over-estimate, as
            // later views revisit voxels of earlier ones) and start again with
            // the table at most half full
            const double pts = (double)ctr[1] + (double)ctr[2];
            const double est = ctr[1] ? (double)ctr[0] * pts / (double)ctr[1] : 4.0 * (double)cap;
            int64_t next = cap << 2;
            while ((double)next < 2.0 * est && next < kMaxCap) next <<= 1;
            cap = next;
            continue;
        }
        if (ctr[2] > 0) return fail(c, MVS_ERR_NOMEM, "fusion table full");
        s->fvoxel = voxel;
        s->fvox = (int64_t)ctr[0];
        s->fpts_ready = false;
        s->fused = true;
        if (n_voxels) *n_voxels = (int64_t)ctr[0];
        return MVS_OK;
    }
}

int mvs_survey_fused_points(mvs_survey *s, float *xyz, uint8_t *rgb, int64_t cap, int64_t *n) {
    if (!s) return MVS_ERR_ARG;
    mvs_ctx *c = s->ctx;
    if (!s->fused) return fail(c, MVS_ERR_ARG, "no fusion has run");
    HIP_TRY(c, hipSetDevice(c->device));
    const int64_t nv = s->fvox;
    if (n) *n = nv;
    if (!xyz && !rgb) return MVS_OK;
    if (cap < nv) return fail(c, MVS_ERR_ARG, "output arrays too small");
    if (nv > 0xffffffffll) return fail(c, MVS_ERR_ARG, "more than 2^32 voxels");
    if (!s->fpts_ready && nv > 0) {
        // compaction, radix sort by voxel key and the means, on the device
        const size_t tb = fuse_points_tmp_bytes(nv);
        if (!tb) return fail(c, MVS_ERR_DEVICE, "radix sort size query failed");
        HIP_TRY(c, s->fo_kin.ensure((size_t)nv * 8));
        HIP_TRY(c, s->fo_kout.ensure((size_t)nv * 8));
        HIP_TRY(c, s->fo_sin.ensure((size_t)nv * 4));
        HIP_TRY(c, s->fo_sout.ensure((size_t)nv * 4));
        HIP_TRY(c, s->fo_tmp.ensure(tb));
        HIP_TRY(c, s->fo_xyz.ensure((size_t)nv * 12));
        HIP_TRY(c, s->fo_rgb.ensure((size_t)nv * 3));
        HIP_TRY(c, s->fo_ctr.ensure(8));
        FusePointsWork w{s->fo_kin.as<unsigned long long>(), s->fo_kout.as<unsigned long long>(),
                         s->fo_sin.as<uint32_t>(), s->fo_sout.as<uint32_t>(), s->fo_tmp.p, tb,
                         s->fo_ctr.as<unsigned long long>()};
        if (fuse_points(fuse_table(s), nv, w, s->fo_xyz.as<float>(), s->fo_rgb.as<uint8_t>(),
                        c->st))
            return hip_fail(c, hipGetLastError(), "fused points");
        unsigned long long got = 0;
        HIP_TRY(c, hipMemcpyAsync(&got, w.ctr, 8, hipMemcpyDeviceToHost, c->st));
        HIP_TRY(c, hipStreamSynchronize(c->st));
        if ((int64_t)got != nv)
            return fail(c, MVS_ERR_DEVICE, "fusion table and voxel count disagree");
        s->fpts_ready = true;
    }
    if (nv > 0) {
        if (xyz)
            HIP_TRY(c, hipMemcpyAsync(xyz, s->fo_xyz.p, (size_t)nv * 12, hipMemcpyDeviceToHost,
                                      c->st));
        if (rgb)
            HIP_TRY(c, hipMemcpyAsync(rgb, s->fo_rgb.p, (size_t)nv * 3, hipMemcpyDeviceToHost,
                                      c->st));
        HIP_TRY(c, hipStreamSynchronize(c->st));
    }
    return MVS_OK;
}

int mvs_fusion_voxel(const double *xs, const double *ys, int64_t npts, const double *ranges,
                     int32_t ncams, double f, int32_t size, int32_t width, int32_t height,
                     int64_t max_points, double *voxel) {
    if (!voxel || npts < 1 || ncams < 1 || !xs || !ys || !ranges || size < 1 || width < 1 ||
        height < 1 || max_points < 1)
        return MVS_ERR_ARG;
    // dense.rs:545-562
    std::vector<double> sx(xs, xs + npts), sy(ys, ys + npts), dm((size_t)ncams);
    std::sort(sx.begin(), sx.end());
    std::sort(sy.begin(), sy.end());
    for (int32_t i = 0; i < ncams; ++i) dm[(size_t)i] = (ranges[2 * i] + ranges[2 * i + 1]) / 2.0;
    std::sort(dm.begin(), dm.end());
    auto pct = [&](const std::vector<double> &v, double q) {
        return v[(size_t)((double)(v.size() - 1) * q)];
    };
    double area = (pct(sx, 0.99) - pct(sx, 0.01)) * (pct(sy, 0.99) - pct(sy, 0.01));
    if (!(area >= 1.0)) area = 1.0;
    const double f0 = f * (double)size / (double)std::max(width, height);
    const double gsd = dm[dm.size() / 2] / f0;
    const double alt = std::sqrt(area * 2.5 / (double)max_points);
    *voxel = std::max(gsd * 1.5, alt);
    return MVS_OK;
}

int mvs_write_ply(const char *path, const float *xyz, const uint8_t *rgb, int64_t n,
                  const char *comment) {
    if (!path || n < 0 || (n > 0 && (!xyz || !rgb))) return MVS_ERR_ARG;
    // write_ply (ply.rs:11-40): ASCII header, then per vertex 3 f32 LE and 3
    // u8, written to path's ".ply.tmp" sibling and renamed
    std::string tmp = std::string(path);
    const size_t slash = tmp.find_last_of('/');
    const size_t dot = tmp.find_last_of('.');
    if (dot != std::string::npos && (slash == std::string::npos || dot > slash))
        tmp = tmp.substr(0, dot);
    tmp += ".ply.tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return MVS_ERR_IO;
    std::string hdr = "ply\nformat binary_little_endian 1.0\ncomment ";
    hdr += comment ? comment : "";
    hdr += "\nelement vertex " + std::to_string(n) +
           "\nproperty float x\nproperty float y\nproperty float z\n"
           "property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n";
    bool ok = fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
    std::vector<uint8_t> rec((size_t)15 * 65536);
    for (int64_t i0 = 0; ok && i0 < n; i0 += 65536) {
        const int64_t m = std::min<int64_t>(65536, n - i0);
        for (int64_t i = 0; i < m; ++i) {
            memcpy(&rec[(size_t)(15 * i)], xyz + 3 * (i0 + i), 12);  // little-endian hosts
            memcpy(&rec[(size_t)(15 * i + 12)], rgb + 3 * (i0 + i), 3);
        }
        ok = fwrite(rec.data(), 15, (size_t)m, f) == (size_t)m;
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok || rename(tmp.c_str(), path) != 0) return MVS_ERR_IO;
    return MVS_OK;
}

int mvs_read_ply(const char *path, float *xyz, uint8_t *rgb, int64_t cap, int64_t *n) {
    if (!path || !n) return MVS_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return MVS_ERR_IO;
    // read_ply (ply.rs:44-70): "element vertex N", then records after end_header
    char line[4096];
    int64_t count = -1;
    bool header = false;
    while (fgets(line, sizeof line, f)) {
        if (strncmp(line, "element vertex ", 15) == 0) count = strtoll(line + 15, nullptr, 10);
        if (strcmp(line, "end_header\n") == 0 || strcmp(line, "end_header\r\n") == 0) {
            header = true;
            break;
        }
    }
    if (!header || count < 0) {
        fclose(f);
        return MVS_ERR_FORMAT;
    }
    *n = count;
    if (xyz || rgb) {
        if (cap < count) {
            fclose(f);
            return MVS_ERR_ARG;
        }
        uint8_t rec[15];
        for (int64_t i = 0; i < count; ++i) {
            if (fread(rec, 1, 15, f) != 15) {
                fclose(f);
                return MVS_ERR_FORMAT;
            }
            if (xyz) memcpy(xyz + 3 * i, rec, 12);
            if (rgb) memcpy(rgb + 3 * i, rec + 12, 3);
        }
    }
    fclose(f);
    return MVS_OK;
}

int mvs_save_depth(const char *path, int32_t w, int32_t h, const float *depth) {
    if (!path || w < 0 || h < 0 || (!depth && (int64_t)w * h > 0)) return MVS_ERR_ARG;
    std::string tmp = std::string(path);
    // path.with_extension("tmp") (dense.rs:482): replace the extension
    const size_t slash = tmp.find_last_of('/');
    const size_t dot = tmp.find_last_of('.');
    if (dot != std::string::npos && (slash == std::string::npos || dot > slash))
        tmp = tmp.substr(0, dot);
    tmp += ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return MVS_ERR_IO;
    const uint32_t hdr[2] = {(uint32_t)w, (uint32_t)h};  // little-endian hosts only
    bool ok = fwrite(hdr, 4, 2, f) == 2;
    ok = ok && fwrite(depth, 4, (size_t)w * h, f) == (size_t)w * h;
    ok = (fclose(f) == 0) && ok;
    if (!ok || rename(tmp.c_str(), path) != 0) return MVS_ERR_IO;
    return MVS_OK;
}

int mvs_load_depth(const char *path, int32_t *w, int32_t *h, float *depth, int64_t cap) {
    if (!path || !w || !h) return MVS_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return MVS_ERR_IO;
    uint32_t hdr[2];
    if (fread(hdr, 4, 2, f) != 2) {
        fclose(f);
        return MVS_ERR_FORMAT;
    }
    fseek(f, 0, SEEK_END);
    const long len = ftell(f);
    const int64_t n = (int64_t)hdr[0] * hdr[1];
    // load_depth (dense.rs:498): the length must be exactly 8 + 4 w h
    if (len != 8 + 4 * n) {
        fclose(f);
        return MVS_ERR_FORMAT;
    }
    *w = (int32_t)hdr[0];
    *h = (int32_t)hdr[1];
    if (depth) {
        if (cap < n) {
            fclose(f);
            return MVS_ERR_ARG;
        }
        fseek(f, 8, SEEK_SET);
        if (fread(depth, 4, (size_t)n, f) != (size_t)n) {
            fclose(f);
            return MVS_ERR_IO;
        }
    }
    fclose(f);
    return MVS_OK;
}

// ---- surface model (recon/dsm.rs) ----

namespace {

// dsm.rs:142-148: the grid from the percentile extent. Rust's `as usize`
// makes a NaN edge count 0; an edge count past 2^31 is refused by the caller.
void dsm_grid(const float ext[4], double voxel_m, int64_t budget, DsmGrid &g) {
    const double xa = ext[0], xb = ext[1], ya = ext[2], yb = ext[3];
    const double area = (xb - xa) * (yb - ya);
    g.cell = std::max(voxel_m * 1.5, std::sqrt(area / (double)budget));
    g.x0 = xa;
    g.y0 = ya;
    auto edge = [](double e) -> int64_t {
        if (std::isnan(e)) return 1;
        if (e > 2147483647.0) return INT64_MAX;
        return (int64_t)e + 1;
    };
    g.nx = edge(std::ceil((xb - xa) / g.cell));
    g.ny = edge(std::ceil((yb - ya) / g.cell));
}

}  // namespace

int mvs_build_dsm(mvs_ctx *c, const float *xyz, int64_t n, double voxel_m, int32_t profile,
                  mvs_dsm_info *info, float *h, int64_t cap) {
    if (!c) return MVS_ERR_ARG;
    static const int64_t budget[3] = {3000000, 8000000, 20000000};  // dsm.rs:121-128
    c->dsm_cells = 0;
    if (!info || profile < 0 || profile > 2 || !(voxel_m > 0.0) || n < 0 || (n > 0 && !xyz))
        return fail(c, MVS_ERR_ARG, "bad surface model arguments");
    if (n < 100) return fail(c, MVS_ERR_ARG, "the point cloud is too small to build a surface");
    if (n >= (1ll << 31)) return fail(c, MVS_ERR_ARG, "surface model of 2^31 points or more");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, c->dsm_xyz.ensure((size_t)n * 12));
    HIP_TRY(c, c->dsm_sel.ensure(kDsmSelBytes));
    HIP_TRY(c, hipMemcpyAsync(c->dsm_xyz.p, xyz, (size_t)n * 12, hipMemcpyHostToDevice, c->st));
    // dsm.rs:142: v[((len - 1) as f64 * f) as usize] of the sorted coordinates
    const uint32_t rlo = (uint32_t)((double)(n - 1) * 0.005);
    const uint32_t rhi = (uint32_t)((double)(n - 1) * 0.995);
    float ext[4];
    if (dsm_extent(c->dsm_xyz.as<float>(), n, rlo, rhi, c->dsm_sel.p, c->st, ext))
        return hip_fail(c, hipGetLastError(), "surface model extent");
    DsmGrid g;
    dsm_grid(ext, voxel_m, budget[profile], g);
    info->x0 = g.x0;
    info->y0 = g.y0;
    info->cell = g.cell;
    info->nx = g.nx;
    info->ny = g.ny;
    info->zmax = NAN;
    info->reserved = 0;
    if (g.nx > (1ll << 31) || g.ny > (1ll << 31) || g.nx * g.ny >= (1ll << 31))
        return fail(c, MVS_ERR_ARG, "surface grid too large (2^31 cells)");
    const int64_t nc = g.nx * g.ny;
    const size_t cb = (size_t)nc * 4;
    HIP_TRY(c, c->dsm_h.ensure(cb));
    DevBuf *bufs[] = {&c->dsm_a, &c->dsm_b, &c->dsm_c, &c->dsm_d, &c->dsm_snap, &c->dsm_parent,
                      &c->dsm_la, &c->dsm_lb};
    for (DevBuf *b : bufs) HIP_TRY(c, b->ensure(cb));
    HIP_TRY(c, c->dsm_ctr.ensure(16));
    DsmWork w{c->dsm_a.as<uint32_t>(),  c->dsm_b.as<uint32_t>(),  c->dsm_c.as<uint32_t>(),
              c->dsm_d.as<uint32_t>(),  c->dsm_snap.as<float>(),  c->dsm_parent.as<int32_t>(),
              c->dsm_la.as<int32_t>(),  c->dsm_lb.as<int32_t>(),  c->dsm_ctr.as<uint32_t>()};
    float zmax = 0.0f;
    if (dsm_build(g, c->dsm_xyz.as<float>(), n, c->dsm_h.as<float>(), w, c->st, &zmax))
        return hip_fail(c, hipGetLastError(), "surface model kernels");
    info->zmax = zmax;
    c->dsm_cells = nc;
    if (!h) return MVS_OK;
    return mvs_dsm_download(c, h, cap);
}

int mvs_dsm_download(mvs_ctx *c, float *h, int64_t cap) {
    if (!c) return MVS_ERR_ARG;
    if (!h || c->dsm_cells == 0) return fail(c, MVS_ERR_ARG, "no surface model built");
    if (cap < c->dsm_cells) return fail(c, MVS_ERR_ARG, "output array too small");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(h, c->dsm_h.p, (size_t)c->dsm_cells * 4, hipMemcpyDeviceToHost,
                              c->st));
    HIP_TRY(c, hipStreamSynchronize(c->st));
    return MVS_OK;
}

int mvs_save_dsm(const char *path, const mvs_dsm_info *info, const float *h) {
    if (!path || !info || info->nx < 0 || info->ny < 0 || info->nx > (1ll << 31) ||
        info->ny > (1ll << 31) || (info->nx * info->ny > 0 && !h))
        return MVS_ERR_ARG;
    // Dsm::save (dsm.rs:27-46): "DSM1", x0 y0 cell f64, nx ny u64, zmax f32,
    // heights f32, all little-endian; tmp + rename
    std::string tmp = std::string(path);
    const size_t slash = tmp.find_last_of('/');
    const size_t dot = tmp.find_last_of('.');
    if (dot != std::string::npos && (slash == std::string::npos || dot > slash))
        tmp = tmp.substr(0, dot);
    tmp += ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return MVS_ERR_IO;
    const double d3[3] = {info->x0, info->y0, info->cell};
    const uint64_t n2[2] = {(uint64_t)info->nx, (uint64_t)info->ny};
    const size_t nc = (size_t)(info->nx * info->ny);
    bool ok = fwrite("DSM1", 1, 4, f) == 4;
    ok = ok && fwrite(d3, 8, 3, f) == 3;
    ok = ok && fwrite(n2, 8, 2, f) == 2;
    ok = ok && fwrite(&info->zmax, 4, 1, f) == 1;
    ok = ok && (nc == 0 || fwrite(h, 4, nc, f) == nc);
    ok = (fclose(f) == 0) && ok;
    if (!ok || rename(tmp.c_str(), path) != 0) return MVS_ERR_IO;
    return MVS_OK;
}

int mvs_load_dsm(const char *path, mvs_dsm_info *info, float *h, int64_t cap) {
    if (!path || !info) return MVS_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return MVS_ERR_IO;
    unsigned char hdr[48];
    const size_t got = fread(hdr, 1, 48, f);
    fseek(f, 0, SEEK_END);
    const long len = ftell(f);
    // Dsm::load (dsm.rs:48-75): "damaged" below 48 bytes or without the tag
    if (got < 48 || memcmp(hdr, "DSM1", 4) != 0) {
        fclose(f);
        return MVS_ERR_FORMAT;
    }
    double d3[3];
    uint64_t n2[2];
    memcpy(d3, hdr + 4, 24);
    memcpy(n2, hdr + 28, 16);
    memcpy(&info->zmax, hdr + 44, 4);
    info->x0 = d3[0];
    info->y0 = d3[1];
    info->cell = d3[2];
    info->nx = (int64_t)n2[0];
    info->ny = (int64_t)n2[1];
    info->reserved = 0;
    // the header's grid against the file length, without overflow
    const uint64_t body = len >= 48 ? (uint64_t)len - 48 : ~0ull;
    const uint64_t nc = n2[0] * n2[1];  // both at most 2^31: no wrap
    if (n2[0] > (1ull << 31) || n2[1] > (1ull << 31) || body == ~0ull || body % 4 != 0 ||
        nc != body / 4) {
        fclose(f);
        return MVS_ERR_FORMAT;  // "truncated surface model file"
    }
    if (h) {
        if (cap < (int64_t)nc) {
            fclose(f);
            return MVS_ERR_ARG;
        }
        fseek(f, 48, SEEK_SET);
        const bool ok = fread(h, 4, (size_t)nc, f) == (size_t)nc;
        fclose(f);
        return ok ? MVS_OK : MVS_ERR_IO;
    }
    fclose(f);
    return MVS_OK;
}

int mvs_dsm_height(const mvs_dsm_info *d, const float *h, double x, double y, float *out) {
    if (!d || !h || !out || d->nx < 1 || d->ny < 1) return MVS_ERR_ARG;
    // Dsm::height (dsm.rs:82-119): bilinear between cell centres over the
    // finite samples, the nearest valid one on an empty sample, None outside
    const double fx = (x - d->x0) / d->cell - 0.5;
    const double fy = (y - d->y0) / d->cell - 0.5;
    if (fx < -0.5 || fy < -0.5 || fx > (double)d->nx - 0.5 || fy > (double)d->ny - 0.5) return 0;
    const int64_t i0 = std::min((int64_t)std::max(std::floor(fx), 0.0), d->nx - 1);
    const int64_t j0 = std::min((int64_t)std::max(std::floor(fy), 0.0), d->ny - 1);
    const int64_t i1 = std::min(i0 + 1, d->nx - 1), j1 = std::min(j0 + 1, d->ny - 1);
    const float ax = (float)std::min(std::max(fx - (double)i0, 0.0), 1.0);
    const float ay = (float)std::min(std::max(fy - (double)j0, 0.0), 1.0);
    const float v[4] = {h[j0 * d->nx + i0], h[j0 * d->nx + i1], h[j1 * d->nx + i0],
                        h[j1 * d->nx + i1]};
    const float wt[4] = {(1.0f - ax) * (1.0f - ay), ax * (1.0f - ay), (1.0f - ax) * ay, ax * ay};
    float s = 0.0f, ws = 0.0f;
    for (int k = 0; k < 4; ++k)
        if (std::isfinite(v[k])) {
            s += v[k] * wt[k];
            ws += wt[k];
        }
    if (ws > 1e-6f) {
        *out = s / ws;
        return 1;
    }
    for (int k = 0; k < 4; ++k)
        if (std::isfinite(v[k])) {
            *out = v[k];
            return 1;
        }
    return 0;
}

int mvs_prepared_size(int32_t w, int32_t h, int32_t max_side, int32_t *out_w, int32_t *out_h,
                      double *scale) {
    if (w < 1 || h < 1 || max_side < 1 || !out_w || !out_h) return MVS_ERR_ARG;
    PrepSizes z;
    prep_sizes(w, h, max_side, z);
    *out_w = z.ow;
    *out_h = z.oh;
    if (scale) *scale = z.scale;
    return MVS_OK;
}

int mvs_prepare_view(mvs_ctx *c, const uint8_t *rgb, int32_t w, int32_t h, int32_t stride,
                     int32_t max_side, float *gray, uint8_t *rgb_out) {
    if (!c) return MVS_ERR_ARG;
    if (!rgb || !gray || !rgb_out || w < 1 || h < 1 || stride < 3 * w || max_side < 1)
        return fail(c, MVS_ERR_ARG, "bad view preparation arguments");
    if ((int64_t)w * h > (1ll << 31) / 3)
        return fail(c, MVS_ERR_ARG, "photo too large (3 w h must stay below 2^31)");
    PrepSizes z;
    prep_sizes(w, h, max_side, z);
    const size_t in_bytes = (size_t)stride * (h - 1) + (size_t)w * 3;
    const size_t n = (size_t)z.ow * z.oh;
    HIP_TRY(c, c->up_u8.ensure(in_bytes));
    HIP_TRY(c, c->prep_box.ensure(z.f >= 2 ? (size_t)z.bw * z.bh * 3 : 1));
    HIP_TRY(c, c->prep_rgb.ensure(n * 3));
    HIP_TRY(c, c->prep_grey.ensure(n * sizeof(float)));
    HIP_TRY(c, hipMemcpyAsync(c->up_u8.p, rgb, in_bytes, hipMemcpyHostToDevice, c->st));
    launch_prepare_view(c->up_u8.as<uint8_t>(), w, h, stride, z, c->prep_box.as<uint8_t>(),
                        c->prep_rgb.as<uint8_t>(), c->prep_grey.as<float>(), c->st);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(gray, c->prep_grey.p, n * sizeof(float), hipMemcpyDeviceToHost,
                              c->st));
    HIP_TRY(c, hipMemcpyAsync(rgb_out, c->prep_rgb.p, n * 3, hipMemcpyDeviceToHost, c->st));
    HIP_TRY(c, hipStreamSynchronize(c->st));
    return MVS_OK;
}

int mvs_save_view(const char *path, int32_t w, int32_t h, const float *gray, const uint8_t *rgb) {
    if (!path || w < 0 || h < 0 || ((int64_t)w * h > 0 && (!gray || !rgb))) return MVS_ERR_ARG;
    const size_t n = (size_t)w * h;
    // prepare_views (dense.rs:101-113): header, grey, RGB, then tmp + rename
    std::string tmp = std::string(path);
    const size_t slash = tmp.find_last_of('/');
    const size_t dot = tmp.find_last_of('.');
    if (dot != std::string::npos && (slash == std::string::npos || dot > slash))
        tmp = tmp.substr(0, dot);
    tmp += ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return MVS_ERR_IO;
    const uint32_t hdr[2] = {(uint32_t)w, (uint32_t)h};  // little-endian hosts only
    bool ok = fwrite(hdr, 4, 2, f) == 2;
    ok = ok && fwrite(gray, 4, n, f) == n;
    ok = ok && fwrite(rgb, 1, 3 * n, f) == 3 * n;
    ok = (fclose(f) == 0) && ok;
    if (!ok || rename(tmp.c_str(), path) != 0) return MVS_ERR_IO;
    return MVS_OK;
}

int mvs_load_view(const char *path, int32_t *w, int32_t *h, float *gray, uint8_t *rgb,
                  int64_t cap) {
    if (!path || !w || !h) return MVS_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return MVS_ERR_IO;
    uint32_t hdr[2];
    if (fread(hdr, 4, 2, f) != 2) {
        fclose(f);
        return MVS_ERR_FORMAT;
    }
    fseek(f, 0, SEEK_END);
    const long len = ftell(f);
    const int64_t n = (int64_t)hdr[0] * hdr[1];
    // load_view (dense.rs:130-132): "damaged cache file" unless 8 + 7 w h bytes
    if (len != 8 + 7 * n) {
        fclose(f);
        return MVS_ERR_FORMAT;
    }
    *w = (int32_t)hdr[0];
    *h = (int32_t)hdr[1];
    if (gray || rgb) {
        if (cap < n) {
            fclose(f);
            return MVS_ERR_ARG;
        }
        bool ok = true;
        if (gray) {
            fseek(f, 8, SEEK_SET);
            ok = fread(gray, 4, (size_t)n, f) == (size_t)n;
        }
        if (ok && rgb) {
            fseek(f, 8 + 4 * n, SEEK_SET);
            ok = fread(rgb, 1, 3 * (size_t)n, f) == 3 * (size_t)n;
        }
        if (!ok) {
            fclose(f);
            return MVS_ERR_IO;
        }
    }
    fclose(f);
    return MVS_OK;
}

int mvs_view_intrinsics(double f, double cx, double cy, int32_t full_width, int32_t w,
                        double out[3]) {
    if (!out || full_width <= 0 || w <= 0) return MVS_ERR_ARG;
    // load_view (dense.rs:139-145): K scaled by s = w / full.width; k1, k2 kept
    const double s = (double)w / (double)full_width;
    out[0] = f * s;
    out[1] = cx * s;
    out[2] = cy * s;
    return MVS_OK;
}

}  // extern "C"
