// C-ABI layer of the dense-MVS depth engine: contexts, device buffers, the
// resident survey, and the host logic of the reference's depth stage
// (crates/stitch-core/src/recon/dense.rs) around the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <map>
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
    double f = 0, cx = 0, cy = 0, k1 = 0, k2 = 0;
    double R[9] = {0}, C[3] = {0};
    bool set = false;
};

// Outputs of one reference view kept on the device.
struct MapDev {
    int32_t w = 0, h = 0;
    DevBuf depth, normal, cost;
    bool valid = false;
};

typedef std::tuple<double, double, double, double, double, int32_t, int32_t> RayKey;

struct TimedLaunch {
    hipEvent_t a, b;
    int kind;  // 0 sweep, 1 init+eval
};

}  // namespace

struct mvs_ctx {
    int32_t device = 0;
    hipStream_t st = nullptr;
    std::string err;
    // work buffers of the current reference view
    DevBuf hyp, cost, textured, list, counts, srcs, evals, ok_count;
    DevBuf up_u8;  // staging of u8 uploads
    // host-view calls
    std::vector<ViewDev> tmp_views;
    MapDev tmp_map;
    DevBuf px, py, hyp_in, cost_out, nbr_cams, agree;
    std::map<RayKey, DevBuf *> rays;
    // timing
    bool timing = false;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> free_events;
    mvs_stats stats{};
};

struct mvs_survey {
    mvs_ctx *ctx = nullptr;
    std::vector<ViewDev> views;
    std::vector<MapDev> maps;
    std::vector<DevBuf> src_tables;  // per reference view
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

#define HIP_TRY(c, expr)                                   \
    do {                                                   \
        hipError_t _e = (expr);                            \
        if (_e != hipSuccess) return hip_fail(c, _e, #expr); \
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

int sync_ctx(mvs_ctx *c) {
    HIP_TRY(c, hipStreamSynchronize(c->st));
    collect_timing(c);
    unsigned long long ev[2] = {0, 0};
    if (c->evals.p) {
        HIP_TRY(c, hipMemcpy(ev, c->evals.p, sizeof ev, hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemset(c->evals.p, 0, sizeof ev));
    }
    c->stats.init_evals += (int64_t)ev[0];
    c->stats.sweep_evals += (int64_t)ev[1];
    c->stats.evals += (int64_t)(ev[0] + ev[1]);
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
        return fail(c, MVS_ERR_ARG, "window taps must be 5 or 7 (got " + std::to_string(p->taps) + ")");
    if (p->stride < 1 || p->stride > 8) return fail(c, MVS_ERR_ARG, "tap stride must be 1..8");
    if (p->iterations < 0 || p->iterations > 64) return fail(c, MVS_ERR_ARG, "iterations out of range");
    return MVS_OK;
}

// Upload a host view into a ViewDev (grey converted to f32 on the device).
int upload_view(mvs_ctx *c, ViewDev &dv, const mvs_view *v) {
    int rc = check_view(c, v, "view");
    if (rc) return rc;
    const size_t n = (size_t)v->w * v->h;
    HIP_TRY(c, dv.gray.ensure(n * sizeof(float)));
    if (v->format == MVS_GRAY_U8) {
        const size_t bytes = (size_t)v->stride * v->h;
        HIP_TRY(c, c->up_u8.ensure(bytes));
        HIP_TRY(c, hipMemcpyAsync(c->up_u8.p, v->pixels, bytes, hipMemcpyHostToDevice, c->st));
        launch_gray_from_u8(c->up_u8.as<uint8_t>(), v->w, v->h, v->stride, dv.gray.as<float>(), c->st);
        HIP_TRY(c, hipGetLastError());
        // staging buffer is reused by the next upload: wait for the conversion
        HIP_TRY(c, hipStreamSynchronize(c->st));
    } else {
        HIP_TRY(c, hipMemcpy2DAsync(dv.gray.p, (size_t)v->w * sizeof(float), v->pixels,
                                    (size_t)v->stride * sizeof(float), (size_t)v->w * sizeof(float),
                                    v->h, hipMemcpyHostToDevice, c->st));
        HIP_TRY(c, hipStreamSynchronize(c->st));
    }
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

void split(double v, float &hi, float &lo) {
    hi = (float)v;
    lo = (float)(v - (double)hi);
}

SrcCam make_src(const ViewDev &ref, const ViewDev &s) {
    SrcCam c;
    memset(&c, 0, sizeof c);
    relative_pose(ref.R, ref.C, s.R, s.C, c.R, c.t);
    split(s.f, c.f, c.f_lo);
    split(s.cx, c.cx, c.cx_lo);
    split(s.cy, c.cy, c.cy_lo);
    split(s.k1, c.k1, c.k1_lo);
    split(s.k2, c.k2, c.k2_lo);
    c.gray = s.gray.as<float>();
    c.w = s.w;
    c.h = s.h;
    c.pinhole_exact = (s.k1 == 0.0 && s.k2 == 0.0 && c.f_lo == 0.0f && c.cx_lo == 0.0f &&
                       c.cy_lo == 0.0f)
                          ? 1
                          : 0;
    return c;
}

// Everything of one reference view's PatchMatch, enqueued on the context
// stream. first_step = -1 initialises; the half-sweeps run from max(first, 0)
// for n_steps. Uses the context work buffers; `table` receives the source table.
int run_map(mvs_ctx *c, const ViewDev &ref, const std::vector<const ViewDev *> &srcs,
            double range0, double range1, const mvs_params *prm, int32_t first_step,
            int32_t n_steps, DevBuf &table, std::vector<SrcCam> &host_table,
            const volatile int32_t *cancel, RefJob *job_out) {
    const int32_t K = (int32_t)srcs.size();
    if (K < 1) return fail(c, MVS_NO_NEIGHBOURS, "no neighbours");
    if (lanes_for(K) < 0) return fail(c, MVS_ERR_ARG, "at most 16 source views per reference view");
    const int32_t w = ref.w, h = ref.h;
    const int64_t n = (int64_t)w * h;
    const int32_t half = (prm->taps / 2) * prm->stride;
    if (w <= 2 * half || h <= 2 * half) return fail(c, MVS_ERR_ARG, "image smaller than the window");
    const int64_t cap = (n + 1) / 2;
    HIP_TRY(c, c->hyp.ensure(n * sizeof(float4)));
    HIP_TRY(c, c->cost.ensure(n * sizeof(float)));
    HIP_TRY(c, c->textured.ensure(n));
    HIP_TRY(c, c->list.ensure(2 * cap * sizeof(int32_t)));
    HIP_TRY(c, c->counts.ensure(4 * sizeof(int32_t)));
    if (!c->evals.p) {
        HIP_TRY(c, c->evals.ensure(2 * sizeof(unsigned long long)));
        HIP_TRY(c, hipMemset(c->evals.p, 0, 2 * sizeof(unsigned long long)));
    }
    host_table.clear();
    bool pinhole = true;
    for (const ViewDev *s : srcs) {
        host_table.push_back(make_src(ref, *s));
        pinhole = pinhole && host_table.back().pinhole_exact;
    }
    HIP_TRY(c, table.ensure(sizeof(SrcCam) * (size_t)K));
    HIP_TRY(c, hipMemcpyAsync(table.p, host_table.data(), sizeof(SrcCam) * (size_t)K,
                              hipMemcpyHostToDevice, c->st));
    const float2 *rays = nullptr;
    int rc = get_rays(c, ref, &rays);
    if (rc) return rc;

    RefJob job;
    memset(&job, 0, sizeof job);
    job.gray = ref.gray.as<float>();
    job.rays = rays;
    job.hyp = c->hyp.as<float4>();
    job.cost = c->cost.as<float>();
    job.textured = c->textured.as<uint8_t>();
    job.list = c->list.as<int32_t>();
    job.counts = c->counts.as<int32_t>();
    job.srcs = table.as<SrcCam>();
    job.w = w;
    job.h = h;
    job.nsrc = K;
    job.list_cap = (int32_t)cap;
    // dense.rs:362-363
    job.dmin = (float)(range0 > 0.5 ? range0 : 0.5);
    job.dmax = (float)range1;
    job.imin = 1.0f / job.dmax;
    job.imax = 1.0f / job.dmin;
    // dense.rs:367-370: R_ref * (0, 0, 1)
    job.up0 = (float)ref.R[2];
    job.up1 = (float)ref.R[5];
    job.up2 = (float)ref.R[8];
    job.texture_min = prm->texture_min;
    job.cost_max = prm->cost_max;
    job.stride = prm->stride;
    job.half = half;
    job.pinhole = pinhole ? 1 : 0;
    if (job_out) *job_out = job;

    unsigned long long *evals = c->evals.as<unsigned long long>();
    HIP_TRY(c, hipMemsetAsync(c->counts.p, 0, 4 * sizeof(int32_t), c->st));
    {
        hipEvent_t a = nullptr, b = nullptr;
        if (c->timing) {
            a = take_event(c);
            b = take_event(c);
            HIP_TRY(c, hipEventRecord(a, c->st));
        }
        launch_init(job, prm->taps, first_step < 0 ? 1 : 0, c->st);
        HIP_TRY(c, hipGetLastError());
        if (first_step < 0) {
            if (launch_eval_current(job, prm->taps, evals + 0, c->st))
                return fail(c, MVS_ERR_ARG, "unsupported window/sources combination");
            HIP_TRY(c, hipGetLastError());
        }
        if (c->timing) {
            HIP_TRY(c, hipEventRecord(b, c->st));
            c->pending.push_back({a, b, 1});
        }
    }
    const int32_t s0 = first_step < 0 ? 0 : first_step;
    for (int32_t s = s0; s < s0 + n_steps; ++s) {
        if (cancel && *cancel) return fail(c, MVS_ERR_CANCELLED, "cancelled");
        hipEvent_t a = nullptr, b = nullptr;
        if (c->timing) {
            a = take_event(c);
            b = take_event(c);
            HIP_TRY(c, hipEventRecord(a, c->st));
        }
        if (launch_sweep(job, prm->taps, s / 2, s % 2, evals + 1, c->st))
            return fail(c, MVS_ERR_ARG, "unsupported window/sources combination");
        HIP_TRY(c, hipGetLastError());
        if (c->timing) {
            HIP_TRY(c, hipEventRecord(b, c->st));
            c->pending.push_back({a, b, 0});
        }
    }
    c->stats.pixels += n;
    return MVS_OK;
}

int ensure_map(mvs_ctx *c, MapDev &m, int32_t w, int32_t h) {
    const size_t n = (size_t)w * h;
    HIP_TRY(c, m.depth.ensure(n * sizeof(float)));
    HIP_TRY(c, m.normal.ensure(3 * n * sizeof(float)));
    HIP_TRY(c, m.cost.ensure(n * sizeof(float)));
    m.w = w;
    m.h = h;
    return MVS_OK;
}

// Threshold (dense.rs:468-477) into a map and keep the final costs.
int finish_map(mvs_ctx *c, const RefJob &job, MapDev &m) {
    int rc = ensure_map(c, m, job.w, job.h);
    if (rc) return rc;
    launch_threshold(job, m.depth.as<float>(), m.normal.as<float>(), c->st);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(m.cost.p, job.cost, sizeof(float) * (size_t)job.w * job.h,
                              hipMemcpyDeviceToDevice, c->st));
    m.valid = true;
    return MVS_OK;
}

int download_map(mvs_ctx *c, const MapDev &m, float *depth, float *normal, float *cost) {
    const size_t n = (size_t)m.w * m.h;
    if (depth) HIP_TRY(c, hipMemcpyAsync(depth, m.depth.p, n * sizeof(float), hipMemcpyDeviceToHost, c->st));
    if (normal)
        HIP_TRY(c, hipMemcpyAsync(normal, m.normal.p, 3 * n * sizeof(float), hipMemcpyDeviceToHost, c->st));
    if (cost) HIP_TRY(c, hipMemcpyAsync(cost, m.cost.p, n * sizeof(float), hipMemcpyDeviceToHost, c->st));
    return sync_ctx(c);
}

int upload_host_views(mvs_ctx *c, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                      std::vector<const ViewDev *> &out) {
    if (nsrc < 0) return fail(c, MVS_ERR_ARG, "negative source count");
    if (nsrc > 0 && !srcs) return fail(c, MVS_ERR_ARG, "null source views");
    if ((int32_t)c->tmp_views.size() < nsrc + 1) c->tmp_views.resize((size_t)nsrc + 1);
    int rc = upload_view(c, c->tmp_views[0], ref);
    if (rc) return rc;
    out.clear();
    for (int32_t i = 0; i < nsrc; ++i) {
        rc = upload_view(c, c->tmp_views[(size_t)i + 1], &srcs[i]);
        if (rc) return rc;
        out.push_back(&c->tmp_views[(size_t)i + 1]);
    }
    return MVS_OK;
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
        return report(MVS_ERR_DEVICE, std::string("built for gfx950, device is ") + prop.gcnArchName);
    e = hipSetDevice(device);
    if (e != hipSuccess) return report(MVS_ERR_DEVICE, hipGetErrorString(e));
    mvs_ctx *c = new mvs_ctx();
    c->device = device;
    e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return report(MVS_ERR_DEVICE, hipGetErrorString(e));
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
    for (auto &kv : c->rays) {
        kv.second->release();
        delete kv.second;
    }
    DevBuf *bufs[] = {&c->hyp, &c->cost, &c->textured, &c->list, &c->counts, &c->srcs,
                      &c->evals, &c->ok_count, &c->up_u8, &c->px, &c->py, &c->hyp_in,
                      &c->cost_out, &c->nbr_cams, &c->agree, &c->tmp_map.depth,
                      &c->tmp_map.normal, &c->tmp_map.cost};
    for (DevBuf *b : bufs) b->release();
    for (auto &v : c->tmp_views) v.gray.release();
    (void)hipStreamDestroy(c->st);
    delete c;
}

const char *mvs_last_error(const mvs_ctx *c) { return c ? c->err.c_str() : "null context"; }

int mvs_synchronize(mvs_ctx *c) {
    if (!c) return MVS_ERR_ARG;
    return sync_ctx(c);
}

int mvs_depth_map(mvs_ctx *c, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                  double range0, double range1, const mvs_params *prm, float *out_depth,
                  float *out_normal, float *out_cost, const volatile int32_t *cancel) {
    if (!c) return MVS_ERR_ARG;
    int rc = check_params(c, prm);
    if (rc) return rc;
    if (nsrc < 1) return fail(c, MVS_NO_NEIGHBOURS, "no neighbours");
    if (!out_depth) return fail(c, MVS_ERR_ARG, "null depth output");
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<const ViewDev *> sv;
    rc = upload_host_views(c, ref, srcs, nsrc, sv);
    if (rc) return rc;
    std::vector<SrcCam> host_table;
    RefJob job;
    rc = run_map(c, c->tmp_views[0], sv, range0, range1, prm, -1, 2 * prm->iterations, c->srcs,
                 host_table, cancel, &job);
    if (rc) {
        (void)hipStreamSynchronize(c->st);
        return rc;
    }
    rc = finish_map(c, job, c->tmp_map);
    if (rc) return rc;
    return download_map(c, c->tmp_map, out_depth, out_normal, out_cost);
}

int mvs_ncc_cost(mvs_ctx *c, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                 double range0, double range1, const mvs_params *prm, int64_t n,
                 const int32_t *px, const int32_t *py, const float *hyp4, float *out_cost) {
    if (!c) return MVS_ERR_ARG;
    int rc = check_params(c, prm);
    if (rc) return rc;
    if (nsrc < 1) return fail(c, MVS_NO_NEIGHBOURS, "no neighbours");
    if (n < 0 || (n > 0 && (!px || !py || !hyp4 || !out_cost)))
        return fail(c, MVS_ERR_ARG, "bad cost list");
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<const ViewDev *> sv;
    rc = upload_host_views(c, ref, srcs, nsrc, sv);
    if (rc) return rc;
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
    // Build the job (rays, sources) without running PatchMatch.
    std::vector<SrcCam> host_table;
    RefJob job;
    rc = run_map(c, rf, sv, range0, range1, prm, 0, 0, c->srcs, host_table, nullptr, &job);
    if (rc) return rc;
    if (launch_cost_list(job, prm->taps, n, c->px.as<int32_t>(), c->py.as<int32_t>(),
                         c->hyp_in.as<float4>(), c->cost_out.as<float>(), c->st))
        return fail(c, MVS_ERR_ARG, "unsupported window/sources combination");
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(out_cost, c->cost_out.p, (size_t)n * 4, hipMemcpyDeviceToHost, c->st));
    return sync_ctx(c);
}

int mvs_patchmatch_steps(mvs_ctx *c, const mvs_view *ref, const mvs_view *srcs, int32_t nsrc,
                         double range0, double range1, const mvs_params *prm,
                         int32_t first_step, int32_t n_steps, float *state, uint8_t *textured,
                         int64_t *evals) {
    if (!c) return MVS_ERR_ARG;
    int rc = check_params(c, prm);
    if (rc) return rc;
    if (nsrc < 1) return fail(c, MVS_NO_NEIGHBOURS, "no neighbours");
    if (!state || n_steps < 0 || first_step < -1) return fail(c, MVS_ERR_ARG, "bad state or steps");
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<const ViewDev *> sv;
    rc = upload_host_views(c, ref, srcs, nsrc, sv);
    if (rc) return rc;
    const ViewDev &rf = c->tmp_views[0];
    const size_t n = (size_t)rf.w * rf.h;
    rc = sync_ctx(c);  // flush earlier evals into stats
    if (rc) return rc;
    const int64_t evals0 = c->stats.evals;
    std::vector<float> hyp4, cst;
    if (first_step >= 0) {
        HIP_TRY(c, c->hyp.ensure(n * sizeof(float4)));
        HIP_TRY(c, c->cost.ensure(n * sizeof(float)));
        hyp4.resize(4 * n);
        cst.resize(n);
        for (size_t i = 0; i < n; ++i) {
            for (int k = 0; k < 4; ++k) hyp4[4 * i + k] = state[5 * i + k];
            cst[i] = state[5 * i + 4];
        }
        HIP_TRY(c, hipMemcpyAsync(c->hyp.p, hyp4.data(), n * 16, hipMemcpyHostToDevice, c->st));
        HIP_TRY(c, hipMemcpyAsync(c->cost.p, cst.data(), n * 4, hipMemcpyHostToDevice, c->st));
    }
    std::vector<SrcCam> host_table;
    RefJob job;
    rc = run_map(c, rf, sv, range0, range1, prm, first_step, n_steps, c->srcs, host_table,
                 nullptr, &job);
    if (rc) {
        (void)hipStreamSynchronize(c->st);
        return rc;
    }
    hyp4.resize(4 * n);
    cst.resize(n);
    HIP_TRY(c, hipMemcpyAsync(hyp4.data(), c->hyp.p, n * 16, hipMemcpyDeviceToHost, c->st));
    HIP_TRY(c, hipMemcpyAsync(cst.data(), c->cost.p, n * 4, hipMemcpyDeviceToHost, c->st));
    if (textured)
        HIP_TRY(c, hipMemcpyAsync(textured, c->textured.p, n, hipMemcpyDeviceToHost, c->st));
    rc = sync_ctx(c);
    if (rc) return rc;
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
    s->src_tables.resize((size_t)nviews);
    *out = s;
    return MVS_OK;
}

void mvs_survey_destroy(mvs_survey *s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->st);
    for (auto &v : s->views) v.gray.release();
    for (auto &m : s->maps) {
        m.depth.release();
        m.normal.release();
        m.cost.release();
    }
    for (auto &t : s->src_tables) t.release();
    delete s;
}

int mvs_survey_set_view(mvs_survey *s, int32_t i, const mvs_view *v) {
    if (!s) return MVS_ERR_ARG;
    mvs_ctx *c = s->ctx;
    if (i < 0 || i >= (int32_t)s->views.size()) return fail(c, MVS_ERR_ARG, "view index out of range");
    HIP_TRY(c, hipSetDevice(c->device));
    return upload_view(c, s->views[(size_t)i], v);
}

static int survey_one(mvs_survey *s, int32_t ref, const int32_t *srcs, int32_t nsrc, double r0,
                      double r1, const mvs_params *prm, const volatile int32_t *cancel,
                      std::vector<SrcCam> &host_table) {
    mvs_ctx *c = s->ctx;
    const int32_t nv = (int32_t)s->views.size();
    if (ref < 0 || ref >= nv || !s->views[(size_t)ref].set)
        return fail(c, MVS_ERR_ARG, "reference view missing");
    if (nsrc < 1) return fail(c, MVS_NO_NEIGHBOURS, "no neighbours");
    std::vector<const ViewDev *> sv;
    for (int32_t k = 0; k < nsrc; ++k) {
        const int32_t j = srcs[k];
        if (j < 0 || j >= nv || !s->views[(size_t)j].set || j == ref)
            return fail(c, MVS_ERR_ARG, "bad source view index");
        sv.push_back(&s->views[(size_t)j]);
    }
    RefJob job;
    int rc = run_map(c, s->views[(size_t)ref], sv, r0, r1, prm, -1, 2 * prm->iterations,
                     s->src_tables[(size_t)ref], host_table, cancel, &job);
    if (rc) return rc;
    return finish_map(c, job, s->maps[(size_t)ref]);
}

int mvs_survey_depth_map(mvs_survey *s, int32_t ref, const int32_t *srcs, int32_t nsrc,
                         double range0, double range1, const mvs_params *prm) {
    if (!s) return MVS_ERR_ARG;
    mvs_ctx *c = s->ctx;
    int rc = check_params(c, prm);
    if (rc) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<SrcCam> host_table;
    rc = survey_one(s, ref, srcs, nsrc, range0, range1, prm, nullptr, host_table);
    if (rc) {
        (void)hipStreamSynchronize(c->st);
        return rc;
    }
    return sync_ctx(c);
}

int mvs_survey_depth_maps(mvs_survey *s, int32_t njobs, const int32_t *refs,
                          const int32_t *src_off, const int32_t *srcs, const double *ranges,
                          const mvs_params *prm, const volatile int32_t *cancel,
                          mvs_progress_fn progress, void *user) {
    if (!s) return MVS_ERR_ARG;
    mvs_ctx *c = s->ctx;
    int rc = check_params(c, prm);
    if (rc) return rc;
    if (njobs < 0 || (njobs > 0 && (!refs || !src_off || !ranges)))
        return fail(c, MVS_ERR_ARG, "bad job arrays");
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<std::vector<SrcCam>> tables((size_t)njobs);
    for (int32_t j = 0; j < njobs; ++j) {
        if (cancel && *cancel) {
            (void)hipStreamSynchronize(c->st);
            return fail(c, MVS_ERR_CANCELLED, "cancelled");
        }
        const int32_t a = src_off[j], b = src_off[j + 1];
        rc = survey_one(s, refs[j], srcs + a, b - a, ranges[2 * j], ranges[2 * j + 1], prm, cancel,
                        tables[(size_t)j]);
        if (rc == MVS_NO_NEIGHBOURS) {
            s->maps[(size_t)refs[j]].valid = false;
            rc = MVS_OK;
        }
        if (rc) {
            (void)hipStreamSynchronize(c->st);
            return rc;
        }
        if (progress && progress(user, "depth", j + 1, njobs)) {
            (void)hipStreamSynchronize(c->st);
            return fail(c, MVS_ERR_CANCELLED, "cancelled");
        }
    }
    return sync_ctx(c);
}

int mvs_survey_download(mvs_survey *s, int32_t view, float *depth, float *normal, float *cost) {
    if (!s) return MVS_ERR_ARG;
    mvs_ctx *c = s->ctx;
    if (view < 0 || view >= (int32_t)s->maps.size() || !s->maps[(size_t)view].valid)
        return fail(c, MVS_ERR_ARG, "no depth map for that view");
    HIP_TRY(c, hipSetDevice(c->device));
    return download_map(c, s->maps[(size_t)view], depth, normal, cost);
}

int mvs_survey_consistency(mvs_survey *s, int32_t ref, const int32_t *nbrs, int32_t nn,
                           int32_t rule, int32_t *agree, int64_t *n_consistent) {
    if (!s) return MVS_ERR_ARG;
    mvs_ctx *c = s->ctx;
    if (rule != MVS_CONS_REFERENCE) return fail(c, MVS_ERR_ARG, "unknown consistency rule");
    const int32_t nv = (int32_t)s->views.size();
    if (ref < 0 || ref >= nv || !s->maps[(size_t)ref].valid)
        return fail(c, MVS_ERR_ARG, "reference view has no depth map");
    if (nn < 0 || (nn > 0 && !nbrs)) return fail(c, MVS_ERR_ARG, "bad neighbour list");
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<NbrCam> cams;
    for (int32_t k = 0; k < nn; ++k) {
        const int32_t j = nbrs[k];
        if (j < 0 || j >= nv || !s->maps[(size_t)j].valid)
            return fail(c, MVS_ERR_ARG, "neighbour without a depth map");
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
    const ViewDev &rv = s->views[(size_t)ref];
    RefCam64 rc64;
    memcpy(rc64.R, rv.R, sizeof rc64.R);
    memcpy(rc64.C, rv.C, sizeof rc64.C);
    rc64.f = rv.f;
    rc64.cx = rv.cx;
    rc64.cy = rv.cy;
    rc64.k1 = rv.k1;
    rc64.k2 = rv.k2;
    const size_t n = (size_t)rv.w * rv.h;
    HIP_TRY(c, c->nbr_cams.ensure(sizeof(NbrCam) * (cams.size() + 1)));
    HIP_TRY(c, c->agree.ensure(n * sizeof(int32_t)));
    HIP_TRY(c, c->ok_count.ensure(sizeof(unsigned long long)));
    if (!cams.empty())
        HIP_TRY(c, hipMemcpyAsync(c->nbr_cams.p, cams.data(), sizeof(NbrCam) * cams.size(),
                                  hipMemcpyHostToDevice, c->st));
    HIP_TRY(c, hipMemsetAsync(c->ok_count.p, 0, sizeof(unsigned long long), c->st));
    // dense.rs:585
    const int32_t need = nn >= 3 ? 2 : 1;
    launch_consistency(rc64, s->maps[(size_t)ref].depth.as<float>(), rv.w, rv.h,
                       c->nbr_cams.as<NbrCam>(), nn, need, c->agree.as<int32_t>(),
                       c->ok_count.as<unsigned long long>(), c->st);
    HIP_TRY(c, hipGetLastError());
    unsigned long long ok = 0;
    if (agree)
        HIP_TRY(c, hipMemcpyAsync(agree, c->agree.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->st));
    HIP_TRY(c, hipMemcpyAsync(&ok, c->ok_count.p, sizeof ok, hipMemcpyDeviceToHost, c->st));
    int rc = sync_ctx(c);
    if (rc) return rc;
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
    int rc = sync_ctx(c);
    if (rc) return rc;
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
    int rc = sync_ctx(c);
    if (rc) return rc;
    *out = c->stats;
    return MVS_OK;
}

int mvs_reset_stats(mvs_ctx *c) {
    if (!c) return MVS_ERR_ARG;
    int rc = sync_ctx(c);
    if (rc) return rc;
    memset(&c->stats, 0, sizeof c->stats);
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

}  // extern "C"
