/*
 * CPU oracle for the dense-MVS depth-map stage — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of crates/stitch-core/src/recon/dense.rs (PatchMatch
 * depth maps, NCC cost, geometric consistency) and of the camera model in
 * crates/stitch-core/src/recon/geometry.rs. Every function cites the lines it
 * follows. Used only by the tests, the smoke check and the CPU-baseline leg of
 * the benchmark; the product library never links or calls it.
 *
 * Build: -O3 -ffp-contract=off, no fast-math (the Rust reference never fuses
 * a*b+c and evaluates left to right; C with FLT_EVAL_METHOD 0 does the same).
 *
 * Parity pins (see tests/test_oracle_*.py): rnd/seeds and random_normal by
 * closed-form known answers; distort/undistort by the reference's own round-trip
 * test (geometry.rs:335-351); gray_at, mean_sd, cost and the sweep by closed
 * forms, invariants and brute force; the full trajectory cannot be pinned
 * against the reference itself (Rust, not buildable here).
 */
#include "dense_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#if defined(__FP_FAST_FMA) && !defined(ORC_ALLOW_FMA)
/* fine: fma availability does not matter, only contraction does */
#endif

/* ------------------------------------------------------------------------ */
/* scalar building blocks                                                   */
/* ------------------------------------------------------------------------ */

/* dense.rs:108 — g = (0.299 R + 0.587 G + 0.114 B) / 255 in f32, left to right. */
float orc_gray_rgb(uint8_t r, uint8_t g, uint8_t b) {
    float v = 0.299f * (float)r + 0.587f * (float)g;
    v = v + 0.114f * (float)b;
    return v / 255.0f;
}

/* A grey u8 pixel is R = G = B (image::to_rgb8 of a luma image). */
void orc_gray_lut(float lut[256]) {
    for (int i = 0; i < 256; ++i) lut[i] = orc_gray_rgb((uint8_t)i, (uint8_t)i, (uint8_t)i);
}

/* dense.rs:273-279 — xorshift64 (13, 7, 17); top 24 bits / 2^24. */
float orc_rnd(uint64_t *s) {
    *s ^= *s << 13;
    *s ^= *s >> 7;
    *s ^= *s << 17;
    return (float)(*s >> 40) / (float)(1u << 24);
}

/* dense.rs:374 */
uint64_t orc_seed_init(uint64_t i) { return (i * 0x9E3779B97F4A7C15ull) | 1ull; }

/* dense.rs:446-448: ((i << 8) ^ (2 it + parity)) * C | 1, wrapping. */
uint64_t orc_seed_refine(uint64_t i, uint64_t it, uint64_t parity) {
    return (((i << 8) ^ (it * 2ull + parity)) * 0x2545F4914F6CDD1Dull) | 1ull;
}

/* dense.rs:281-297 with Some(around) (every caller passes one: :376, :456). */
void orc_random_normal(uint64_t *s, const float ray[2], const float around[3], float spread,
                       float out[3]) {
    float n[3];
    for (int k = 0; k < 3; ++k) {
        float r = orc_rnd(s);
        float t = (r - 0.5f) * 2.0f;
        t = t * spread;
        n[k] = around[k] + t;
    }
    float facing = n[0] * ray[0] + n[1] * ray[1];
    facing = facing + n[2];
    if (facing > 0.0f) {
        n[0] = -n[0];
        n[1] = -n[1];
        n[2] = -n[2];
    }
    float l2 = n[0] * n[0] + n[1] * n[1];
    l2 = l2 + n[2] * n[2];
    float l = sqrtf(l2);
    if (!(l >= 1e-9f)) l = 1e-9f; /* f32::max(l, 1e-9); l is never NaN here */
    out[0] = n[0] / l;
    out[1] = n[1] / l;
    out[2] = n[2] / l;
}

/* geometry.rs:26-32 — normalised -> distorted pixel, f64. */
void orc_distort(double f, double cx, double cy, double k1, double k2, double x, double y,
                 double out[2]) {
    double r2 = x * x + y * y;
    double d = 1.0 + k1 * r2;
    d = d + (k2 * r2) * r2;
    double fd = f * d;
    out[0] = fd * x + cx;
    out[1] = fd * y + cy;
}

/* geometry.rs:34-45 — pixel -> normalised, 8 fixed-point steps, f64. */
void orc_undistort(double f, double cx, double cy, double k1, double k2, double u, double v,
                   double out[2]) {
    double qx = (u - cx) / f, qy = (v - cy) / f;
    double px = qx, py = qy;
    for (int i = 0; i < 8; ++i) {
        double r2 = px * px + py * py;
        double d = 1.0 + k1 * r2;
        d = d + (k2 * r2) * r2;
        px = qx / d;
        py = qy / d;
    }
    out[0] = px;
    out[1] = py;
}

/* dense.rs:299-309 — R = Rs * Rr^T, t = Rs * (Cr - Cs), f64 then cast to f32. */
void orc_relative(const double Rr[9], const double Cr[3], const double Rs[9], const double Cs[3],
                  float R[9], float t[3]) {
    double dc[3] = {Cr[0] - Cs[0], Cr[1] - Cs[1], Cr[2] - Cs[2]};
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

/* dense.rs:69-85 — bilinear grey at pixel-centre coordinates. Returns 0 for
 * None. Non-finite coordinates are treated as None (the reference would read
 * pixel 0 for NaN and panic for values beyond usize; neither occurs for finite
 * projections inside the image). */
int orc_gray_at(const float *g, int32_t w, int32_t h, float x, float y, float *out) {
    float fx = x - 0.5f;
    float fy = y - 0.5f;
    if (!(fx >= 0.0f) || !(fy >= 0.0f)) return 0;
    if (!(fx < 1.8e19f) || !(fy < 1.8e19f)) return 0;
    uint64_t x0 = (uint64_t)fx, y0 = (uint64_t)fy;
    if (x0 + 1 >= (uint64_t)w || y0 + 1 >= (uint64_t)h) return 0;
    float ax = fx - (float)x0, ay = fy - (float)y0;
    size_t i = (size_t)y0 * (size_t)w + (size_t)x0;
    float top = g[i] * (1.0f - ax) + g[i + 1] * ax;
    float bot = g[i + (size_t)w] * (1.0f - ax) + g[i + (size_t)w + 1] * ax;
    *out = top * (1.0f - ay) + bot * ay;
    return 1;
}

/* dense.rs:266-271 — two-pass mean and population standard deviation, with the
 * window size generalised from 25 to n (sums in index order). */
void orc_mean_sd(const float *v, int32_t n, float *m, float *sd) {
    float s = 0.0f;
    for (int32_t i = 0; i < n; ++i) s = s + v[i];
    float mean = s / (float)n;
    float var = 0.0f;
    for (int32_t i = 0; i < n; ++i) {
        float d = v[i] - mean;
        var = var + d * d;
    }
    var = var / (float)n;
    *m = mean;
    *sd = sqrtf(var);
}

/* dense.rs:324-334 — neighbours in list order whose baseline exceeds
 * 0.02 * max(range0, 1), first `take` of them. centres: ncams x 3 (f64).
 * Returns the count written to out. */
int orc_select_neighbours(const double *centres, int32_t ncams, int32_t ci,
                          const int32_t *neighbours, int32_t nn, double range0, int32_t take,
                          int32_t *out) {
    (void)ncams;
    int32_t k = 0;
    double lim = 0.02 * (range0 > 1.0 ? range0 : 1.0);
    for (int32_t a = 0; a < nn && k < take; ++a) {
        int32_t j = neighbours[a];
        const double *cj = centres + 3 * (size_t)j, *cc = centres + 3 * (size_t)ci;
        double dx = cj[0] - cc[0], dy = cj[1] - cc[1], dz = cj[2] - cc[2];
        double base = sqrt(dx * dx + dy * dy + dz * dz);
        if (base > lim) out[k++] = j;
    }
    return k;
}

/* ------------------------------------------------------------------------ */
/* depth-map context                                                        */
/* ------------------------------------------------------------------------ */

struct orc_ctx {
    orc_view rf;
    int32_t nsrc;
    orc_view *srcs;
    float *rays;   /* w*h*2 */
    float *Rs;     /* nsrc*9 */
    float *ts;     /* nsrc*3 */
    int32_t n;     /* window samples */
    int32_t half;  /* window half footprint in pixels */
    int32_t *wdx, *wdy;
    orc_params prm;
    float dmin, dmax, imin, imax;
    float up[3];
};

orc_ctx *orc_ctx_create(const orc_view *ref, const orc_view *srcs, int32_t nsrc,
                        const orc_params *prm, double range0, double range1) {
    if (prm->taps < 1 || (prm->taps & 1) == 0 || prm->taps * prm->taps > ORC_MAX_TAPS)
        return NULL;
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->rf = *ref;
    c->nsrc = nsrc;
    c->srcs = (orc_view *)malloc(sizeof(orc_view) * (size_t)(nsrc > 0 ? nsrc : 1));
    memcpy(c->srcs, srcs, sizeof(orc_view) * (size_t)nsrc);
    c->prm = *prm;
    int32_t w = ref->w, h = ref->h;
    /* dense.rs:343-349 — undistorted ray per pixel centre, f64 -> f32. */
    c->rays = (float *)malloc(sizeof(float) * 2 * (size_t)w * (size_t)h);
    for (int64_t i = 0; i < (int64_t)w * h; ++i) {
        double p[2];
        orc_undistort(ref->f, ref->cx, ref->cy, ref->k1, ref->k2, (double)(i % w) + 0.5,
                      (double)(i / w) + 0.5, p);
        c->rays[2 * i] = (float)p[0];
        c->rays[2 * i + 1] = (float)p[1];
    }
    c->Rs = (float *)malloc(sizeof(float) * 9 * (size_t)(nsrc > 0 ? nsrc : 1));
    c->ts = (float *)malloc(sizeof(float) * 3 * (size_t)(nsrc > 0 ? nsrc : 1));
    for (int32_t s = 0; s < nsrc; ++s)
        orc_relative(ref->R, ref->C, srcs[s].R, srcs[s].C, c->Rs + 9 * s, c->ts + 3 * s);
    /* dense.rs:170 generalised: taps at -(taps/2)*stride .. +(taps/2)*stride. */
    c->n = prm->taps * prm->taps;
    c->half = (prm->taps / 2) * prm->stride;
    c->wdx = (int32_t *)malloc(sizeof(int32_t) * (size_t)c->n);
    c->wdy = (int32_t *)malloc(sizeof(int32_t) * (size_t)c->n);
    int32_t k = 0;
    for (int32_t a = 0; a < prm->taps; ++a)
        for (int32_t b = 0; b < prm->taps; ++b) {
            c->wdy[k] = (a - prm->taps / 2) * prm->stride;
            c->wdx[k] = (b - prm->taps / 2) * prm->stride;
            ++k;
        }
    /* dense.rs:362-363 */
    c->dmin = (float)(range0 > 0.5 ? range0 : 0.5);
    c->dmax = (float)range1;
    c->imin = 1.0f / c->dmax;
    c->imax = 1.0f / c->dmin;
    /* dense.rs:367-370 — R_ref * (0,0,1) is R's third column. */
    c->up[0] = (float)ref->R[2];
    c->up[1] = (float)ref->R[5];
    c->up[2] = (float)ref->R[8];
    return c;
}

void orc_ctx_destroy(orc_ctx *c) {
    if (!c) return;
    free(c->srcs);
    free(c->rays);
    free(c->Rs);
    free(c->ts);
    free(c->wdx);
    free(c->wdy);
    free(c);
}

const float *orc_ctx_rays(const orc_ctx *c) { return c->rays; }

void orc_ctx_relpose(const orc_ctx *c, int32_t s, float R[9], float t[3]) {
    memcpy(R, c->Rs + 9 * s, 9 * sizeof(float));
    memcpy(t, c->ts + 3 * s, 3 * sizeof(float));
}

void orc_ctx_limits(const orc_ctx *c, float out[4]) {
    out[0] = c->dmin;
    out[1] = c->dmax;
    out[2] = c->imin;
    out[3] = c->imax;
}

static int cmp_f32(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

/* dense.rs:181-263 — Ctx2::cost. */
float orc_cost(const orc_ctx *c, int32_t x, int32_t y, const float hyp[4]) {
    const int32_t w = c->rf.w, h = c->rf.h, N = c->n;
    const float *ray = c->rays + 2 * ((size_t)y * (size_t)w + (size_t)x);
    const float d = hyp[0], n0 = hyp[1], n1 = hyp[2], n2 = hyp[3];
    const float xp0 = ray[0] * d, xp1 = ray[1] * d, xp2 = d;
    float cc = n0 * xp0 + n1 * xp1;
    cc = cc + n2 * xp2;
    float refv[ORC_MAX_TAPS], p0[ORC_MAX_TAPS], p1[ORC_MAX_TAPS], p2[ORC_MAX_TAPS];
    for (int32_t k = 0; k < N; ++k) {
        int32_t qx = x + c->wdx[k], qy = y + c->wdy[k];
        if (qx < 0 || qy < 0 || qx >= w || qy >= h) return 2.0f;
        size_t qi = (size_t)qy * (size_t)w + (size_t)qx;
        refv[k] = c->rf.gray[qi];
        const float r0 = c->rays[2 * qi], r1 = c->rays[2 * qi + 1];
        float den = n0 * r0 + n1 * r1;
        den = den + n2;
        if (fabsf(den) < 1e-6f) return 2.0f;
        float dq = cc / den;
        if (dq <= 0.0f) return 2.0f;
        p0[k] = r0 * dq;
        p1[k] = r1 * dq;
        p2[k] = dq;
    }
    float mr, sr;
    orc_mean_sd(refv, N, &mr, &sr);
    if (sr < c->prm.texture_min) return 2.0f;
    float costs_stack[16];
    float *costs = c->nsrc <= 16 ? costs_stack : (float *)malloc(sizeof(float) * (size_t)c->nsrc);
    float sv[ORC_MAX_TAPS];
    for (int32_t si = 0; si < c->nsrc; ++si) {
        costs[si] = 2.0f;
        const float *R = c->Rs + 9 * si, *t = c->ts + 3 * si;
        const orc_view *v = &c->srcs[si];
        int ok = 1;
        for (int32_t j = 0; j < N; ++j) {
            float xs0 = R[0] * p0[j] + R[1] * p1[j];
            xs0 = xs0 + R[2] * p2[j];
            xs0 = xs0 + t[0];
            float xs1 = R[3] * p0[j] + R[4] * p1[j];
            xs1 = xs1 + R[5] * p2[j];
            xs1 = xs1 + t[1];
            float xs2 = R[6] * p0[j] + R[7] * p1[j];
            xs2 = xs2 + R[8] * p2[j];
            xs2 = xs2 + t[2];
            if (xs2 <= 0.0f) {
                ok = 0;
                break;
            }
            double u[2];
            orc_distort(v->f, v->cx, v->cy, v->k1, v->k2, (double)(xs0 / xs2), (double)(xs1 / xs2),
                        u);
            if (!orc_gray_at(v->gray, v->w, v->h, (float)u[0], (float)u[1], &sv[j])) {
                ok = 0;
                break;
            }
        }
        if (!ok) continue;
        float ms, ss;
        orc_mean_sd(sv, N, &ms, &ss);
        if (ss < 1e-3f) continue;
        float cov = 0.0f;
        for (int32_t j = 0; j < N; ++j) cov = cov + (refv[j] - mr) * (sv[j] - ms);
        float den = (float)N * sr;
        den = den * ss;
        costs[si] = 1.0f - cov / den;
    }
    int32_t m = c->nsrc;
    float r;
    if (m <= 0) {
        r = 2.0f;
    } else {
        qsort(costs, (size_t)m, sizeof(float), cmp_f32);
        r = m >= 2 ? (costs[0] + costs[1]) * 0.5f : costs[0];
    }
    if (costs != costs_stack) free(costs);
    return r;
}

/* ------------------------------------------------------------------------ */
/* threading: rows split over pthreads (the rayon row model, dense.rs:417)   */
/* ------------------------------------------------------------------------ */

typedef void (*row_fn)(void *arg, int32_t y0, int32_t y1, int64_t *counter);
typedef struct {
    row_fn fn;
    void *arg;
    int32_t y0, y1;
    int64_t count;
} row_job;

static void *row_job_main(void *p) {
    row_job *j = (row_job *)p;
    j->fn(j->arg, j->y0, j->y1, &j->count);
    return NULL;
}

static int64_t run_rows(row_fn fn, void *arg, int32_t rows, int32_t nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > rows) nthreads = rows > 0 ? rows : 1;
    row_job *jobs = (row_job *)calloc((size_t)nthreads, sizeof(row_job));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int32_t t = 0; t < nthreads; ++t) {
        jobs[t].fn = fn;
        jobs[t].arg = arg;
        jobs[t].y0 = (int32_t)((int64_t)rows * t / nthreads);
        jobs[t].y1 = (int32_t)((int64_t)rows * (t + 1) / nthreads);
    }
    for (int32_t t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, row_job_main, &jobs[t]);
    row_job_main(&jobs[0]);
    int64_t total = jobs[0].count;
    for (int32_t t = 1; t < nthreads; ++t) {
        pthread_join(th[t], NULL);
        total += jobs[t].count;
    }
    free(jobs);
    free(th);
    return total;
}

typedef struct {
    const orc_ctx *c;
    int64_t n;
    const int32_t *px, *py;
    const float *hyp4;
    float *out;
} batch_arg;

static void batch_rows(void *p, int32_t a, int32_t b, int64_t *cnt) {
    batch_arg *ba = (batch_arg *)p;
    for (int64_t i = a; i < b; ++i) ba->out[i] = orc_cost(ba->c, ba->px[i], ba->py[i], ba->hyp4 + 4 * i);
    *cnt += b - a;
}

void orc_cost_batch(const orc_ctx *c, int64_t n, const int32_t *px, const int32_t *py,
                    const float *hyp4, float *out, int32_t nthreads) {
    batch_arg a = {c, n, px, py, hyp4, out};
    run_rows(batch_rows, &a, (int32_t)n, nthreads);
}

/* ------------------------------------------------------------------------ */
/* PatchMatch                                                               */
/* ------------------------------------------------------------------------ */

typedef struct {
    const orc_ctx *c;
    float *state;
    uint8_t *textured;
    const uint8_t *tex_ro;
    int32_t it, parity;
} sweep_arg;

/* dense.rs:371-381 (init) and :385-402 (texture mask). */
static void init_rows(void *p, int32_t y0, int32_t y1, int64_t *cnt) {
    sweep_arg *a = (sweep_arg *)p;
    const orc_ctx *c = a->c;
    const int32_t w = c->rf.w, h = c->rf.h, half = c->half;
    float v[ORC_MAX_TAPS];
    for (int32_t y = y0; y < y1; ++y)
        for (int32_t x = 0; x < w; ++x) {
            size_t i = (size_t)y * (size_t)w + (size_t)x;
            uint64_t s = orc_seed_init((uint64_t)i);
            float r = orc_rnd(&s);
            float dd = c->imin + r * (c->imax - c->imin);
            float *st = a->state + 5 * i;
            st[0] = 1.0f / dd;
            orc_random_normal(&s, c->rays + 2 * i, c->up, 0.3f, st + 1);
            st[4] = orc_cost(c, x, y, st);
            *cnt += 1;
            uint8_t tx = 0;
            if (!(x < half || y < half || x >= w - half || y >= h - half)) {
                for (int32_t k = 0; k < c->n; ++k)
                    v[k] = c->rf.gray[(size_t)(y + c->wdy[k]) * (size_t)w + (size_t)(x + c->wdx[k])];
                float m, sd;
                orc_mean_sd(v, c->n, &m, &sd);
                tx = sd >= c->prm.texture_min;
            }
            a->textured[i] = tx;
        }
}

void orc_init(const orc_ctx *c, float *state, uint8_t *textured, int32_t nthreads) {
    sweep_arg a = {c, state, textured, textured, 0, 0};
    run_rows(init_rows, &a, c->rf.h, nthreads);
}

static const int32_t CAND[8][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1},
                                   {-5, 0}, {5, 0}, {0, -5}, {0, 5}}; /* dense.rs:403-412 */

/* dense.rs:413-467 — one half-sweep (one parity of one iteration). Candidates
 * are all of the opposite colour, so reading the live state equals reading the
 * snapshot the reference clones at :416. Returns cost evaluations done. */
static void sweep_rows(void *p, int32_t y0, int32_t y1, int64_t *cnt) {
    sweep_arg *a = (sweep_arg *)p;
    const orc_ctx *c = a->c;
    const int32_t w = c->rf.w, h = c->rf.h;
    const float range = ldexpf(1.0f, -a->it); /* 0.5f32.powi(it), exact */
    const float span = c->imax - c->imin;
    for (int32_t y = y0; y < y1; ++y)
        for (int32_t x = 0; x < w; ++x) {
            if ((x + y) % 2 != a->parity) continue;
            size_t i = (size_t)y * (size_t)w + (size_t)x;
            if (!a->tex_ro[i]) continue;
            float best[5];
            memcpy(best, a->state + 5 * i, sizeof best);
            for (int k = 0; k < 8; ++k) {
                int32_t nx = x + CAND[k][0], ny = y + CAND[k][1];
                if (nx < 0 || ny < 0 || nx >= w || ny >= h) continue;
                const float *cand = a->state + 5 * ((size_t)ny * (size_t)w + (size_t)nx);
                if (cand[0] == best[0] && cand[1] == best[1] && cand[2] == best[2] &&
                    cand[3] == best[3])
                    continue;
                float cost = orc_cost(c, x, y, cand);
                *cnt += 1;
                if (cost < best[4]) {
                    memcpy(best, cand, 4 * sizeof(float));
                    best[4] = cost;
                }
            }
            uint64_t s = orc_seed_refine((uint64_t)i, (uint64_t)a->it, (uint64_t)a->parity);
            for (int rep = 0; rep < 2; ++rep) {
                float r = orc_rnd(&s);
                float step = (r - 0.5f) * range;
                step = step * 0.25f;
                step = step * span;
                float id = 1.0f / best[0] + step;
                float hyp[4];
                hyp[0] = (id > c->imin * 0.5f && id < c->imax * 1.5f) ? 1.0f / id : best[0];
                orc_random_normal(&s, c->rays + 2 * i, best + 1, 0.25f * range, hyp + 1);
                float cost = orc_cost(c, x, y, hyp);
                *cnt += 1;
                if (cost < best[4]) {
                    memcpy(best, hyp, 4 * sizeof(float));
                    best[4] = cost;
                }
            }
            memcpy(a->state + 5 * i, best, sizeof best);
        }
}

int64_t orc_half_sweep(const orc_ctx *c, float *state, const uint8_t *textured, int32_t it,
                       int32_t parity, int32_t nthreads) {
    sweep_arg a = {c, state, NULL, textured, it, parity};
    return run_rows(sweep_rows, &a, c->rf.h, nthreads);
}

/* dense.rs:468-477 */
void orc_threshold(const orc_ctx *c, const float *state, float *depth) {
    const float lo = c->dmin * 0.7f, hi = c->dmax * 1.4f;
    for (int64_t i = 0; i < (int64_t)c->rf.w * c->rf.h; ++i) {
        const float *st = state + 5 * i;
        depth[i] = (st[4] < c->prm.cost_max && st[0] >= lo && st[0] <= hi) ? st[0] : 0.0f;
    }
}

int64_t orc_depth_map(const orc_ctx *c, int32_t iterations, int32_t nthreads, float *state,
                      uint8_t *textured, float *depth) {
    int64_t evals = 0;
    orc_init(c, state, textured, nthreads);
    for (int32_t it = 0; it < iterations; ++it)
        for (int32_t parity = 0; parity < 2; ++parity)
            evals += orc_half_sweep(c, state, textured, it, parity, nthreads);
    if (depth) orc_threshold(c, state, depth);
    return evals;
}

/* ------------------------------------------------------------------------ */
/* geometric consistency (fusion's filter)                                  */
/* ------------------------------------------------------------------------ */

typedef struct {
    const orc_view *ref;
    const float *depth;
    const orc_nbr_depth *nbrs;
    int32_t nnbr;
    int32_t *agree;
    int32_t need;
    int32_t rule; /* 0: dense.rs:585-615; 1: reprojection <= 1 px */
} cons_arg;

/* dense.rs:585-612 for one pixel: -1 where depth <= 0, else the number of
 * neighbours that agree (rule 0: relative depth difference < 1%; rule 1:
 * reprojection within 1 px); xw receives the pixel's world point. */
static int32_t cons_pixel(const cons_arg *a, int32_t x, int32_t y, double xw[3]) {
    const orc_view *rf = a->ref;
    const size_t i = (size_t)y * (size_t)rf->w + (size_t)x;
    const float di = a->depth[i];
    if (di <= 0.0f) return -1;
    double pp[2];
    orc_undistort(rf->f, rf->cx, rf->cy, rf->k1, rf->k2, (double)x + 0.5, (double)y + 0.5, pp);
    double xc[3] = {pp[0] * (double)di, pp[1] * (double)di, (double)di};
    for (int r = 0; r < 3; ++r) {
        /* rwc = R^T: row r of R^T is column r of R */
        double acc = rf->R[0 * 3 + r] * xc[0];
        acc = acc + rf->R[1 * 3 + r] * xc[1];
        acc = acc + rf->R[2 * 3 + r] * xc[2];
        xw[r] = acc + rf->C[r];
    }
    int32_t agree = 0;
    for (int32_t k = 0; k < a->nnbr; ++k) {
        const orc_view *v = a->nbrs[k].view;
        double dx[3] = {xw[0] - v->C[0], xw[1] - v->C[1], xw[2] - v->C[2]};
        double xs[3];
        for (int r = 0; r < 3; ++r) {
            double acc = v->R[r * 3 + 0] * dx[0];
            acc = acc + v->R[r * 3 + 1] * dx[1];
            acc = acc + v->R[r * 3 + 2] * dx[2];
            xs[r] = acc;
        }
        if (xs[2] <= 0.0) continue;
        double u[2];
        orc_distort(v->f, v->cx, v->cy, v->k1, v->k2, xs[0] / xs[2], xs[1] / xs[2], u);
        double fx = floor(u[0]), fy = floor(u[1]);
        if (!(fx >= 0.0) || !(fy >= 0.0) || !(fx < (double)v->w) || !(fy < (double)v->h))
            continue;
        int64_t ux = (int64_t)fx, uy = (int64_t)fy;
        double dn = (double)a->nbrs[k].depth[(size_t)uy * (size_t)v->w + (size_t)ux];
        if (a->rule == 0) {
            if (dn > 0.0 && fabs((dn - xs[2]) / xs[2]) < 0.01) agree += 1;
            continue;
        }
        /* reprojection: the neighbour's own point at that pixel, back in
         * the reference image, within 1 px of the pixel centre */
        if (!(dn > 0.0)) continue;
        double pn[2];
        orc_undistort(v->f, v->cx, v->cy, v->k1, v->k2, (double)ux + 0.5, (double)uy + 0.5, pn);
        double xn[3] = {pn[0] * dn, pn[1] * dn, dn}, xwn[3];
        for (int r = 0; r < 3; ++r) {
            double acc = v->R[0 * 3 + r] * xn[0];
            acc = acc + v->R[1 * 3 + r] * xn[1];
            acc = acc + v->R[2 * 3 + r] * xn[2];
            xwn[r] = acc + v->C[r];
        }
        double er[3] = {xwn[0] - rf->C[0], xwn[1] - rf->C[1], xwn[2] - rf->C[2]}, xr[3];
        for (int r = 0; r < 3; ++r) {
            double acc = rf->R[r * 3 + 0] * er[0];
            acc = acc + rf->R[r * 3 + 1] * er[1];
            acc = acc + rf->R[r * 3 + 2] * er[2];
            xr[r] = acc;
        }
        if (!(xr[2] > 0.0)) continue;
        double ur[2];
        orc_distort(rf->f, rf->cx, rf->cy, rf->k1, rf->k2, xr[0] / xr[2], xr[1] / xr[2], ur);
        double ex = ur[0] - ((double)x + 0.5), ey = ur[1] - ((double)y + 0.5);
        if (ex * ex + ey * ey <= 1.0) agree += 1;
    }
    return agree;
}

/* dense.rs:585-615 */
static void cons_rows(void *p, int32_t y0, int32_t y1, int64_t *cnt) {
    cons_arg *a = (cons_arg *)p;
    const int32_t w = a->ref->w;
    for (int32_t y = y0; y < y1; ++y)
        for (int32_t x = 0; x < w; ++x) {
            double xw[3];
            const int32_t agree = cons_pixel(a, x, y, xw);
            a->agree[(size_t)y * (size_t)w + (size_t)x] = agree;
            if (agree >= a->need) *cnt += 1;
        }
}

int64_t orc_consistency(const orc_view *ref, const float *depth, const orc_nbr_depth *nbrs,
                        int32_t nnbr, int32_t *agree, int32_t nthreads) {
    cons_arg a = {ref, depth, nbrs, nnbr, agree, nnbr >= 3 ? 2 : 1, 0};
    return run_rows(cons_rows, &a, ref->h, nthreads);
}

int64_t orc_consistency_reproj(const orc_view *ref, const float *depth, const orc_nbr_depth *nbrs,
                               int32_t nnbr, int32_t need, int32_t *agree, int32_t nthreads) {
    cons_arg a = {ref, depth, nbrs, nnbr, agree, need, 1};
    return run_rows(cons_rows, &a, ref->h, nthreads);
}

/* dense.rs:589-632 for one reference view: every pixel whose depth the
 * neighbours confirm (agree >= need under `rule`) emits its voxel key
 * floor(xw / voxel) (dense.rs:617-621), its world point (f64) and its pixel
 * index (for the colour, dense.rs:622). Sequential; returns the count. */
int64_t orc_fuse_points(const orc_view *ref, const float *depth, const orc_nbr_depth *nbrs,
                        int32_t nnbr, int32_t rule, int32_t need, double inv_voxel,
                        int32_t *keys, double *xw_out, int32_t *pix) {
    cons_arg a = {ref, depth, nbrs, nnbr, NULL, need, rule};
    int64_t n = 0;
    for (int32_t y = 0; y < ref->h; ++y)
        for (int32_t x = 0; x < ref->w; ++x) {
            double xw[3];
            const int32_t agree = cons_pixel(&a, x, y, xw);
            if (agree < 0 || agree < need) continue;
            for (int r = 0; r < 3; ++r) {
                keys[3 * n + r] = (int32_t)floor(xw[r] * inv_voxel);
                xw_out[3 * n + r] = xw[r];
            }
            pix[n] = y * ref->w + x;
            ++n;
        }
    return n;
}

/* dense.rs:545-562: the voxel edge from the sparse points' 1%-99% plan extent,
 * the cameras' mid depths and the working-size focal length. */
double orc_fusion_voxel(const double *xs, const double *ys, int64_t npts, const double *ranges,
                        int32_t ncams, double f, int32_t size, int32_t width, int32_t height,
                        int64_t max_points) {
    double *sx = (double *)malloc(sizeof(double) * (size_t)npts);
    double *sy = (double *)malloc(sizeof(double) * (size_t)npts);
    double *dm = (double *)malloc(sizeof(double) * (size_t)ncams);
    for (int64_t i = 0; i < npts; ++i) {
        sx[i] = xs[i];
        sy[i] = ys[i];
    }
    for (int32_t c = 0; c < ncams; ++c) dm[c] = (ranges[2 * c] + ranges[2 * c + 1]) / 2.0;
    /* insertion sorts keep the oracle free of library ordering subtleties */
    for (int64_t i = 1; i < npts; ++i)
        for (int64_t j = i; j > 0 && sx[j - 1] > sx[j]; --j) {
            double t = sx[j]; sx[j] = sx[j - 1]; sx[j - 1] = t;
        }
    for (int64_t i = 1; i < npts; ++i)
        for (int64_t j = i; j > 0 && sy[j - 1] > sy[j]; --j) {
            double t = sy[j]; sy[j] = sy[j - 1]; sy[j - 1] = t;
        }
    for (int32_t i = 1; i < ncams; ++i)
        for (int32_t j = i; j > 0 && dm[j - 1] > dm[j]; --j) {
            double t = dm[j]; dm[j] = dm[j - 1]; dm[j - 1] = t;
        }
#define ORC_PCT(v, q) (v)[(size_t)((double)(npts - 1) * (q))]
    double area = (ORC_PCT(sx, 0.99) - ORC_PCT(sx, 0.01)) * (ORC_PCT(sy, 0.99) - ORC_PCT(sy, 0.01));
#undef ORC_PCT
    if (!(area >= 1.0)) area = 1.0;
    const double f0 = f * (double)size / (double)(width > height ? width : height);
    const double gsd = dm[ncams / 2] / f0;
    double voxel = gsd * 1.5;
    const double alt = sqrt(area * 2.5 / (double)max_points);
    if (alt > voxel) voxel = alt;
    free(sx);
    free(sy);
    free(dm);
    return voxel;
}

/* ------------------------------------------------------------------------ */
/* fixed-pair f64 cost and gradient (checks of the GPU refinement)          */
/* ------------------------------------------------------------------------ */

/* Bilinear grey at pixel-centre coordinates in f64 with its derivatives
 * (the same cell convention as gray_at, dense.rs:69-85). */
static int gray_at64(const float *g, int32_t w, int32_t h, double x, double y, double *val,
                     double *gu, double *gv, uint64_t *cell) {
    double fx = x - 0.5, fy = y - 0.5;
    if (!(fx >= 0.0) || !(fy >= 0.0) || !(fx < (double)(w - 1)) || !(fy < (double)(h - 1)))
        return 0;
    int64_t x0 = (int64_t)fx, y0 = (int64_t)fy;
    double ax = fx - (double)x0, ay = fy - (double)y0;
    size_t i = (size_t)y0 * (size_t)w + (size_t)x0;
    double g00 = g[i], g10 = g[i + 1], g01 = g[i + (size_t)w], g11 = g[i + (size_t)w + 1];
    double top = g00 * (1.0 - ax) + g10 * ax, bot = g01 * (1.0 - ax) + g11 * ax;
    *val = top * (1.0 - ay) + bot * ay;
    *gu = (1.0 - ay) * (g10 - g00) + ay * (g11 - g01);
    *gv = (1.0 - ax) * (g01 - g00) + ax * (g11 - g10);
    *cell = *cell * 1000003ull + (uint64_t)(y0 * 65536 + x0);
    return 1;
}

/* The window of pixel (x, y) under plane hyp = (d, n0, n1, n2): reference
 * grey, plane depth dq and its derivatives d dq / d(d, n0, n1, n2) per tap,
 * and the tap rays. 0 where the window leaves the image or the plane is
 * degenerate. */
typedef struct {
    double refv[ORC_MAX_TAPS], dq[ORC_MAX_TAPS], rk[ORC_MAX_TAPS][3];
    double ddq[ORC_MAX_TAPS][4];
    double mr, na;
} pair_window64;

static int pair_window(const orc_ctx *c, int32_t x, int32_t y, const double hyp[4],
                       pair_window64 *pw) {
    const int32_t w = c->rf.w, h = c->rf.h, N = c->n;
    const double d = hyp[0], n[3] = {hyp[1], hyp[2], hyp[3]};
    const float *rc = c->rays + 2 * ((size_t)y * (size_t)w + (size_t)x);
    const double ch[3] = {rc[0], rc[1], 1.0};
    const double nc = n[0] * ch[0] + n[1] * ch[1] + n[2] * ch[2];
    const double cc = d * nc;
    for (int32_t k = 0; k < N; ++k) {
        int32_t qx = x + c->wdx[k], qy = y + c->wdy[k];
        if (qx < 0 || qy < 0 || qx >= w || qy >= h) return 0;
        size_t qi = (size_t)qy * (size_t)w + (size_t)qx;
        pw->refv[k] = c->rf.gray[qi];
        pw->rk[k][0] = c->rays[2 * qi];
        pw->rk[k][1] = c->rays[2 * qi + 1];
        pw->rk[k][2] = 1.0;
        double den = n[0] * pw->rk[k][0] + n[1] * pw->rk[k][1] + n[2];
        if (fabs(den) < 1e-6) return 0;
        pw->dq[k] = cc / den;
        if (pw->dq[k] <= 0.0) return 0;
        pw->ddq[k][0] = nc / den;
        for (int a = 0; a < 3; ++a) pw->ddq[k][1 + a] = (d * ch[a] - pw->dq[k] * pw->rk[k][a]) / den;
    }
    double mr = 0.0;
    for (int32_t k = 0; k < N; ++k) mr += pw->refv[k];
    mr /= N;
    double na = 0.0;
    for (int32_t k = 0; k < N; ++k) na += (pw->refv[k] - mr) * (pw->refv[k] - mr);
    pw->mr = mr;
    pw->na = sqrt(na);
    return pw->na >= 1e-12;
}

/* Samples of source si over the window and their derivatives w.r.t.
 * (d, n0, n1, n2): the projective warp, the radial model (geometry.rs:26-32)
 * and the bilinear grey, all in f64. 0 where a sample fails. */
static int pair_source(const orc_ctx *c, const pair_window64 *pw, int32_t si, double *sv,
                       double (*dsv)[4], uint64_t *cell) {
    const int32_t N = c->n;
    const orc_view *v = &c->srcs[si];
    double Rr[9], tt[3];
    /* relative pose in f64 (the f32 cast is the kernels' business) */
    {
        const double *Rr_ = c->rf.R, *Cr = c->rf.C, *Rs = v->R, *Cs = v->C;
        double dc[3] = {Cr[0] - Cs[0], Cr[1] - Cs[1], Cr[2] - Cs[2]};
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j)
                Rr[i * 3 + j] = Rs[i * 3] * Rr_[j * 3] + Rs[i * 3 + 1] * Rr_[j * 3 + 1] +
                                Rs[i * 3 + 2] * Rr_[j * 3 + 2];
            tt[i] = Rs[i * 3] * dc[0] + Rs[i * 3 + 1] * dc[1] + Rs[i * 3 + 2] * dc[2];
        }
    }
    for (int32_t k = 0; k < N; ++k) {
        const double *rk = pw->rk[k];
        double p[3] = {pw->dq[k] * rk[0], pw->dq[k] * rk[1], pw->dq[k]};
        double xs[3], wk[3];
        for (int i = 0; i < 3; ++i) {
            xs[i] = Rr[i * 3] * p[0] + Rr[i * 3 + 1] * p[1] + Rr[i * 3 + 2] * p[2] + tt[i];
            wk[i] = Rr[i * 3] * rk[0] + Rr[i * 3 + 1] * rk[1] + Rr[i * 3 + 2];
        }
        if (xs[2] <= 0.0) return 0;
        const double qx = xs[0] / xs[2], qy = xs[1] / xs[2];
        const double r2 = qx * qx + qy * qy;
        const double D = 1.0 + v->k1 * r2 + v->k2 * r2 * r2;
        const double dD = 2.0 * v->k1 + 4.0 * v->k2 * r2; /* dD/dx = dD * x */
        const double u = v->f * D * qx + v->cx, vv = v->f * D * qy + v->cy;
        double val, gu, gv;
        if (!gray_at64(v->gray, v->w, v->h, u, vv, &val, &gu, &gv, cell)) return 0;
        sv[k] = val;
        /* d(u, v)/d(qx, qy) */
        const double ux = v->f * (D + qx * dD * qx), uy = v->f * qx * dD * qy;
        const double vx = v->f * qy * dD * qx, vy = v->f * (D + qy * dD * qy);
        /* d(qx, qy)/d dq = (w0 - qx w2, w1 - qy w2) / xs2 */
        const double dqx = (wk[0] - qx * wk[2]) / xs[2], dqy = (wk[1] - qy * wk[2]) / xs[2];
        const double dI = gu * (ux * dqx + uy * dqy) + gv * (vx * dqx + vy * dqy);
        for (int a = 0; a < 4; ++a) dsv[k][a] = dI * pw->ddq[k][a];
    }
    return 1;
}

int orc_pair_terms64(const orc_ctx *c, int32_t x, int32_t y, const double hyp[4], int32_t s0,
                     int32_t s1, double *cost, double grad[4], double jtj[16], uint64_t *cells) {
    const int32_t N = c->n;
    if (s0 < 0 || s1 < 0 || s0 >= c->nsrc || s1 >= c->nsrc) return 0;
    pair_window64 pw;
    if (!pair_window(c, x, y, hyp, &pw)) return 0;
    uint64_t cell = 0;
    double total = 0.0, gsum[4] = {0, 0, 0, 0}, H[16];
    for (int a = 0; a < 16; ++a) H[a] = 0.0;
    const int32_t pair[2] = {s0, s1};
    for (int ps = 0; ps < 2; ++ps) {
        double sv[ORC_MAX_TAPS], dsv[ORC_MAX_TAPS][4];
        if (!pair_source(c, &pw, pair[ps], sv, dsv, &cell)) return 0;
        double ms = 0.0;
        for (int32_t k = 0; k < N; ++k) ms += sv[k];
        ms /= N;
        double nb = 0.0, ab = 0.0;
        for (int32_t k = 0; k < N; ++k) {
            nb += (sv[k] - ms) * (sv[k] - ms);
            ab += (pw.refv[k] - pw.mr) * (sv[k] - ms);
        }
        nb = sqrt(nb);
        if (nb < 1e-12) return 0;
        const double ncc = ab / (pw.na * nb);
        total += ncc;
        /* d ncc / d sv_k = (a_k / |a| - ncc b_k / |b|) / |b| */
        for (int32_t k = 0; k < N; ++k) {
            const double g = ((pw.refv[k] - pw.mr) / pw.na - ncc * (sv[k] - ms) / nb) / nb;
            for (int a = 0; a < 4; ++a) gsum[a] += g * dsv[k][a];
        }
        if (jtj) {
            /* residual r = a/|a| - b/|b| (1 - ncc = |r|^2 / 2); its Jacobian
             * is -(I - u u^T) Dc / |b| with u = b/|b| and Dc the derivatives
             * minus their window means, so J^T J = (Dc^T Dc - (Dc^T u)(u^T Dc)) / |b|^2 */
            double dm[4] = {0, 0, 0, 0}, du[4] = {0, 0, 0, 0};
            for (int32_t k = 0; k < N; ++k)
                for (int a = 0; a < 4; ++a) dm[a] += dsv[k][a];
            for (int a = 0; a < 4; ++a) dm[a] /= N;
            for (int32_t k = 0; k < N; ++k) {
                const double u = (sv[k] - ms) / nb;
                for (int a = 0; a < 4; ++a) du[a] += u * (dsv[k][a] - dm[a]);
            }
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) {
                    double dd = 0.0;
                    for (int32_t k = 0; k < N; ++k) dd += (dsv[k][a] - dm[a]) * (dsv[k][b] - dm[b]);
                    H[4 * a + b] += (dd - du[a] * du[b]) / (nb * nb);
                }
        }
    }
    *cost = 1.0 - 0.5 * total;
    if (grad)
        for (int a = 0; a < 4; ++a) grad[a] = -0.5 * gsum[a];
    if (jtj)
        for (int a = 0; a < 16; ++a) jtj[a] = H[a];
    if (cells) *cells = cell;
    return 1;
}

int orc_pair_cost64(const orc_ctx *c, int32_t x, int32_t y, const double hyp[4], int32_t s0,
                    int32_t s1, double *cost, double grad[4], uint64_t *cells) {
    return orc_pair_terms64(c, x, y, hyp, s0, s1, cost, grad, NULL, cells);
}

int orc_pair_unit64(const orc_ctx *c, int32_t x, int32_t y, const double hyp[4], int32_t s,
                    double *unit, uint64_t *cells) {
    const int32_t N = c->n;
    if (s < 0 || s >= c->nsrc) return 0;
    pair_window64 pw;
    if (!pair_window(c, x, y, hyp, &pw)) return 0;
    uint64_t cell = 0;
    double sv[ORC_MAX_TAPS], dsv[ORC_MAX_TAPS][4];
    if (!pair_source(c, &pw, s, sv, dsv, &cell)) return 0;
    double ms = 0.0;
    for (int32_t k = 0; k < N; ++k) ms += sv[k];
    ms /= N;
    double nb = 0.0;
    for (int32_t k = 0; k < N; ++k) nb += (sv[k] - ms) * (sv[k] - ms);
    nb = sqrt(nb);
    if (nb < 1e-12) return 0;
    for (int32_t k = 0; k < N; ++k) unit[k] = (sv[k] - ms) / nb;
    if (cells) *cells = cell;
    return 1;
}


/* ---- view preparation ---------------------------------------------------- */

/* imgops.rs:145-171 shrink_to: the sizes only. */
void orc_shrink_size(int32_t w, int32_t h, int32_t max_side, int32_t *f, int32_t *bw,
                     int32_t *bh, int32_t *ow, int32_t *oh, double *scale) {
    const int32_t lng = w > h ? w : h;
    if (lng <= max_side) {
        *f = 0;
        *bw = *ow = w;
        *bh = *oh = h;
        *scale = 1.0;
        return;
    }
    const int32_t fac = lng / max_side;
    *f = fac >= 2 ? fac : 0;
    *bw = fac >= 2 ? w / fac : w; /* box_down: w / f, h / f (imgops.rs:94-95) */
    *bh = fac >= 2 ? h / fac : h;
    *scale = (double)max_side / (double)lng;
    const int32_t bl = *bw > *bh ? *bw : *bh;
    const double k = (double)max_side / (double)bl;
    *ow = (int32_t)round((double)*bw * k);
    *oh = (int32_t)round((double)*bh * k);
    if (*ow < 1) *ow = 1;
    if (*oh < 1) *oh = 1;
}

/* imgops.rs:89-118 box_down: u32 sums over f x f, (sum + area/2) / area. */
void orc_box_down(const uint8_t *src, int32_t w, int32_t h, int32_t stride, int32_t f,
                  uint8_t *dst) {
    const int32_t ow = w / f, oh = h / f;
    const uint32_t area = (uint32_t)(f * f);
    for (int32_t oy = 0; oy < oh; ++oy)
        for (int32_t ox = 0; ox < ow; ++ox)
            for (int c = 0; c < 3; ++c) {
                uint32_t a = 0;
                for (int32_t sy = oy * f; sy < oy * f + f; ++sy)
                    for (int32_t k = 0; k < f; ++k)
                        a += src[(int64_t)sy * stride + (int64_t)(ox * f + k) * 3 + c];
                dst[((int64_t)oy * ow + ox) * 3 + c] = (uint8_t)((a + area / 2) / area);
            }
}

/* imgops.rs:120-143 resize_bilinear, f64, left to right, rounded half away. */
void orc_resize_bilinear(const uint8_t *src, int32_t sw, int32_t sh, int32_t stride,
                         uint8_t *dst, int32_t w, int32_t h) {
    const double sx = (double)sw / (double)w, sy = (double)sh / (double)h;
    for (int32_t y = 0; y < h; ++y) {
        double fy = ((double)y + 0.5) * sy - 0.5;
        fy = fy < 0.0 ? 0.0 : (fy > (double)(sh - 1) ? (double)(sh - 1) : fy);
        const int32_t y0 = (int32_t)fy;
        const int32_t y1 = y0 + 1 < sh - 1 ? y0 + 1 : sh - 1;
        const double ay = fy - (double)y0;
        for (int32_t x = 0; x < w; ++x) {
            double fx = ((double)x + 0.5) * sx - 0.5;
            fx = fx < 0.0 ? 0.0 : (fx > (double)(sw - 1) ? (double)(sw - 1) : fx);
            const int32_t x0 = (int32_t)fx;
            const int32_t x1 = x0 + 1 < sw - 1 ? x0 + 1 : sw - 1;
            const double ax = fx - (double)x0;
            for (int c = 0; c < 3; ++c) {
                const double p00 = src[(int64_t)y0 * stride + (int64_t)x0 * 3 + c];
                const double p10 = src[(int64_t)y0 * stride + (int64_t)x1 * 3 + c];
                const double p01 = src[(int64_t)y1 * stride + (int64_t)x0 * 3 + c];
                const double p11 = src[(int64_t)y1 * stride + (int64_t)x1 * 3 + c];
                double top = p00 * (1.0 - ax);
                top = top + p10 * ax;
                double bot = p01 * (1.0 - ax);
                bot = bot + p11 * ax;
                double v = top * (1.0 - ay);
                v = v + bot * ay;
                dst[((int64_t)y * w + x) * 3 + c] = (uint8_t)round(v);
            }
        }
    }
}

/* dense.rs:100-112: shrink_to, then the grey of every pixel (dense.rs:108). */
void orc_prepare_view(const uint8_t *rgb, int32_t w, int32_t h, int32_t stride,
                      int32_t max_side, uint8_t *out_rgb, float *grey) {
    int32_t f, bw, bh, ow, oh;
    double scale;
    orc_shrink_size(w, h, max_side, &f, &bw, &bh, &ow, &oh, &scale);
    uint8_t *boxed = NULL;
    const uint8_t *b = rgb;
    int32_t bstride = stride;
    if (f >= 2) {
        boxed = (uint8_t *)malloc((size_t)bw * bh * 3);
        orc_box_down(rgb, w, h, stride, f, boxed);
        b = boxed;
        bstride = bw * 3;
    }
    if (ow == bw && oh == bh) {
        for (int32_t y = 0; y < oh; ++y)
            memcpy(out_rgb + (int64_t)y * ow * 3, b + (int64_t)y * bstride, (size_t)ow * 3);
    } else {
        orc_resize_bilinear(b, bw, bh, bstride, out_rgb, ow, oh);
    }
    free(boxed);
    const int64_t n = (int64_t)ow * oh;
    for (int64_t i = 0; i < n; ++i)
        grey[i] = orc_gray_rgb(out_rgb[3 * i], out_rgb[3 * i + 1], out_rgb[3 * i + 2]);
}


/* ---- DSM (recon/dsm.rs) ---------------------------------------------------- */

void orc_free(void *p) { free(p); }

static int cmp_f32_total(const void *a, const void *b) {
    /* f32::total_cmp on the values the DSM sorts (finite ones) */
    int32_t x, y;
    memcpy(&x, a, 4);
    memcpy(&y, b, 4);
    x ^= (int32_t)(((uint32_t)(x >> 31)) >> 1);
    y ^= (int32_t)(((uint32_t)(y >> 31)) >> 1);
    return (x > y) - (x < y);
}

/* dsm.rs:183-201: median of the finite values in the (2r+1)^2 window, None
 * (returns 0) when fewer than 5. */
static int dsm_med(const float *h, int64_t nx, int64_t ny, int64_t i, int64_t j, int64_t r,
                   float *out) {
    float v[25];
    int n = 0;
    for (int64_t dj = -r; dj <= r; ++dj)
        for (int64_t di = -r; di <= r; ++di) {
            const int64_t a = i + di, b = j + dj;
            if (a >= 0 && b >= 0 && a < nx && b < ny) {
                const float x = h[b * nx + a];
                if (isfinite(x)) v[n++] = x;
            }
        }
    if (n < 5) return 0;
    qsort(v, (size_t)n, sizeof(float), cmp_f32_total);
    *out = v[n / 2];
    return 1;
}

/* dsm.rs:293-331: 4-connected components of valid cells smaller than
 * min_cells become empty. */
void orc_remove_islands(float *h, int64_t nx, int64_t ny, int64_t min_cells) {
    const int64_t n = nx * ny;
    uint8_t *seen = (uint8_t *)calloc((size_t)n, 1);
    int64_t *stack = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *comp = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    for (int64_t start = 0; start < n; ++start) {
        if (!isfinite(h[start]) || seen[start]) continue;
        int64_t sp = 0, nc = 0;
        stack[sp++] = start;
        seen[start] = 1;
        while (sp) {
            const int64_t k = stack[--sp];
            comp[nc++] = k;
            const int64_t i = k % nx, j = k / nx;
            const int64_t nb[4] = {i > 0 ? k - 1 : -1, i + 1 < nx ? k + 1 : -1,
                                   j > 0 ? k - nx : -1, j + 1 < ny ? k + nx : -1};
            for (int q = 0; q < 4; ++q) {
                const int64_t kk = nb[q];
                if (kk >= 0 && isfinite(h[kk]) && !seen[kk]) {
                    seen[kk] = 1;
                    stack[sp++] = kk;
                }
            }
        }
        if (nc < min_cells)
            for (int64_t q = 0; q < nc; ++q) h[comp[q]] = NAN;
    }
    free(seen);
    free(stack);
    free(comp);
}

/* dsm.rs:335-390: empty cells not reachable from the border through empty
 * cells are filled by repeated 4-neighbour averaging of a snapshot. */
void orc_fill_enclosed(float *h, int64_t nx, int64_t ny) {
    const int64_t n = nx * ny;
    uint8_t *outside = (uint8_t *)calloc((size_t)n, 1);
    /* every cell is marked once and pushes at most 4: border + 4 n entries */
    int64_t *stack = (int64_t *)malloc((size_t)(4 * n + 2 * nx + 2 * ny + 4) * sizeof(int64_t));
    int64_t sp = 0;
    for (int64_t i = 0; i < nx; ++i) {
        stack[sp++] = i;
        stack[sp++] = (ny - 1) * nx + i;
    }
    for (int64_t j = 0; j < ny; ++j) {
        stack[sp++] = j * nx;
        stack[sp++] = j * nx + nx - 1;
    }
    while (sp) {
        const int64_t k = stack[--sp];
        if (isfinite(h[k]) || outside[k]) continue;
        outside[k] = 1;
        const int64_t i = k % nx, j = k / nx;
        if (i > 0) stack[sp++] = k - 1;
        if (i + 1 < nx) stack[sp++] = k + 1;
        if (j > 0) stack[sp++] = k - nx;
        if (j + 1 < ny) stack[sp++] = k + nx;
    }
    int64_t *todo = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t nt = 0;
    for (int64_t k = 0; k < n; ++k)
        if (!isfinite(h[k]) && !outside[k]) todo[nt++] = k;
    float *snap = (float *)malloc((size_t)n * sizeof(float));
    for (int guard = 0; nt > 0 && guard < 10000; ++guard) {
        memcpy(snap, h, (size_t)n * sizeof(float));
        int64_t kept = 0;
        for (int64_t t = 0; t < nt; ++t) {
            const int64_t k = todo[t], i = k % nx, j = k / nx;
            float s = 0.0f;
            int c = 0;
            const int64_t d[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
            for (int q = 0; q < 4; ++q) {
                const int64_t a = i + d[q][0], b = j + d[q][1];
                if (a >= 0 && b >= 0 && a < nx && b < ny) {
                    const float x = snap[b * nx + a];
                    if (isfinite(x)) {
                        s += x;
                        ++c;
                    }
                }
            }
            if (c > 0)
                h[k] = s / (float)c;
            else
                todo[kept++] = k;
        }
        nt = kept;
    }
    free(snap);
    free(todo);
    free(outside);
    free(stack);
}

static double dsm_q(float *v, int64_t n, double f) {
    /* dsm.rs:141: v[((len - 1) as f64 * f) as usize] of the sorted values */
    const int64_t k = (int64_t)((double)(n - 1) * f);
    qsort(v, (size_t)n, sizeof(float), cmp_f32_total);
    return (double)v[k];
}

int orc_build_dsm(const float *xyz, int64_t n, double voxel_m, int64_t budget, orc_dsm *out,
                  float **hout) {
    *hout = NULL;
    if (n < 100) return -1;
    float *xs = (float *)malloc((size_t)n * sizeof(float));
    float *ys = (float *)malloc((size_t)n * sizeof(float));
    for (int64_t p = 0; p < n; ++p) {
        xs[p] = xyz[3 * p];
        ys[p] = xyz[3 * p + 1];
    }
    const double xa = dsm_q(xs, n, 0.005), xb = dsm_q(xs, n, 0.995);
    const double ya = dsm_q(ys, n, 0.005), yb = dsm_q(ys, n, 0.995);
    free(xs);
    free(ys);
    const double area = (xb - xa) * (yb - ya);
    double cell = voxel_m * 1.5;
    const double c2 = sqrt(area / (double)budget);
    if (c2 > cell) cell = c2;
    // Rust's `as usize` makes a NaN edge count (a NaN percentile) 0
    const double ex = ceil((xb - xa) / cell), ey = ceil((yb - ya) / cell);
    const int64_t nx = (isnan(ex) ? 0 : (int64_t)ex) + 1;
    const int64_t ny = (isnan(ey) ? 0 : (int64_t)ey) + 1;
    const int64_t nc = nx * ny;
    float *top = (float *)malloc((size_t)nc * 3 * sizeof(float));
    uint16_t *count = (uint16_t *)calloc((size_t)nc, sizeof(uint16_t));
    for (int64_t k = 0; k < 3 * nc; ++k) top[k] = -INFINITY;
    for (int64_t p = 0; p < n; ++p) {
        const double i = floor(((double)xyz[3 * p] - xa) / cell);
        const double j = floor(((double)xyz[3 * p + 1] - ya) / cell);
        if (i < 0.0 || j < 0.0 || i >= (double)nx || j >= (double)ny) continue;
        // a NaN coordinate passes the test above; Rust's `as usize` makes it 0
        const int64_t k = (int64_t)(isnan(j) ? 0.0 : j) * nx + (int64_t)(isnan(i) ? 0.0 : i);
        if (count[k] < 65535) ++count[k];
        float *t = top + 3 * k;
        const float z = xyz[3 * p + 2];
        if (z > t[0]) {
            t[2] = t[1];
            t[1] = t[0];
            t[0] = z;
        } else if (z > t[1]) {
            t[2] = t[1];
            t[1] = z;
        } else if (z > t[2]) {
            t[2] = z;
        }
    }
    float *h = (float *)malloc((size_t)nc * sizeof(float));
    for (int64_t k = 0; k < nc; ++k)
        h[k] = count[k] == 0 ? NAN : (count[k] == 1 ? top[3 * k] : top[3 * k + 1]);
    free(top);
    /* despike (dsm.rs:202-216) */
    float *snap = (float *)malloc((size_t)nc * sizeof(float));
    memcpy(snap, h, (size_t)nc * sizeof(float));
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            const int64_t k = j * nx + i;
            if (!isfinite(snap[k])) continue;
            float m;
            if (dsm_med(snap, nx, ny, i, j, 2, &m)) {
                if (fabsf(snap[k] - m) > 3.0f && count[k] <= 2) h[k] = m;
            } else if (count[k] <= 2) {
                h[k] = NAN;
            }
        }
    free(count);
    orc_remove_islands(h, nx, ny, (int64_t)(25.0 / (cell * cell)));
    orc_fill_enclosed(h, nx, ny);
    /* small holes (dsm.rs:226-257): 8 passes, 8 neighbours, at least 3 */
    static const int64_t d8[8][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1},
                                     {-1, -1}, {1, 1}, {-1, 1}, {1, -1}};
    for (int pass = 0; pass < 8; ++pass) {
        memcpy(snap, h, (size_t)nc * sizeof(float));
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                const int64_t k = j * nx + i;
                if (isfinite(snap[k])) continue;
                float s = 0.0f;
                int c = 0;
                for (int q = 0; q < 8; ++q) {
                    const int64_t a = i + d8[q][0], b = j + d8[q][1];
                    if (a >= 0 && b >= 0 && a < nx && b < ny) {
                        const float x = snap[b * nx + a];
                        if (isfinite(x)) {
                            s += x;
                            ++c;
                        }
                    }
                }
                if (c >= 3) h[k] = s / (float)c;
            }
    }
    /* 3x3 median (dsm.rs:259-270) */
    memcpy(snap, h, (size_t)nc * sizeof(float));
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            const int64_t k = j * nx + i;
            float m;
            if (isfinite(snap[k]) && dsm_med(snap, nx, ny, i, j, 1, &m)) h[k] = m;
        }
    free(snap);
    float zmax = -FLT_MAX;
    for (int64_t k = 0; k < nc; ++k)
        if (isfinite(h[k])) zmax = fmaxf(zmax, h[k]);
    out->x0 = xa;
    out->y0 = ya;
    out->cell = cell;
    out->nx = nx;
    out->ny = ny;
    out->zmax = zmax;
    *hout = h;
    return 0;
}

/* dsm.rs:82-119 */
int orc_dsm_height(const orc_dsm *d, const float *h, double x, double y, float *out) {
    const double fx = (x - d->x0) / d->cell - 0.5;
    const double fy = (y - d->y0) / d->cell - 0.5;
    if (fx < -0.5 || fy < -0.5 || fx > (double)d->nx - 0.5 || fy > (double)d->ny - 0.5) return 0;
    double ffx = floor(fx), ffy = floor(fy);
    if (ffx < 0.0) ffx = 0.0;
    if (ffy < 0.0) ffy = 0.0;
    int64_t i0 = (int64_t)ffx, j0 = (int64_t)ffy;
    if (i0 > d->nx - 1) i0 = d->nx - 1;
    if (j0 > d->ny - 1) j0 = d->ny - 1;
    const int64_t i1 = i0 + 1 < d->nx - 1 ? i0 + 1 : d->nx - 1;
    const int64_t j1 = j0 + 1 < d->ny - 1 ? j0 + 1 : d->ny - 1;
    double ax = fx - (double)i0, ay = fy - (double)j0;
    ax = ax < 0.0 ? 0.0 : (ax > 1.0 ? 1.0 : ax);
    ay = ay < 0.0 ? 0.0 : (ay > 1.0 ? 1.0 : ay);
    const float fax = (float)ax, fay = (float)ay;
    const float v[4] = {h[j0 * d->nx + i0], h[j0 * d->nx + i1], h[j1 * d->nx + i0],
                        h[j1 * d->nx + i1]};
    const float w[4] = {(1.0f - fax) * (1.0f - fay), fax * (1.0f - fay), (1.0f - fax) * fay,
                        fax * fay};
    float s = 0.0f, ws = 0.0f;
    for (int k = 0; k < 4; ++k)
        if (isfinite(v[k])) {
            s += v[k] * w[k];
            ws += w[k];
        }
    if (ws > 1e-6f) {
        *out = s / ws;
        return 1;
    }
    for (int k = 0; k < 4; ++k)
        if (isfinite(v[k])) {
            *out = v[k];
            return 1;
        }
    return 0;
}
