/*
 * Seeded synthetic aerial survey: fractal heightfield (optionally with box
 * extrusions), band-limited value-noise texture, pinhole ray-cast rendering
 * to 8-bit grey with per-view Gaussian sensor noise, and the FP64 camera-z
 * depth of every pixel's hit point (the ground truth that depth errors are
 * measured against).
 *
 * Workload generator for the tests and the benchmark; it is not on the depth
 * path. Cameras follow the reference's convention (geometry.rs:1-5, 47-93):
 * world = local east/north/up, camera x right, y down, z forward, R maps
 * world to camera and X_cam = R (X - C).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double x0, y0;   /* world position of grid node (0, 0) */
    double step;     /* grid spacing, m */
    int32_t nx, ny;  /* nodes */
    double *hgt;     /* nx*ny heights, m */
    double hmin, hmax;
    double slope;    /* Lipschitz bound of the bilinear surface */
    uint64_t seed;
    int32_t tex_octaves;
    double tex_lambda0; /* finest texture wavelength, m */
} synth_scene;

static uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* lattice value in [-1, 1] */
static double lattice(int64_t ix, int64_t iy, uint64_t salt) {
    uint64_t h = mix64((uint64_t)ix * 0x632BE59BD9B4E019ull ^ mix64((uint64_t)iy + salt));
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

static double quintic(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

/* smooth value noise with lattice spacing `lambda` */
static double value_noise(double x, double y, double lambda, uint64_t salt) {
    double gx = x / lambda, gy = y / lambda;
    double fx = floor(gx), fy = floor(gy);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy;
    double tx = quintic(gx - fx), ty = quintic(gy - fy);
    double a = lattice(ix, iy, salt), b = lattice(ix + 1, iy, salt);
    double c = lattice(ix, iy + 1, salt), d = lattice(ix + 1, iy + 1, salt);
    double top = a + (b - a) * tx, bot = c + (d - c) * tx;
    return top + (bot - top) * ty;
}

typedef struct {
    synth_scene *s;
    double base_lambda;
    int32_t octaves;
    double persistence;
    int32_t j0, j1;
} hf_job;

static void *hf_rows(void *p) {
    hf_job *j = (hf_job *)p;
    synth_scene *s = j->s;
    for (int32_t iy = j->j0; iy < j->j1; ++iy)
        for (int32_t ix = 0; ix < s->nx; ++ix) {
            double x = s->x0 + ix * s->step, y = s->y0 + iy * s->step;
            double amp = 1.0, lam = j->base_lambda, v = 0.0;
            for (int32_t o = 0; o < j->octaves; ++o) {
                v += amp * value_noise(x, y, lam, s->seed * 31u + (uint64_t)o);
                amp *= j->persistence;
                lam *= 0.5;
            }
            s->hgt[(size_t)iy * s->nx + ix] = v;
        }
    return NULL;
}

static void run_threads(void *(*fn)(void *), void *jobs, size_t job_size, int32_t n) {
    pthread_t *th = (pthread_t *)calloc((size_t)n, sizeof(pthread_t));
    for (int32_t t = 1; t < n; ++t) pthread_create(&th[t], NULL, fn, (char *)jobs + job_size * t);
    fn(jobs);
    for (int32_t t = 1; t < n; ++t) pthread_join(th[t], NULL);
    free(th);
}

static void finish_scene(synth_scene *s) {
    double lo = 1e300, hi = -1e300, g = 0.0;
    for (int32_t iy = 0; iy < s->ny; ++iy)
        for (int32_t ix = 0; ix < s->nx; ++ix) {
            double v = s->hgt[(size_t)iy * s->nx + ix];
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
            if (ix + 1 < s->nx) {
                double d = fabs(s->hgt[(size_t)iy * s->nx + ix + 1] - v);
                g = d > g ? d : g;
            }
            if (iy + 1 < s->ny) {
                double d = fabs(s->hgt[(size_t)(iy + 1) * s->nx + ix] - v);
                g = d > g ? d : g;
            }
        }
    s->hmin = lo;
    s->hmax = hi;
    /* |grad| of a bilinear cell <= sqrt(2) * max edge slope */
    s->slope = sqrt(2.0) * g / s->step;
}

/* Heightfield over [x0, x0+width] x [y0, y0+height]: fBm of smooth value noise
 * (octaves, base wavelength, persistence), rescaled to `relief` metres peak to
 * peak with its minimum at z = 0. */
synth_scene *synth_scene_create(double x0, double y0, double width, double height, double step,
                                double base_lambda, int32_t octaves, double persistence,
                                double relief, uint64_t seed, double tex_lambda0,
                                int32_t tex_octaves, int32_t nthreads) {
    synth_scene *s = (synth_scene *)calloc(1, sizeof(synth_scene));
    s->x0 = x0;
    s->y0 = y0;
    s->step = step;
    s->nx = (int32_t)ceil(width / step) + 1;
    s->ny = (int32_t)ceil(height / step) + 1;
    s->seed = seed;
    s->tex_lambda0 = tex_lambda0;
    s->tex_octaves = tex_octaves;
    s->hgt = (double *)malloc(sizeof(double) * (size_t)s->nx * (size_t)s->ny);
    if (nthreads < 1) nthreads = 1;
    hf_job *jobs = (hf_job *)calloc((size_t)nthreads, sizeof(hf_job));
    for (int32_t t = 0; t < nthreads; ++t) {
        jobs[t].s = s;
        jobs[t].base_lambda = base_lambda;
        jobs[t].octaves = octaves;
        jobs[t].persistence = persistence;
        jobs[t].j0 = (int32_t)((int64_t)s->ny * t / nthreads);
        jobs[t].j1 = (int32_t)((int64_t)s->ny * (t + 1) / nthreads);
    }
    run_threads(hf_rows, jobs, sizeof(hf_job), nthreads);
    free(jobs);
    finish_scene(s);
    double lo = s->hmin, scale = relief / (s->hmax - s->hmin);
    for (size_t i = 0; i < (size_t)s->nx * s->ny; ++i) s->hgt[i] = (s->hgt[i] - lo) * scale;
    finish_scene(s);
    return s;
}

/* Raise axis-aligned boxes (cx, cy, half_x, half_y, height above local max
 * ground) onto the grid: building-like extrusions. boxes: n x 5. */
void synth_scene_add_boxes(synth_scene *s, const double *boxes, int32_t n) {
    for (int32_t b = 0; b < n; ++b) {
        const double *bx = boxes + 5 * b;
        int32_t i0 = (int32_t)floor((bx[0] - bx[2] - s->x0) / s->step);
        int32_t i1 = (int32_t)ceil((bx[0] + bx[2] - s->x0) / s->step);
        int32_t j0 = (int32_t)floor((bx[1] - bx[3] - s->y0) / s->step);
        int32_t j1 = (int32_t)ceil((bx[1] + bx[3] - s->y0) / s->step);
        i0 = i0 < 0 ? 0 : i0;
        j0 = j0 < 0 ? 0 : j0;
        i1 = i1 >= s->nx ? s->nx - 1 : i1;
        j1 = j1 >= s->ny ? s->ny - 1 : j1;
        double top = -1e300;
        for (int32_t j = j0; j <= j1; ++j)
            for (int32_t i = i0; i <= i1; ++i) {
                double v = s->hgt[(size_t)j * s->nx + i];
                top = v > top ? v : top;
            }
        top += bx[4];
        for (int32_t j = j0; j <= j1; ++j)
            for (int32_t i = i0; i <= i1; ++i) s->hgt[(size_t)j * s->nx + i] = top;
    }
    finish_scene(s);
}

void synth_scene_destroy(synth_scene *s) {
    if (!s) return;
    free(s->hgt);
    free(s);
}

void synth_scene_info(const synth_scene *s, double out[6]) {
    out[0] = s->hmin;
    out[1] = s->hmax;
    out[2] = s->slope;
    out[3] = s->step;
    out[4] = (double)s->nx;
    out[5] = (double)s->ny;
}

/* bilinear height; outside the grid the border value is held */
double synth_height(const synth_scene *s, double x, double y) {
    double gx = (x - s->x0) / s->step, gy = (y - s->y0) / s->step;
    if (gx < 0) gx = 0;
    if (gy < 0) gy = 0;
    if (gx > s->nx - 1.000001) gx = s->nx - 1.000001;
    if (gy > s->ny - 1.000001) gy = s->ny - 1.000001;
    int32_t ix = (int32_t)gx, iy = (int32_t)gy;
    double tx = gx - ix, ty = gy - iy;
    const double *r0 = s->hgt + (size_t)iy * s->nx + ix, *r1 = r0 + s->nx;
    double top = r0[0] + (r0[1] - r0[0]) * tx, bot = r1[0] + (r1[1] - r1[0]) * tx;
    return top + (bot - top) * ty;
}

/* band-limited texture: equal-weight octaves from tex_lambda0 upwards, mapped
 * to [16, 240] grey levels */
double synth_texture(const synth_scene *s, double x, double y) {
    double v = 0.0, lam = s->tex_lambda0;
    for (int32_t o = 0; o < s->tex_octaves; ++o) {
        v += value_noise(x, y, lam, s->seed * 977u + 1000u + (uint64_t)o);
        lam *= 2.0;
    }
    v = v / (0.5 * s->tex_octaves);
    if (v > 1.0) v = 1.0;
    if (v < -1.0) v = -1.0;
    return 128.0 + 112.0 * v;
}

/* First hit of the ray C + t D (t > 0) with the surface; returns t or -1. */
static double cast(const synth_scene *s, const double C[3], const double D[3]) {
    double dxy = sqrt(D[0] * D[0] + D[1] * D[1]);
    double rate = fabs(D[2]) + s->slope * dxy;
    double t = 0.0;
    if (C[2] > s->hmax) {
        if (D[2] >= 0.0) return -1.0;
        t = (C[2] - s->hmax) / (-D[2]);
    }
    double tmax = 1e6;
    if (D[2] < 0.0) tmax = (C[2] - s->hmin) / (-D[2]) + 1.0;
    /* conservative marching: f(t) = z(t) - H(xy(t)) never crosses zero within
     * a step of f / rate */
    double f = 0.0;
    int it = 0;
    for (; it < 100000 && t < tmax; ++it) {
        f = C[2] + t * D[2] - synth_height(s, C[0] + t * D[0], C[1] + t * D[1]);
        if (f < 1e-5) break;
        double dt = f / rate;
        if (dt < 1e-5) dt = 1e-5;
        t += dt;
    }
    if (t >= tmax) return -1.0;
    /* bracket [t - small, t] and bisect to ~1e-12 m */
    double lo = t, hi = t;
    double flo = f;
    for (int k = 0; flo < 0.0 && k < 1000; ++k) {
        lo -= 1e-4;
        flo = C[2] + lo * D[2] - synth_height(s, C[0] + lo * D[0], C[1] + lo * D[1]);
    }
    double fhi = f;
    for (int k = 0; fhi > 0.0 && k < 1000; ++k) {
        hi += 1e-4;
        fhi = C[2] + hi * D[2] - synth_height(s, C[0] + hi * D[0], C[1] + hi * D[1]);
    }
    if (flo < 0.0 || fhi > 0.0) return t; /* grazing touch: keep the march point */
    for (int k = 0; k < 28; ++k) {
        double mid = 0.5 * (lo + hi);
        double fm = C[2] + mid * D[2] - synth_height(s, C[0] + mid * D[0], C[1] + mid * D[1]);
        if (fm > 0.0)
            lo = mid;
        else
            hi = mid;
    }
    return 0.5 * (lo + hi);
}

typedef struct {
    const synth_scene *s;
    int32_t w, h;
    double f, cx, cy, k1, k2;
    const double *R, *C;
    uint8_t *img;
    double *depth;
    uint64_t noise_seed;
    double noise_dn;
    int32_t y0, y1;
} render_job;

static double gauss(uint64_t key) {
    /* Box-Muller from two hashed uniforms */
    uint64_t a = mix64(key), b = mix64(key ^ 0xD1B54A32D192ED03ull);
    double u1 = ((double)(a >> 11) + 0.5) / 9007199254740992.0;
    double u2 = (double)(b >> 11) / 9007199254740992.0;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static void *render_rows(void *p) {
    render_job *j = (render_job *)p;
    const double *R = j->R;
    for (int32_t y = j->y0; y < j->y1; ++y)
        for (int32_t x = 0; x < j->w; ++x) {
            /* pixel centre -> normalised (undistort by fixed point, as the
             * reference's geometry.rs:34-45) */
            double qx = (x + 0.5 - j->cx) / j->f, qy = (y + 0.5 - j->cy) / j->f;
            double px = qx, py = qy;
            for (int k = 0; k < 8; ++k) {
                double r2 = px * px + py * py;
                double d = 1.0 + j->k1 * r2 + j->k2 * r2 * r2;
                px = qx / d;
                py = qy / d;
            }
            /* D = R^T (px, py, 1) */
            double D[3];
            for (int r = 0; r < 3; ++r) D[r] = R[0 * 3 + r] * px + R[1 * 3 + r] * py + R[2 * 3 + r];
            double t = cast(j->s, j->C, D);
            size_t i = (size_t)y * j->w + x;
            double g = 0.0, z = 0.0;
            if (t > 0.0) {
                double X = j->C[0] + t * D[0], Y = j->C[1] + t * D[1];
                g = synth_texture(j->s, X, Y);
                z = t; /* camera z of the hit: D has unit z in the camera frame */
            }
            if (j->noise_dn > 0.0) g += j->noise_dn * gauss(j->noise_seed * 0x100000001B3ull + i);
            g = floor(g + 0.5);
            j->img[i] = (uint8_t)(g < 0.0 ? 0.0 : (g > 255.0 ? 255.0 : g));
            if (j->depth) j->depth[i] = z;
        }
    return NULL;
}

/* Render one view. R: world->camera row-major, C: centre. img: w*h u8;
 * depth (nullable): w*h f64 camera-z of each pixel-centre hit (0 = sky). */
void synth_render(const synth_scene *s, int32_t w, int32_t h, double f, double cx, double cy,
                  double k1, double k2, const double R[9], const double C[3], uint64_t noise_seed,
                  double noise_dn, uint8_t *img, double *depth, int32_t nthreads) {
    if (nthreads < 1) nthreads = 1;
    render_job *jobs = (render_job *)calloc((size_t)nthreads, sizeof(render_job));
    for (int32_t t = 0; t < nthreads; ++t) {
        render_job *j = &jobs[t];
        j->s = s;
        j->w = w;
        j->h = h;
        j->f = f;
        j->cx = cx;
        j->cy = cy;
        j->k1 = k1;
        j->k2 = k2;
        j->R = R;
        j->C = C;
        j->img = img;
        j->depth = depth;
        j->noise_seed = noise_seed;
        j->noise_dn = noise_dn;
        j->y0 = (int32_t)((int64_t)h * t / nthreads);
        j->y1 = (int32_t)((int64_t)h * (t + 1) / nthreads);
    }
    run_threads(render_rows, jobs, sizeof(render_job), nthreads);
    free(jobs);
}
