// Seeded synthetic aerial survey on the GPU: the generator of synth.c (the
// same fractal heightfield, box extrusions, band-limited texture and sensor
// noise) for surveys too large to render on the host (thousands of views of
// several megapixels). Workload generator for the benchmark and the tests; it
// is not on the depth path and not part of libaerial_mvs.
//
// Heights are computed in f64 exactly as synth.c does and stored as f32 on
// the device. A pixel's ray is intersected exactly with the bilinear surface
// (f64): a DDA over 16x16-cell tiles skips every tile whose highest node lies
// below the ray's segment over it, and a DDA over the cells of the remaining
// tiles brackets the first sign change of z(t) - H(x(t), y(t)), which is a
// quadratic in t inside a cell, refined by bisection. Outside the grid there
// is no surface (synth.c holds the border height there): scenes carry margins
// wide enough that every ray of their views lands inside. Cameras follow the
// reference's convention (geometry.rs:1-5): X_cam = R (X - C), camera z
// forward; the depth of a pixel is the camera z of its hit.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>

namespace {

constexpr int kTile = 16;  // cells per tile side

struct Scene {
    double x0 = 0, y0 = 0, step = 0;
    int32_t nx = 0, ny = 0;      // nodes
    int32_t tnx = 0, tny = 0;    // tiles
    float *hgt = nullptr;        // nx*ny
    float *tmax = nullptr;       // tnx*tny, highest node of each tile
    double hmin = 0, hmax = 0;
    uint64_t seed = 0;
    int32_t tex_octaves = 0;
    double tex_lambda0 = 0;
    uint8_t *img = nullptr;      // render staging
    float *depth = nullptr;
    int64_t stage_px = 0;
    double *red = nullptr;       // reduction scratch
};

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ inline double lattice(int64_t ix, int64_t iy, uint64_t salt) {
    const uint64_t h = mix64((uint64_t)ix * 0x632BE59BD9B4E019ull ^ mix64((uint64_t)iy + salt));
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

__device__ inline double quintic(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

__device__ double value_noise(double x, double y, double lambda, uint64_t salt) {
    const double gx = x / lambda, gy = y / lambda;
    const double fx = floor(gx), fy = floor(gy);
    const int64_t ix = (int64_t)fx, iy = (int64_t)fy;
    const double tx = quintic(gx - fx), ty = quintic(gy - fy);
    const double a = lattice(ix, iy, salt), b = lattice(ix + 1, iy, salt);
    const double c = lattice(ix, iy + 1, salt), d = lattice(ix + 1, iy + 1, salt);
    const double top = a + (b - a) * tx, bot = c + (d - c) * tx;
    return top + (bot - top) * ty;
}

__global__ void k_fbm(double x0, double y0, double step, int32_t nx, int32_t ny,
                      double base_lambda, int32_t octaves, double persistence, uint64_t seed,
                      double *out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nx * ny) return;
    const int32_t iy = (int32_t)(k / nx), ix = (int32_t)(k - (int64_t)iy * nx);
    const double x = x0 + ix * step, y = y0 + iy * step;
    double amp = 1.0, lam = base_lambda, v = 0.0;
    for (int32_t o = 0; o < octaves; ++o) {
        v += amp * value_noise(x, y, lam, seed * 31u + (uint64_t)o);
        amp *= persistence;
        lam *= 0.5;
    }
    out[k] = v;
}

// min and max of n values: wave reductions, then atomics on order-preserving keys
__host__ __device__ inline unsigned long long okey(double v) {
    unsigned long long b;
    memcpy(&b, &v, 8);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double oval(unsigned long long k) {
    const unsigned long long b =
        (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    double v;
    memcpy(&v, &b, 8);
    return v;
}

template <class T>
__global__ void k_minmax(const T *v, int64_t n, unsigned long long *mm) {
    double lo = 1e300, hi = -1e300;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const double x = (double)v[k];
        lo = x < lo ? x : lo;
        hi = x > hi ? x : hi;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off));
        hi = fmax(hi, __shfl_xor(hi, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&mm[0], okey(lo));
        atomicMax(&mm[1], okey(hi));
    }
}

__global__ void k_normalise(const double *v, int64_t n, double lo, double scale, float *h) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) h[k] = (float)((v[k] - lo) * scale);
}

// one box: the highest node of its range, then the range raised to it + height
__global__ void k_box_max(const float *h, int32_t nx, int32_t i0, int32_t i1, int32_t j0,
                          int32_t j1, unsigned long long *mm) {
    double hi = -1e300;
    const int32_t bw = i1 - i0 + 1;
    const int64_t n = (int64_t)bw * (j1 - j0 + 1);
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) {
        const int32_t j = j0 + (int32_t)(k / bw), i = i0 + (int32_t)(k % bw);
        const double x = (double)h[(int64_t)j * nx + i];
        hi = x > hi ? x : hi;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) hi = fmax(hi, __shfl_xor(hi, off));
    if ((threadIdx.x & 63) == 0) atomicMax(&mm[1], okey(hi));
}

__global__ void k_box_fill(float *h, int32_t nx, int32_t i0, int32_t i1, int32_t j0, int32_t j1,
                           const unsigned long long *mm, double add) {
    const int32_t bw = i1 - i0 + 1;
    const int64_t n = (int64_t)bw * (j1 - j0 + 1);
    const float top = (float)(oval(mm[1]) + add);
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t j = j0 + (int32_t)(k / bw), i = i0 + (int32_t)(k % bw);
        h[(int64_t)j * nx + i] = top;
    }
}

__global__ void k_tile_max(const float *h, int32_t nx, int32_t ny, int32_t tnx, int32_t tny,
                           float *tmax) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)tnx * tny) return;
    const int32_t ty = (int32_t)(t / tnx), tx = (int32_t)(t - (int64_t)ty * tnx);
    const int32_t i1 = std::min(tx * kTile + kTile, nx - 1), j1 = std::min(ty * kTile + kTile, ny - 1);
    float m = -INFINITY;
    for (int32_t j = ty * kTile; j <= j1; ++j)
        for (int32_t i = tx * kTile; i <= i1; ++i) m = fmaxf(m, h[(int64_t)j * nx + i]);
    tmax[t] = m;
}

struct Ray {
    double cz, dz;        // z(t) = cz + t dz
    double gx0, gy0;      // grid coordinates at t = 0
    double dgx, dgy;      // grid coordinates per unit t
};

// f(t) = z(t) - H(t) inside cell (i, j)
__device__ inline double cell_f(const Ray &r, const float *h, int32_t nx, int32_t i, int32_t j,
                                double t) {
    const float *p = h + (int64_t)j * nx + i;
    const double h00 = p[0], h10 = p[1], h01 = p[nx], h11 = p[nx + 1];
    const double u = r.gx0 + t * r.dgx - i, v = r.gy0 + t * r.dgy - j;
    const double top = h00 + (h10 - h00) * u, bot = h01 + (h11 - h01) * u;
    return r.cz + t * r.dz - (top + (bot - top) * v);
}

// First t in [ta, tb] with f(t) <= 0 in cell (i, j), or -1. f is a quadratic
// in t here; its minimum over the segment is at an end or at the vertex.
__device__ double cell_hit(const Ray &r, const float *h, int32_t nx, int32_t i, int32_t j,
                           double ta, double tb) {
    const float *p = h + (int64_t)j * nx + i;
    const double cmax = fmax(fmax((double)p[0], (double)p[1]), fmax((double)p[nx], (double)p[nx + 1]));
    if (fmin(r.cz + ta * r.dz, r.cz + tb * r.dz) > cmax) return -1.0;
    const double fa = cell_f(r, h, nx, i, j, ta);
    if (fa <= 0.0) return ta;
    double hi = -1.0;
    if (cell_f(r, h, nx, i, j, tb) <= 0.0) {
        hi = tb;
    } else {
        // f = A t^2 + B t + c: A = -e3 a b with e3 the bilinear twist
        const double e3 = (double)p[0] - p[1] - p[nx] + p[nx + 1];
        const double A = -e3 * r.dgx * r.dgy;
        const double e1 = (double)p[1] - p[0], e2 = (double)p[nx] - p[0];
        const double u0 = r.gx0 - i, v0 = r.gy0 - j;
        const double B = r.dz - (e1 * r.dgx + e2 * r.dgy + e3 * (u0 * r.dgy + v0 * r.dgx));
        if (A > 0.0) {
            const double tv = -B / (2.0 * A);
            if (tv > ta && tv < tb && cell_f(r, h, nx, i, j, tv) <= 0.0) hi = tv;
        }
    }
    if (hi < 0.0) return -1.0;
    double lo = ta;  // f(lo) > 0 >= f(hi)
    for (int k = 0; k < 60 && hi - lo > 1e-12 * fmax(1.0, hi); ++k) {
        const double mid = 0.5 * (lo + hi);
        if (cell_f(r, h, nx, i, j, mid) > 0.0)
            lo = mid;
        else
            hi = mid;
    }
    return hi;
}

// DDA over cells [c0x, c1x] x [c0y, c1y] (inclusive) from t0 to t1 (units:
// cells of `scale` grid units). Calls visit(i, j, ta, tb) until it returns
// true; returns whether one did.
template <class F>
__device__ bool dda(const Ray &r, double scale, int32_t c0x, int32_t c1x, int32_t c0y,
                    int32_t c1y, double t0, double t1, F visit) {
    const double px = (r.gx0 + t0 * r.dgx) / scale, py = (r.gy0 + t0 * r.dgy) / scale;
    int32_t i = (int32_t)floor(px), j = (int32_t)floor(py);
    i = std::min(std::max(i, c0x), c1x);
    j = std::min(std::max(j, c0y), c1y);
    const double dx = r.dgx / scale, dy = r.dgy / scale;
    const int32_t sx = dx > 0 ? 1 : -1, sy = dy > 0 ? 1 : -1;
    // t at which the ray leaves cell i (j) along x (y)
    auto next_x = [&](int32_t ii) -> double {
        if (dx == 0.0) return HUGE_VAL;
        const double edge = (double)(dx > 0 ? ii + 1 : ii) * scale;
        return (edge - r.gx0) / r.dgx;
    };
    auto next_y = [&](int32_t jj) -> double {
        if (dy == 0.0) return HUGE_VAL;
        const double edge = (double)(dy > 0 ? jj + 1 : jj) * scale;
        return (edge - r.gy0) / r.dgy;
    };
    double tx = next_x(i), ty = next_y(j);
    double t = t0;
    const int32_t cap = (c1x - c0x + 1) + (c1y - c0y + 1) + 4;
    for (int32_t it = 0; it < cap && t < t1; ++it) {
        const double te = fmin(fmin(tx, ty), t1);
        if (te >= t && visit(i, j, t, te)) return true;
        t = te;
        if (tx <= ty) {
            i += sx;
            tx = next_x(i);
        } else {
            j += sy;
            ty = next_y(j);
        }
        if (i < c0x || i > c1x || j < c0y || j > c1y) break;
    }
    return false;
}

__device__ double texture(double x, double y, uint64_t seed, double lambda0, int32_t octaves) {
    double v = 0.0, lam = lambda0;
    for (int32_t o = 0; o < octaves; ++o) {
        v += value_noise(x, y, lam, seed * 977u + 1000u + (uint64_t)o);
        lam *= 2.0;
    }
    v = v / (0.5 * octaves);
    if (v > 1.0) v = 1.0;
    if (v < -1.0) v = -1.0;
    return 128.0 + 112.0 * v;
}

__device__ double gauss(uint64_t key) {
    const uint64_t a = mix64(key), b = mix64(key ^ 0xD1B54A32D192ED03ull);
    const double u1 = ((double)(a >> 11) + 0.5) / 9007199254740992.0;
    const double u2 = (double)(b >> 11) / 9007199254740992.0;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

struct RenderArgs {
    const float *hgt, *tmax;
    double x0, y0, step, hmin, hmax;
    int32_t nx, ny, tnx, tny;
    int32_t w, h;
    double f, cx, cy;
    double R[9], C[3];
    uint64_t seed, noise_seed;
    double noise_dn, tex_lambda0;
    int32_t tex_octaves;
    uint8_t *img;
    float *depth;
};

__global__ __launch_bounds__(256) void k_render(RenderArgs a) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)a.w * a.h) return;
    const int32_t y = (int32_t)(k / a.w), x = (int32_t)(k - (int64_t)y * a.w);
    const double qx = (x + 0.5 - a.cx) / a.f, qy = (y + 0.5 - a.cy) / a.f;
    double D[3];
    for (int r = 0; r < 3; ++r) D[r] = a.R[0 * 3 + r] * qx + a.R[1 * 3 + r] * qy + a.R[2 * 3 + r];
    Ray ray;
    ray.cz = a.C[2];
    ray.dz = D[2];
    ray.gx0 = (a.C[0] - a.x0) / a.step;
    ray.gy0 = (a.C[1] - a.y0) / a.step;
    ray.dgx = D[0] / a.step;
    ray.dgy = D[1] / a.step;
    // the part of the ray inside the height slab and over the grid
    double t0 = 0.0, t1 = INFINITY;
    bool miss = false;
    if (D[2] < 0.0) {
        t0 = fmax(t0, (a.C[2] - a.hmax) / (-D[2]));
        t1 = fmin(t1, (a.C[2] - a.hmin) / (-D[2]) + 1e-9);
    } else {
        miss = a.C[2] > a.hmax;  // the camera sits above every height
    }
    const double gmax[2] = {(double)(a.nx - 1), (double)(a.ny - 1)};
    const double g0[2] = {ray.gx0, ray.gy0}, dg[2] = {ray.dgx, ray.dgy};
    for (int d = 0; d < 2 && !miss; ++d) {
        if (dg[d] == 0.0) {
            miss = g0[d] < 0.0 || g0[d] > gmax[d];
        } else {
            double ta = (0.0 - g0[d]) / dg[d], tb = (gmax[d] - g0[d]) / dg[d];
            if (ta > tb) {
                const double t = ta;
                ta = tb;
                tb = t;
            }
            t0 = fmax(t0, ta);
            t1 = fmin(t1, tb);
        }
    }
    double hit = -1.0;
    if (!miss && t0 < t1 && isfinite(t1)) {
        const float *hg = a.hgt;
        const float *tm = a.tmax;
        const int32_t nx = a.nx, tnx = a.tnx;
        const int32_t ncx = a.nx - 2, ncy = a.ny - 2;  // last cell index
        dda(ray, (double)kTile, 0, a.tnx - 1, 0, a.tny - 1, t0, t1,
            [&](int32_t ti, int32_t tj, double ta, double tb) {
                const double zlo = fmin(ray.cz + ta * ray.dz, ray.cz + tb * ray.dz);
                if (zlo > (double)tm[(int64_t)tj * tnx + ti]) return false;
                const int32_t c0x = ti * kTile, c0y = tj * kTile;
                const int32_t c1x = std::min(c0x + kTile - 1, ncx);
                const int32_t c1y = std::min(c0y + kTile - 1, ncy);
                return dda(ray, 1.0, c0x, c1x, c0y, c1y, ta, tb,
                           [&](int32_t i, int32_t j, double ca, double cb) {
                               const double th = cell_hit(ray, hg, nx, i, j, ca, cb);
                               if (th < 0.0) return false;
                               hit = th;
                               return true;
                           });
            });
    }
    double g = 0.0, z = 0.0;
    if (hit > 0.0) {
        const double X = a.C[0] + hit * D[0], Y = a.C[1] + hit * D[1];
        g = texture(X, Y, a.seed, a.tex_lambda0, a.tex_octaves);
        z = hit;  // D has unit camera z
    }
    if (a.noise_dn > 0.0) g += a.noise_dn * gauss(a.noise_seed * 0x100000001B3ull + (uint64_t)k);
    g = floor(g + 0.5);
    a.img[k] = (uint8_t)(g < 0.0 ? 0.0 : (g > 255.0 ? 255.0 : g));
    if (a.depth) a.depth[k] = (float)z;
}

unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

int minmax_f32(Scene *s, double *lo, double *hi) {
    unsigned long long mm[2] = {okey(1e300), okey(-1e300)};
    unsigned long long *d = (unsigned long long *)s->red;
    if (hipMemcpy(d, mm, 16, hipMemcpyHostToDevice) != hipSuccess) return -1;
    const int64_t n = (int64_t)s->nx * s->ny;
    hipLaunchKernelGGL(k_minmax<float>, dim3(4096), dim3(256), 0, 0, s->hgt, n, d);
    if (hipMemcpy(mm, d, 16, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    *lo = oval(mm[0]);
    *hi = oval(mm[1]);
    return 0;
}

int finish(Scene *s) {
    if (minmax_f32(s, &s->hmin, &s->hmax)) return -1;
    hipLaunchKernelGGL(k_tile_max, dim3(nblk((int64_t)s->tnx * s->tny)), dim3(256), 0, 0,
                       s->hgt, s->nx, s->ny, s->tnx, s->tny, s->tmax);
    return hipDeviceSynchronize() == hipSuccess ? 0 : -1;
}

}  // namespace

extern "C" {

typedef struct Scene sg_scene;

// Heightfield over [x0, x0+width] x [y0, y0+height] (synth_scene_create):
// fBm of value noise rescaled to `relief` metres with its minimum at 0.
int sg_scene_create(double x0, double y0, double width, double height, double step,
                    double base_lambda, int32_t octaves, double persistence, double relief,
                    uint64_t seed, double tex_lambda0, int32_t tex_octaves, sg_scene **out) {
    if (!out || !(step > 0) || !(width > 0) || !(height > 0)) return -1;
    *out = nullptr;
    Scene *s = new Scene();
    s->x0 = x0;
    s->y0 = y0;
    s->step = step;
    s->nx = (int32_t)ceil(width / step) + 1;
    s->ny = (int32_t)ceil(height / step) + 1;
    s->tnx = (s->nx - 2) / kTile + 1;
    s->tny = (s->ny - 2) / kTile + 1;
    s->seed = seed;
    s->tex_lambda0 = tex_lambda0;
    s->tex_octaves = tex_octaves;
    const int64_t n = (int64_t)s->nx * s->ny;
    double *tmp = nullptr;
    bool ok = hipMalloc((void **)&s->hgt, (size_t)n * 4) == hipSuccess &&
              hipMalloc((void **)&s->tmax, (size_t)s->tnx * s->tny * 4) == hipSuccess &&
              hipMalloc((void **)&s->red, 64) == hipSuccess &&
              hipMalloc((void **)&tmp, (size_t)n * 8) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_fbm, dim3(nblk(n)), dim3(256), 0, 0, x0, y0, step, s->nx, s->ny,
                           base_lambda, octaves, persistence, seed, tmp);
        unsigned long long mm[2] = {okey(1e300), okey(-1e300)};
        ok = hipMemcpy(s->red, mm, 16, hipMemcpyHostToDevice) == hipSuccess;
        hipLaunchKernelGGL(k_minmax<double>, dim3(4096), dim3(256), 0, 0, tmp, n,
                           (unsigned long long *)s->red);
        ok = ok && hipMemcpy(mm, s->red, 16, hipMemcpyDeviceToHost) == hipSuccess;
        const double lo = oval(mm[0]), hi = oval(mm[1]);
        hipLaunchKernelGGL(k_normalise, dim3(nblk(n)), dim3(256), 0, 0, tmp, n, lo,
                           relief / (hi - lo), s->hgt);
        ok = ok && hipDeviceSynchronize() == hipSuccess && finish(s) == 0;
    }
    if (tmp) (void)hipFree(tmp);
    if (!ok) {
        (void)hipFree(s->hgt);
        (void)hipFree(s->tmax);
        (void)hipFree(s->red);
        delete s;
        return -2;
    }
    *out = s;
    return 0;
}

// Boxes (cx, cy, half_x, half_y, height above the local highest node), n x 5,
// raised in order (synth_scene_add_boxes).
int sg_scene_add_boxes(sg_scene *s, const double *boxes, int32_t n) {
    if (!s || (n > 0 && !boxes)) return -1;
    unsigned long long *d = (unsigned long long *)s->red;
    for (int32_t b = 0; b < n; ++b) {
        const double *bx = boxes + 5 * b;
        int32_t i0 = (int32_t)floor((bx[0] - bx[2] - s->x0) / s->step);
        int32_t i1 = (int32_t)ceil((bx[0] + bx[2] - s->x0) / s->step);
        int32_t j0 = (int32_t)floor((bx[1] - bx[3] - s->y0) / s->step);
        int32_t j1 = (int32_t)ceil((bx[1] + bx[3] - s->y0) / s->step);
        i0 = std::max(i0, 0);
        j0 = std::max(j0, 0);
        i1 = std::min(i1, s->nx - 1);
        j1 = std::min(j1, s->ny - 1);
        if (i0 > i1 || j0 > j1) continue;
        unsigned long long mm[2] = {okey(1e300), okey(-1e300)};
        if (hipMemcpy(d, mm, 16, hipMemcpyHostToDevice) != hipSuccess) return -2;
        hipLaunchKernelGGL(k_box_max, dim3(1), dim3(256), 0, 0, s->hgt, s->nx, i0, i1, j0, j1, d);
        const int64_t cnt = (int64_t)(i1 - i0 + 1) * (j1 - j0 + 1);
        hipLaunchKernelGGL(k_box_fill, dim3((unsigned)std::min<int64_t>(nblk(cnt), 1024)),
                           dim3(256), 0, 0, s->hgt, s->nx, i0, i1, j0, j1, d, bx[4]);
    }
    return finish(s) == 0 ? 0 : -2;
}

void sg_scene_info(const sg_scene *s, double out[6]) {
    out[0] = s->hmin;
    out[1] = s->hmax;
    out[2] = 0.0;
    out[3] = s->step;
    out[4] = (double)s->nx;
    out[5] = (double)s->ny;
}

// Render one pinhole view into host memory (synth_render without lens
// distortion). R: world->camera row-major; img: w*h u8; depth (nullable):
// w*h f32 camera z of each pixel-centre hit, 0 where the ray misses.
int sg_render(sg_scene *s, int32_t w, int32_t h, double f, double cx, double cy,
              const double R[9], const double C[3], uint64_t noise_seed, double noise_dn,
              uint8_t *img, float *depth) {
    if (!s || w < 1 || h < 1 || !img || !R || !C) return -1;
    const int64_t n = (int64_t)w * h;
    if (n > s->stage_px) {
        (void)hipFree(s->img);
        (void)hipFree(s->depth);
        s->img = nullptr;
        s->depth = nullptr;
        if (hipMalloc((void **)&s->img, (size_t)n) != hipSuccess ||
            hipMalloc((void **)&s->depth, (size_t)n * 4) != hipSuccess) {
            s->stage_px = 0;
            return -2;
        }
        s->stage_px = n;
    }
    RenderArgs a;
    a.hgt = s->hgt;
    a.tmax = s->tmax;
    a.x0 = s->x0;
    a.y0 = s->y0;
    a.step = s->step;
    a.hmin = s->hmin;
    a.hmax = s->hmax;
    a.nx = s->nx;
    a.ny = s->ny;
    a.tnx = s->tnx;
    a.tny = s->tny;
    a.w = w;
    a.h = h;
    a.f = f;
    a.cx = cx;
    a.cy = cy;
    memcpy(a.R, R, sizeof a.R);
    memcpy(a.C, C, sizeof a.C);
    a.seed = s->seed;
    a.noise_seed = noise_seed;
    a.noise_dn = noise_dn;
    a.tex_lambda0 = s->tex_lambda0;
    a.tex_octaves = s->tex_octaves;
    a.img = s->img;
    a.depth = depth ? s->depth : nullptr;
    hipLaunchKernelGGL(k_render, dim3(nblk(n)), dim3(256), 0, 0, a);
    if (hipGetLastError() != hipSuccess) return -2;
    if (hipMemcpy(img, s->img, (size_t)n, hipMemcpyDeviceToHost) != hipSuccess) return -2;
    if (depth && hipMemcpy(depth, s->depth, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return -2;
    return 0;
}

void sg_scene_destroy(sg_scene *s) {
    if (!s) return;
    (void)hipDeviceSynchronize();
    (void)hipFree(s->hgt);
    (void)hipFree(s->tmax);
    (void)hipFree(s->red);
    (void)hipFree(s->img);
    (void)hipFree(s->depth);
    delete s;
}

}  // extern "C"
