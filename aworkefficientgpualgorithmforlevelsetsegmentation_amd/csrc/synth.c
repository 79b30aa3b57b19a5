/*
 * synth.c -- seeded synthetic volumes for the benchmark and test workloads.
 *
 * The reference ships no input volumes for its published runs (the BrainWeb
 * raws are missing blobs), so every workload here is generated from a seed.
 * Random numbers: SplitMix64 seeds one PCG32 stream per z-slice, Gaussians use
 * Box-Muller.  Per-slice streams keep the output independent of the thread
 * count.  All volumes are linear, x fastest, (z, y, x) row-major.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define SYN_PI 3.14159265358979323846

static uint64_t splitmix64(uint64_t* x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct {
    uint64_t state, inc;
    int has_spare;
    double spare;
} rng_t;

static uint32_t pcg32(rng_t* r)
{
    const uint64_t old = r->state;
    r->state = old * 6364136223846793005ull + r->inc;
    const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
    const uint32_t rot = (uint32_t)(old >> 59);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

static void rng_seed(rng_t* r, uint64_t seed, uint64_t stream)
{
    uint64_t sm = seed ^ (stream * 0xD1B54A32D192ED03ull);
    r->state = 0;
    r->inc = (splitmix64(&sm) << 1) | 1u;
    pcg32(r);
    r->state += splitmix64(&sm);
    pcg32(r);
    r->has_spare = 0;
}

static double uniform01(rng_t* r) { return (double)(pcg32(r) >> 8) * (1.0 / 16777216.0); }

static double normal01(rng_t* r)
{
    if (r->has_spare) {
        r->has_spare = 0;
        return r->spare;
    }
    const double u1 = 1.0 - uniform01(r);
    const double u2 = uniform01(r);
    const double rad = sqrt(-2.0 * log(u1));
    r->spare = rad * sin(2.0 * SYN_PI * u2);
    r->has_spare = 1;
    return rad * cos(2.0 * SYN_PI * u2);
}

/* exported: uniform and normal draws for host-side parameter sampling */
double lsseg_synth_uniform(uint64_t seed, uint64_t stream, uint64_t index)
{
    rng_t r;
    rng_seed(&r, seed, stream);
    for (uint64_t i = 0; i < index; ++i) pcg32(&r);
    return uniform01(&r);
}

static float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

/*
 * Background / foreground ellipsoids with additive Gaussian noise.
 * ell: n_ell records of (cx, cy, cz, ax, ay, az) in voxels.
 * out_u8 (if non-null): round(value) clamped to [0, 255].
 * out_f32 (if non-null): clamp(value, lo, hi) * out_scale (no rounding).
 * gt (if non-null): 1 inside any ellipsoid.
 */
void lsseg_synth_ellipsoids(int32_t nx, int32_t ny, int32_t nz, uint64_t seed,
                            float bg, float fg, float sigma, float lo, float hi,
                            int32_t n_ell, const float* ell,
                            uint8_t* out_u8, float* out_f32, float out_scale, uint8_t* gt)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t z = 0; z < nz; ++z) {
        rng_t r;
        rng_seed(&r, seed, (uint64_t)z);
        for (int32_t y = 0; y < ny; ++y)
            for (int32_t x = 0; x < nx; ++x) {
                int inside = 0;
                for (int32_t k = 0; k < n_ell && !inside; ++k) {
                    const float* e = ell + 6 * k;
                    const double dx = (x - e[0]) / e[3], dy = (y - e[1]) / e[4], dz = (z - e[2]) / e[5];
                    if (dx * dx + dy * dy + dz * dz <= 1.0) inside = 1;
                }
                const double v = (inside ? fg : bg) + sigma * normal01(&r);
                const uint64_t i = ((uint64_t)z * ny + y) * nx + x;
                if (out_u8) out_u8[i] = (uint8_t)clampf(floorf((float)v + 0.5f), lo, hi);
                if (out_f32) out_f32[i] = clampf((float)v, lo, hi) * out_scale;
                if (gt) gt[i] = (uint8_t)inside;
            }
    }
}

/* phi0 = clamp(|x - c| - radius, -clampv, clampv): signed distance to a sphere */
void lsseg_synth_sphere_sdf(int32_t nx, int32_t ny, int32_t nz, float cx, float cy, float cz,
                            float radius, float clampv, float* phi)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t z = 0; z < nz; ++z)
        for (int32_t y = 0; y < ny; ++y)
            for (int32_t x = 0; x < nx; ++x) {
                const double dx = x - cx, dy = y - cy, dz = z - cz;
                const double d = sqrt(dx * dx + dy * dy + dz * dz) - radius;
                phi[((uint64_t)z * ny + y) * nx + x] = clampf((float)d, -clampv, clampv);
            }
}

/* separable Gaussian blur along one axis (stride, length), radius ceil(3 sigma),
   clamp-to-edge */
static void blur_axis(float* dst, const float* src, int32_t nx, int32_t ny, int32_t nz, int axis,
                      const float* kern, int32_t rad)
{
    const int32_t dims[3] = {nx, ny, nz};
    const int64_t strides[3] = {1, nx, (int64_t)nx * ny};
    const int32_t L = dims[axis];
    const int64_t st = strides[axis];
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t z = 0; z < nz; ++z)
        for (int32_t y = 0; y < ny; ++y)
            for (int32_t x = 0; x < nx; ++x) {
                const int32_t c[3] = {x, y, z};
                const int64_t base = ((int64_t)z * ny + y) * nx + x - (int64_t)c[axis] * st;
                double acc = 0.0;
                for (int32_t k = -rad; k <= rad; ++k) {
                    int32_t q = c[axis] + k;
                    q = q < 0 ? 0 : (q >= L ? L - 1 : q);
                    acc += kern[k + rad] * src[base + (int64_t)q * st];
                }
                dst[((int64_t)z * ny + y) * nx + x] = (float)acc;
            }
}

static int cmp_float(const void* a, const void* b)
{
    const float x = *(const float*)a, y = *(const float*)b;
    return (x > y) - (x < y);
}

/*
 * Four nested smooth classes from Gaussian-smoothed white noise thresholded at
 * the quantiles q[0] < q[1] < q[2] (0 = background, 3 = innermost), class means
 * means[4], additive noise sigma, multiplicative bias 1 + bias_amp * mean of three
 * random 3D cosines (frequencies up to 2 cycles per volume).
 * Writes out (float), labels (uint8), and picks a seed centre: the first point of
 * a seeded random scan whose ball of radius seed_clear lies inside class
 * seed_class.  Returns 0 on success, -1 if no such point exists.
 */
int32_t lsseg_synth_phantom(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, float smooth_sigma,
                            const float* q, const float* means, float sigma, float bias_amp,
                            int32_t seed_class, float seed_clear,
                            float* out, uint8_t* labels, float* seed_xyz)
{
    const int64_t N = (int64_t)nx * ny * nz;
    float* a = (float*)malloc((size_t)N * sizeof(float));
    float* b = (float*)malloc((size_t)N * sizeof(float));
    if (!a || !b) {
        free(a);
        free(b);
        return -1;
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t z = 0; z < nz; ++z) {
        rng_t r;
        rng_seed(&r, seed, (uint64_t)z);
        for (int64_t i = (int64_t)z * nx * ny; i < (int64_t)(z + 1) * nx * ny; ++i)
            a[i] = (float)normal01(&r);
    }
    const int32_t rad = (int32_t)ceil(3.0 * smooth_sigma);
    float* kern = (float*)malloc((size_t)(2 * rad + 1) * sizeof(float));
    double ks = 0.0;
    for (int32_t k = -rad; k <= rad; ++k) {
        kern[k + rad] = (float)exp(-0.5 * (double)k * k / ((double)smooth_sigma * smooth_sigma));
        ks += kern[k + rad];
    }
    for (int32_t k = 0; k <= 2 * rad; ++k) kern[k] = (float)(kern[k] / ks);
    blur_axis(b, a, nx, ny, nz, 0, kern, rad);
    blur_axis(a, b, nx, ny, nz, 1, kern, rad);
    blur_axis(b, a, nx, ny, nz, 2, kern, rad);
    free(kern);
    /* quantile thresholds of the smoothed field */
    memcpy(a, b, (size_t)N * sizeof(float));
    qsort(a, (size_t)N, sizeof(float), cmp_float);
    float thr[3];
    for (int k = 0; k < 3; ++k) {
        int64_t idx = (int64_t)(q[k] * (double)(N - 1));
        thr[k] = a[idx];
    }
    /* bias field parameters */
    rng_t pr;
    rng_seed(&pr, seed, 0xB1A5ull);
    double fx[3], fy[3], fz[3], ph[3];
    for (int k = 0; k < 3; ++k) {
        fx[k] = 2.0 * uniform01(&pr);
        fy[k] = 2.0 * uniform01(&pr);
        fz[k] = 2.0 * uniform01(&pr);
        ph[k] = 2.0 * SYN_PI * uniform01(&pr);
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t z = 0; z < nz; ++z) {
        rng_t r;
        rng_seed(&r, seed ^ 0x5A5A5A5Aull, (uint64_t)z);
        for (int32_t y = 0; y < ny; ++y)
            for (int32_t x = 0; x < nx; ++x) {
                const int64_t i = ((int64_t)z * ny + y) * nx + x;
                const float f = b[i];
                const int cls = (f > thr[0]) + (f > thr[1]) + (f > thr[2]);
                double bias = 0.0;
                for (int k = 0; k < 3; ++k)
                    bias += cos(2.0 * SYN_PI * (fx[k] * x / nx + fy[k] * y / ny + fz[k] * z / nz) + ph[k]);
                bias = 1.0 + bias_amp * (bias / 3.0);
                labels[i] = (uint8_t)cls;
                out[i] = (float)((means[cls] + sigma * normal01(&r)) * bias);
            }
    }
    free(a);
    free(b);
    /* seed search */
    rng_t sr;
    rng_seed(&sr, seed, 0x5EEDull);
    const int32_t cr = (int32_t)ceil(seed_clear);
    for (int tries = 0; tries < 2000000; ++tries) {
        const int32_t cx = cr + (int32_t)(uniform01(&sr) * (nx - 2 * cr));
        const int32_t cy = cr + (int32_t)(uniform01(&sr) * (ny - 2 * cr));
        const int32_t cz = cr + (int32_t)(uniform01(&sr) * (nz - 2 * cr));
        if (labels[((int64_t)cz * ny + cy) * nx + cx] != seed_class) continue;
        int ok = 1;
        for (int32_t dz = -cr; dz <= cr && ok; ++dz)
            for (int32_t dy = -cr; dy <= cr && ok; ++dy)
                for (int32_t dx = -cr; dx <= cr && ok; ++dx) {
                    if (dx * dx + dy * dy + dz * dz > seed_clear * seed_clear) continue;
                    if (labels[((int64_t)(cz + dz) * ny + (cy + dy)) * nx + (cx + dx)] != seed_class) ok = 0;
                }
        if (ok) {
            seed_xyz[0] = (float)cx;
            seed_xyz[1] = (float)cy;
            seed_xyz[2] = (float)cz;
            return 0;
        }
    }
    return -1;
}

static void rand_unit(rng_t* r, double v[3])
{
    double n2;
    do {
        v[0] = 2.0 * uniform01(r) - 1.0;
        v[1] = 2.0 * uniform01(r) - 1.0;
        v[2] = 2.0 * uniform01(r) - 1.0;
        n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    } while (n2 > 1.0 || n2 < 1e-6);
    const double inv = 1.0 / sqrt(n2);
    v[0] *= inv; v[1] *= inv; v[2] *= inv;
}

/* distance^2 from p to segment [a, b]; t_out = projection parameter in [0, 1] */
static double seg_dist2(const double p[3], const double a[3], const double b[3], double* t_out)
{
    double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    const double l2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
    double t = l2 > 0 ? (ap[0] * ab[0] + ap[1] * ab[1] + ap[2] * ab[2]) / l2 : 0.0;
    t = t < 0 ? 0 : (t > 1 ? 1 : t);
    const double d[3] = {ap[0] - t * ab[0], ap[1] - t * ab[1], ap[2] - t * ab[2]};
    *t_out = t;
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

/*
 * Vessel tree: n_tubes tubes along seeded random walks (step `step` voxels,
 * steps drawn uniformly in [min_steps, max_steps]); tube radius drawn uniformly
 * in [r_lo, r_hi] (the root tube in [max(r_lo, root_r_lo), r_hi]) and tapering
 * linearly to `taper` of it along the tube.  Tube 0 starts at the root and heads
 * +z; every later tube branches from a random point of an earlier tube.
 * Intensities: vessel N(fg, fg_sigma), background N(bg, bg_sigma), rounded and
 * clamped to [0, 65535].  Writes vol (uint16), optional mask, and the root.
 */
void lsseg_synth_vessels(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, int32_t n_tubes,
                         float step, int32_t min_steps, int32_t max_steps, float r_lo, float r_hi,
                         float root_r_lo, float taper, float bg, float bg_sigma, float fg, float fg_sigma,
                         uint16_t* vol, uint8_t* mask_out, float* root_xyz)
{
    const int64_t N = (int64_t)nx * ny * nz;
    uint8_t* mask = mask_out ? mask_out : (uint8_t*)calloc((size_t)N, 1);
    if (mask_out) memset(mask, 0, (size_t)N);
    rng_t r;
    rng_seed(&r, seed, 0x7BEEull);
    const int32_t maxpts = max_steps + 1;
    double* pts = (double*)malloc((size_t)n_tubes * maxpts * 3 * sizeof(double));
    int32_t* npts = (int32_t*)calloc((size_t)n_tubes, sizeof(int32_t));
    double* r0 = (double*)calloc((size_t)n_tubes, sizeof(double));
    const double margin = 12.0;
    const double root[3] = {nx * 0.5, ny * 0.5, margin + 8.0};
    for (int32_t k = 0; k < n_tubes; ++k) {
        double p[3], d[3];
        if (k == 0) {
            p[0] = root[0]; p[1] = root[1]; p[2] = root[2];
            d[0] = 0; d[1] = 0; d[2] = 1;
            const double lo = r_lo > root_r_lo ? r_lo : root_r_lo;
            r0[k] = lo + (r_hi - lo) * uniform01(&r);
        } else {
            const int32_t parent = (int32_t)(uniform01(&r) * k);
            const int32_t at = (int32_t)(uniform01(&r) * npts[parent]);
            const double* q = pts + ((int64_t)parent * maxpts + at) * 3;
            p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
            rand_unit(&r, d);
            r0[k] = r_lo + (r_hi - r_lo) * uniform01(&r);
        }
        const int32_t steps = min_steps + (int32_t)(uniform01(&r) * (max_steps - min_steps + 1));
        double* P = pts + (int64_t)k * maxpts * 3;
        P[0] = p[0]; P[1] = p[1]; P[2] = p[2];
        int32_t m = 1;
        for (int32_t s = 0; s < steps; ++s) {
            double w[3];
            rand_unit(&r, w);
            for (int i = 0; i < 3; ++i) d[i] = d[i] + 0.35 * w[i];
            double nd = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            for (int i = 0; i < 3; ++i) d[i] /= nd;
            const double lim[3] = {nx, ny, nz};
            for (int i = 0; i < 3; ++i) {
                double np = p[i] + step * d[i];
                if (np < margin || np > lim[i] - 1 - margin) {
                    d[i] = -d[i];
                    np = p[i] + step * d[i];
                }
                p[i] = np;
            }
            P[3 * m + 0] = p[0]; P[3 * m + 1] = p[1]; P[3 * m + 2] = p[2];
            ++m;
        }
        npts[k] = m;
        /* rasterize capsules */
        for (int32_t s = 0; s + 1 < m; ++s) {
            const double* a = P + 3 * s;
            const double* b = P + 3 * (s + 1);
            const double ra = r0[k] * (1.0 - (1.0 - taper) * (double)s / (m - 1));
            const double rb = r0[k] * (1.0 - (1.0 - taper) * (double)(s + 1) / (m - 1));
            const double rmax = ra > rb ? ra : rb;
            int32_t lo[3], hi[3];
            for (int i = 0; i < 3; ++i) {
                const double mn = (a[i] < b[i] ? a[i] : b[i]) - rmax - 1;
                const double mx = (a[i] > b[i] ? a[i] : b[i]) + rmax + 1;
                const int32_t lim = i == 0 ? nx : (i == 1 ? ny : nz);
                lo[i] = mn < 0 ? 0 : (int32_t)mn;
                hi[i] = mx > lim - 1 ? lim - 1 : (int32_t)mx;
            }
            for (int32_t z = lo[2]; z <= hi[2]; ++z)
                for (int32_t y = lo[1]; y <= hi[1]; ++y)
                    for (int32_t x = lo[0]; x <= hi[0]; ++x) {
                        const double pp[3] = {x, y, z};
                        double t;
                        const double d2 = seg_dist2(pp, a, b, &t);
                        const double rr = ra + (rb - ra) * t;
                        if (d2 <= rr * rr) mask[((int64_t)z * ny + y) * nx + x] = 1;
                    }
        }
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t z = 0; z < nz; ++z) {
        rng_t q;
        rng_seed(&q, seed, (uint64_t)z);
        for (int64_t i = (int64_t)z * nx * ny; i < (int64_t)(z + 1) * nx * ny; ++i) {
            const double v = mask[i] ? fg + fg_sigma * normal01(&q) : bg + bg_sigma * normal01(&q);
            vol[i] = (uint16_t)clampf(floorf((float)v + 0.5f), 0.0f, 65535.0f);
        }
    }
    root_xyz[0] = (float)root[0];
    root_xyz[1] = (float)root[1];
    root_xyz[2] = (float)root[2];
    free(pts);
    free(npts);
    free(r0);
    if (!mask_out) free(mask);
}
