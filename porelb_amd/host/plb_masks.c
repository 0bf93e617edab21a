/*
 * plb_masks.c -- seeded synthetic voxel masks (u8 [nz][ny][nx], 1 = pore,
 * 0 = solid, x fastest: the flag.bin layout preprocess.c:61-69 reads).
 *
 * The reference ships no mask generator; these produce the benchmark and
 * test domains.  Every value is a pure function of (dims, parameters, seed):
 * random numbers come from a counter-based splitmix64 hash, so a voxel's
 * noise does not depend on evaluation order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "plb.h"

static uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
/* uniform in [0,1) from stream `seed`, counter `k` */
static double u01(uint64_t seed, uint64_t k)
{
    return (double)(mix64(seed ^ mix64(k)) >> 11) * (1.0 / 9007199254740992.0);
}

/* ---------------------------------------------------------------- noise */
/* Periodic 1-D convolution with weights w[0..2R] along one axis; the loops
 * walk rows and planes so every access is contiguous along x. */
static void blur_axis(const float *in, float *out, int nx, int ny, int nz, int axis,
                      const double *w, int R)
{
    const int64_t plane = (int64_t)nx * ny;
    if (axis == 0) {
        double *acc = (double *)malloc(sizeof(double) * (size_t)nx);
        for (int64_t r = 0; r < (int64_t)ny * nz; r++) {
            const float *src = in + r * nx;
            for (int x = 0; x < nx; x++) {
                double s = 0.0;
                for (int k = -R; k <= R; k++) s += w[k + R] * src[((x + k) % nx + nx) % nx];
                acc[x] = s;
            }
            for (int x = 0; x < nx; x++) out[r * nx + x] = (float)acc[x];
        }
        free(acc);
        return;
    }
    const int n = axis == 1 ? ny : nz;
    const int64_t len = axis == 1 ? nx : plane;      /* contiguous block per index */
    const int64_t outer = axis == 1 ? nz : 1;        /* independent groups */
    const int64_t group = axis == 1 ? plane : 0;
    double *acc = (double *)malloc(sizeof(double) * (size_t)len);
    for (int64_t o = 0; o < outer; o++) {
        const float *gin = in + o * group;
        float *gout = out + o * group;
        for (int i = 0; i < n; i++) {
            for (int64_t t = 0; t < len; t++) acc[t] = 0.0;
            for (int k = -R; k <= R; k++) {
                const float *src = gin + (int64_t)(((i + k) % n + n) % n) * len;
                const double wk = w[k + R];
                for (int64_t t = 0; t < len; t++) acc[t] += wk * src[t];
            }
            for (int64_t t = 0; t < len; t++) gout[(int64_t)i * len + t] = (float)acc[t];
        }
    }
    free(acc);
}

static void swap_f(float *a, float *b)
{
    float t = *a;
    *a = *b;
    *b = t;
}

/* k-th smallest (0-based) of v[0..n), reorders v. */
static float select_kth(float *v, int64_t n, int64_t k)
{
    int64_t lo = 0, hi = n - 1;
    while (hi > lo) {
        int64_t mid = lo + (hi - lo) / 2;
        if (v[mid] < v[lo]) swap_f(&v[mid], &v[lo]);
        if (v[hi] < v[lo]) swap_f(&v[hi], &v[lo]);
        if (v[hi] < v[mid]) swap_f(&v[hi], &v[mid]);
        float piv = v[mid];
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (v[i] < piv) i++;
            while (v[j] > piv) j--;
            if (i <= j) {
                swap_f(&v[i], &v[j]);
                i++;
                j--;
            }
        }
        if (k <= j)
            hi = j;
        else if (k >= i)
            lo = i;
        else
            return v[k];
    }
    return v[k];
}

int plb_mask_noise(uint8_t *flag, int nx, int ny, int nz, double sigma, double porosity,
                   uint64_t seed)
{
    if (!flag || nx <= 0 || ny <= 0 || nz <= 0 || !(sigma >= 0.0) || !(porosity >= 0.0) ||
        porosity > 1.0)
        return PLB_ERR_INVALID;
    const int64_t V = (int64_t)nx * ny * nz;
    float *a = (float *)malloc(sizeof(float) * (size_t)V);
    float *b = (float *)malloc(sizeof(float) * (size_t)V);
    if (!a || !b) {
        free(a);
        free(b);
        return PLB_ERR_NOMEM;
    }
    const double two_pi = 6.283185307179586;
    for (int64_t k = 0; k < V; k++) {
        /* Box-Muller, cosine branch: one N(0,1) value per voxel. */
        double u1 = 1.0 - u01(seed, 2 * (uint64_t)k);
        double u2 = u01(seed, 2 * (uint64_t)k + 1);
        a[k] = (float)(sqrt(-2.0 * log(u1)) * cos(two_pi * u2));
    }
    if (sigma > 0.0) {
        int R = (int)ceil(4.0 * sigma);
        double *w = (double *)malloc(sizeof(double) * (size_t)(2 * R + 1));
        double s = 0.0;
        for (int k = -R; k <= R; k++) s += (w[k + R] = exp(-(double)k * k / (2.0 * sigma * sigma)));
        for (int k = 0; k < 2 * R + 1; k++) w[k] /= s;
        blur_axis(a, b, nx, ny, nz, 0, w, R);
        blur_axis(b, a, nx, ny, nz, 1, w, R);
        blur_axis(a, b, nx, ny, nz, 2, w, R);
        memcpy(a, b, sizeof(float) * (size_t)V);
        free(w);
    }
    /* Exact porosity: the lowest round(porosity*V) values are pore, ties
     * broken by voxel index. */
    int64_t npore = llround(porosity * (double)V);
    memset(flag, 0, (size_t)V);
    if (npore > 0) {
        memcpy(b, a, sizeof(float) * (size_t)V);
        float t = select_kth(b, V, npore - 1);
        int64_t below = 0;
        for (int64_t k = 0; k < V; k++) below += (a[k] < t);
        int64_t ties = npore - below;
        for (int64_t k = 0; k < V; k++) {
            if (a[k] < t)
                flag[k] = 1;
            else if (a[k] == t && ties > 0) {
                flag[k] = 1;
                ties--;
            }
        }
    }
    free(a);
    free(b);
    return PLB_OK;
}

/* ---------------------------------------------------------------- spheres */
int plb_mask_spheres(uint8_t *flag, int nx, int ny, int nz, double rmin, double rmax,
                     double porosity, uint64_t seed)
{
    if (!flag || nx <= 0 || ny <= 0 || nz <= 0 || !(rmin > 0.0) || rmax < rmin ||
        !(porosity >= 0.0) || porosity > 1.0)
        return PLB_ERR_INVALID;
    const int64_t V = (int64_t)nx * ny * nz;
    memset(flag, 1, (size_t)V);
    int64_t pores = V;
    const double target = porosity * (double)V;
    for (uint64_t s = 0; (double)pores > target; s++) {
        const double cx = u01(seed, 4 * s) * nx, cy = u01(seed, 4 * s + 1) * ny,
                     cz = u01(seed, 4 * s + 2) * nz;
        const double r = rmin + (rmax - rmin) * u01(seed, 4 * s + 3);
        const double r2 = r * r;
        for (int z = (int)ceil(cz - r); z <= (int)floor(cz + r); z++) {
            const double dz = z - cz, rz = r2 - dz * dz;
            if (rz < 0) continue;
            const double sy = sqrt(rz);
            const int zz = ((z % nz) + nz) % nz;
            for (int y = (int)ceil(cy - sy); y <= (int)floor(cy + sy); y++) {
                const double dy = y - cy, ry = rz - dy * dy;
                if (ry < 0) continue;
                const double sx = sqrt(ry);
                const int yy = ((y % ny) + ny) % ny;
                uint8_t *row = flag + ((int64_t)zz * ny + yy) * nx;
                for (int x = (int)ceil(cx - sx); x <= (int)floor(cx + sx); x++) {
                    const double dx = x - cx;
                    if (dx * dx + dy * dy + dz * dz > r2) continue;
                    const int xx = ((x % nx) + nx) % nx;
                    if (row[xx]) {
                        row[xx] = 0;
                        pores--;
                    }
                }
            }
        }
    }
    return PLB_OK;
}

/* ---------------------------------------------------------------- BCC array */
int plb_mask_bcc(uint8_t *flag, int n, double chi)
{
    if (!flag || n <= 1 || !(chi >= 0.0) || chi > 1.0) return PLB_ERR_INVALID;
    const double a = sqrt(3.0) * n * chi / 4.0, a2 = a * a;
    const double c = n / 2.0;
    for (int z = 0; z < n; z++)
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) {
                /* periodic distance to the corner sphere and the centre sphere */
                double d0 = 0.0, d1 = 0.0;
                const int q[3] = {x, y, z};
                for (int k = 0; k < 3; k++) {
                    double t0 = fabs((double)q[k]);
                    if (t0 > n / 2.0) t0 = n - t0;
                    double t1 = fabs((double)q[k] - c);
                    if (t1 > n / 2.0) t1 = n - t1;
                    d0 += t0 * t0;
                    d1 += t1 * t1;
                }
                flag[((int64_t)z * n + y) * n + x] = (d0 <= a2 || d1 <= a2) ? 0 : 1;
            }
    return PLB_OK;
}
