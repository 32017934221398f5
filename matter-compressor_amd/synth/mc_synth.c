/*
 * mc_synth.c -- deterministic synthetic scroll-like volumes for the codec
 * workloads (layered sheets + Gaussian noise + contiguous air regions).
 *
 *   v = clamp(round(128 + 70 sin(2 pi z/23 + 0.4 sin(2 pi x/97)
 *                               + 0.3 sin(2 pi y/61)) + 8 n), 1, 255)
 *
 * in global voxel coordinates, n ~ N(0,1) by Box-Muller over a counter-based
 * SplitMix64 stream keyed by (seed, voxel index in the chunk).  Air: a 32^3
 * grid of U(0,1) per chunk (its own SplitMix64 stream), smoothed by a
 * [1,2,1]^3 kernel with clamped edges, trilinearly upsampled x8 at cell
 * centres; voxels whose field value is below the air-fraction quantile of
 * the chunk (the k-th smallest value, k = floor(frac * 256^3)) become 0.
 *
 * Built with -ffp-contract=off and no fast-math, so the bytes depend only on
 * IEEE double arithmetic and libm's sin/cos/log/sqrt.  Callers pin the output
 * with a per-chunk xxh64.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define CH 256
#define NV (CH * CH * CH)

static uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static double unit_open(uint64_t r) { return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

/* k-th smallest (0-based) of n floats; reorders buf (Wirth's selection). */
static float kth_smallest(float *a, long n, long k) {
    long l = 0, m = n - 1;
    while (l < m) {
        float x = a[k];
        long i = l, j = m;
        do {
            while (a[i] < x) i++;
            while (x < a[j]) j--;
            if (i <= j) { float t = a[i]; a[i] = a[j]; a[j] = t; i++; j--; }
        } while (i <= j);
        if (j < k) l = i;
        if (k < i) m = j;
    }
    return a[k];
}

/* Air field of one chunk as float values (smoothed, upsampled). */
static void air_field(uint64_t seed, float *field) {
    static const double w[3] = {1.0, 2.0, 1.0};
    double g[32][32][32], s[32][32][32];
    uint64_t key = splitmix(seed ^ 0xA1A1A1A1A1A1A1A1ull);
    for (int i = 0; i < 32 * 32 * 32; ++i)
        (&g[0][0][0])[i] = unit_open(splitmix(key + (uint64_t)i));
    for (int z = 0; z < 32; ++z)
        for (int y = 0; y < 32; ++y)
            for (int x = 0; x < 32; ++x) {
                double acc = 0.0, wt = 0.0;
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            int zz = z + dz, yy = y + dy, xx = x + dx;
                            zz = zz < 0 ? 0 : (zz > 31 ? 31 : zz);
                            yy = yy < 0 ? 0 : (yy > 31 ? 31 : yy);
                            xx = xx < 0 ? 0 : (xx > 31 ? 31 : xx);
                            double ww = w[dz + 1] * w[dy + 1] * w[dx + 1];
                            acc += ww * g[zz][yy][xx];
                            wt += ww;
                        }
                s[z][y][x] = acc / wt;
            }
    /* per-axis interpolation weights: voxel v sits at grid coordinate (v+0.5)/8-0.5 */
    int i0[CH], i1[CH];
    double f[CH];
    for (int v = 0; v < CH; ++v) {
        double c = (v + 0.5) / 8.0 - 0.5;
        if (c < 0) c = 0;
        if (c > 31) c = 31;
        int a = (int)floor(c);
        if (a > 30) a = 30;
        i0[v] = a; i1[v] = a + 1; f[v] = c - a;
    }
    for (int z = 0; z < CH; ++z)
        for (int y = 0; y < CH; ++y) {
            float *row = field + ((size_t)z * CH + y) * CH;
            double fz = f[z], fy = f[y];
            for (int x = 0; x < CH; ++x) {
                double fx = f[x];
                double c00 = s[i0[z]][i0[y]][i0[x]] * (1 - fx) + s[i0[z]][i0[y]][i1[x]] * fx;
                double c01 = s[i0[z]][i1[y]][i0[x]] * (1 - fx) + s[i0[z]][i1[y]][i1[x]] * fx;
                double c10 = s[i1[z]][i0[y]][i0[x]] * (1 - fx) + s[i1[z]][i0[y]][i1[x]] * fx;
                double c11 = s[i1[z]][i1[y]][i0[x]] * (1 - fx) + s[i1[z]][i1[y]][i1[x]] * fx;
                double c0 = c00 * (1 - fy) + c01 * fy, c1 = c10 * (1 - fy) + c11 * fy;
                row[x] = (float)(c0 * (1 - fz) + c1 * fz);
            }
        }
}

/* One 256^3 chunk at chunk coordinates (cz,cy,cx). */
void mcs_chunk(uint64_t seed, int cz, int cy, int cx, double air_frac, uint8_t *out) {
    const double two_pi = 6.283185307179586476925;
    double px[CH], py[CH], pz[CH];
    for (int i = 0; i < CH; ++i) {
        px[i] = 0.4 * sin(two_pi * (double)(cx * CH + i) / 97.0);
        py[i] = 0.3 * sin(two_pi * (double)(cy * CH + i) / 61.0);
        pz[i] = two_pi * (double)(cz * CH + i) / 23.0;
    }
    uint64_t key = splitmix(seed);
    for (size_t z = 0; z < CH; ++z)
        for (size_t y = 0; y < CH; ++y)
            for (size_t x = 0; x < CH; ++x) {
                size_t i = (z * CH + y) * CH + x;
                uint64_t r1 = splitmix(key + 2 * (uint64_t)i), r2 = splitmix(key + 2 * (uint64_t)i + 1);
                double n = sqrt(-2.0 * log(unit_open(r1))) * cos(two_pi * unit_open(r2));
                double v = 128.0 + 70.0 * sin(pz[z] + px[x] + py[y]) + 8.0 * n;
                v = round(v);
                if (v < 1) v = 1;
                if (v > 255) v = 255;
                out[i] = (uint8_t)v;
            }
    if (air_frac > 0.0) {
        float *field = malloc((size_t)NV * sizeof(float)), *tmp = malloc((size_t)NV * sizeof(float));
        air_field(seed, field);
        memcpy(tmp, field, (size_t)NV * sizeof(float));
        long k = (long)floor(air_frac * (double)NV);
        if (k >= (long)NV) memset(out, 0, NV);
        else if (k > 0) {
            float thr = kth_smallest(tmp, NV, k);
            for (size_t i = 0; i < (size_t)NV; ++i) if (field[i] < thr) out[i] = 0;
        }
        free(field); free(tmp);
    }
}

typedef struct {
    int n; const uint64_t *seeds; const int *cz, *cy, *cx; const double *air; uint8_t *out;
    atomic_int next;
} job_t;
static void *worker(void *p) {
    job_t *j = p;
    for (;;) {
        int i = atomic_fetch_add(&j->next, 1);
        if (i >= j->n) break;
        mcs_chunk(j->seeds[i], j->cz ? j->cz[i] : 0, j->cy ? j->cy[i] : 0, j->cx ? j->cx[i] : 0,
                  j->air[i], j->out + (size_t)i * NV);
    }
    return NULL;
}
/* n chunks in parallel (chunk coordinate arrays may be NULL = origin). */
void mcs_chunks(int n, const uint64_t *seeds, const int *cz, const int *cy, const int *cx,
                const double *air, uint8_t *out, int nthreads) {
    job_t j = { .n = n, .seeds = seeds, .cz = cz, .cy = cy, .cx = cx, .air = air, .out = out };
    atomic_init(&j.next, 0);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > n) nthreads = n;
    pthread_t th[64];
    if (nthreads > 64) nthreads = 64;
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, worker, &j);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}
