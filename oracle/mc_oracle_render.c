/*
 * mc_oracle_render.c -- CPU restatement of the matter-compressor LOD
 * decimation and quad-surface render subset (TEST INFRASTRUCTURE ONLY; see
 * mc_oracle.h).
 *
 * Restated: decimate (src/mc_archive.c:61-72), mc_quad_gen
 * (src/mc_render.c:1981-2052) as driven row by row by quad_rowgen
 * (:2097-2102), the LOD remap of lod_rowgen / render_lod (:2255-2288),
 * mc_render_pick_lod and mc_quad_spacing (:2222-2250), make_cfg's step count
 * (:827-840), render_pixel for composites NONE, MIN, MEAN, MAX, ALPHA (with
 * zero alpha_min / alpha_opacity, the defaults) and STDDEV (:1150-1281; the
 * trilinear filter walks the ray in groups of four -- the mean takes the
 * group sum ((v0+v1)+v2)+v3 -- then one sample at a time), the trilinear and
 * nearest samplers (:176-217, :86-89, :58-67; every path of the reference's
 * sampler -- dense, in-block, straddling, edge, and the 4-wide groups --
 * computes the same corner values and the same mul/add sequence) and to_u8
 * (:793-795).
 *
 * A level is read either from a dense raster or from 16^3 blocks through a
 * per-block table (absent blocks read 0), as the reference's dense and
 * block sources do (:614-639).
 */
#include "mc_oracle.h"
#include <math.h>
#include <string.h>

/* ---- LOD decimation (src/mc_archive.c:61-72) ----------------------------- */
void orc_decimate(const uint8_t *src, int dz, int dy, int dx, uint8_t *out) {
    const int hz = (dz + 1) / 2, hy = (dy + 1) / 2, hx = (dx + 1) / 2;
    for (int z = 0; z < hz; ++z)
        for (int y = 0; y < hy; ++y)
            for (int x = 0; x < hx; ++x) {
                int s = 0, c = 0;
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b)
                        for (int e = 0; e < 2; ++e) {
                            const int sz = 2 * z + a, sy = 2 * y + b, sx = 2 * x + e;
                            if (sz >= dz || sy >= dy || sx >= dx) continue;
                            const uint8_t v = src[((size_t)sz * dy + sy) * dx + sx];
                            if (v) { s += v; c++; }
                        }
                out[((size_t)z * hy + y) * hx + x] = c ? (uint8_t)((s + c / 2) / c) : 0;
            }
}

/* ---- sampling ------------------------------------------------------------ */
static float voxel(const orc_level *L, int z, int y, int x) {
    if ((unsigned)z >= (unsigned)L->nz || (unsigned)y >= (unsigned)L->ny || (unsigned)x >= (unsigned)L->nx)
        return 0.0f;
    if (L->dense) return (float)L->dense[((size_t)z * L->ny + y) * L->nx + x];
    const int gby = (L->ny + 15) >> 4, gbx = (L->nx + 15) >> 4;
    const int64_t o = L->blk_off[((size_t)(z >> 4) * gby + (y >> 4)) * gbx + (x >> 4)];
    if (o < 0) return 0.0f;
    return (float)L->blk_data[(size_t)o + (((z & 15) << 8) | ((y & 15) << 4) | (x & 15))];
}

static int isnan_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return (u & 0x7FFFFFFFu) > 0x7F800000u;
}

static float trilinear(const orc_level *L, float z, float y, float x) {
    const float zf = floorf(z), yf = floorf(y), xf = floorf(x);
    const int z0 = (int)zf, y0 = (int)yf, x0 = (int)xf;
    const float dz = z - zf, dy = y - yf, dx = x - xf;
    const float c000 = voxel(L, z0, y0, x0), c001 = voxel(L, z0, y0, x0 + 1);
    const float c010 = voxel(L, z0, y0 + 1, x0), c011 = voxel(L, z0, y0 + 1, x0 + 1);
    const float c100 = voxel(L, z0 + 1, y0, x0), c101 = voxel(L, z0 + 1, y0, x0 + 1);
    const float c110 = voxel(L, z0 + 1, y0 + 1, x0), c111 = voxel(L, z0 + 1, y0 + 1, x0 + 1);
    const float c00 = c000 + (c001 - c000) * dx;
    const float c01 = c010 + (c011 - c010) * dx;
    const float c10 = c100 + (c101 - c100) * dx;
    const float c11 = c110 + (c111 - c110) * dx;
    const float c0 = c00 + (c01 - c00) * dy;
    const float c1 = c10 + (c11 - c10) * dy;
    return c0 + (c1 - c0) * dz;
}

static float nearest(const orc_level *L, float z, float y, float x) {
    return voxel(L, (int)floorf(z + 0.5f), (int)floorf(y + 0.5f), (int)floorf(x + 0.5f));
}

static float sample(const orc_level *L, float z, float y, float x, int filter) {
    if (isnan_bits(z) || isnan_bits(y) || isnan_bits(x)) return 0.0f;
    return filter == ORC_FILTER_NEAREST ? nearest(L, z, y, x) : trilinear(L, z, y, x);
}

static uint8_t to_u8(float v) { return (uint8_t)(v < 0.0f ? 0 : v > 255.0f ? 255 : (int)(v + 0.5f)); }

static int pt_valid(const float *p) {
    if (p[0] != p[0] || p[1] != p[1] || p[2] != p[2]) return 0;
    return p[0] >= 0.0f && p[1] >= 0.0f && p[2] >= 0.0f;
}

/* ---- quad surface (src/mc_render.c:1981-2052) --------------------------- */
static int qvalid(const float *p) {
    return !(p[0] == -1.0f && p[1] == -1.0f && p[2] == -1.0f) && p[0] == p[0] && p[1] == p[1] && p[2] == p[2];
}

static void v3_norm(float *v) {
    const float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 1e-12f) { v[0] /= l; v[1] /= l; v[2] /= l; }
}

/* Point and unit normal of pixel (row, j): the reference generates row `row`
 * with mc_quad_gen(q, x0, y0 + row*step, step, w, 1, ...).  The per-cell
 * values it caches along a row depend only on the cell and the row, so they
 * are recomputed per pixel here. */
void orc_quad_point(const float *grid, int gw, int gh, float x0, float y0, float step, int row, int j, float *P,
                    float *N) {
    P[0] = P[1] = P[2] = -1.0f;
    N[0] = N[1] = N[2] = 0.0f;
    const float yrow = y0 + (float)row * step;
    const float gy = yrow + (float)0 * step;
    if (!(gy >= 0.0f) || gy > (float)(gh - 1)) return;
    int y0i = (int)gy;
    if (y0i > gh - 2) y0i = gh - 2;
    if (y0i < 0) y0i = 0;
    const float fy = gy - (float)y0i;
    const float *r0 = grid + (size_t)y0i * gw * 3;
    const float *r1 = grid + (size_t)(y0i + (gh > 1)) * gw * 3;
    const float gx = x0 + (float)j * step;
    if (!(gx >= 0.0f) || gx > (float)(gw - 1)) return;
    int x0i = (int)gx;
    if (x0i > gw - 2) x0i = gw - 2;
    if (x0i < 0) x0i = 0;
    const float *p00 = r0 + (size_t)x0i * 3, *p01 = r0 + (size_t)(x0i + (gw > 1)) * 3;
    const float *p10 = r1 + (size_t)x0i * 3, *p11 = r1 + (size_t)(x0i + (gw > 1)) * 3;
    if (!(qvalid(p00) && qvalid(p01) && qvalid(p10) && qvalid(p11))) return;
    float A[3], B[3], du[3], dv0[3], dv1[3];
    for (int k = 0; k < 3; k++) {
        A[k] = p00[k] + (p10[k] - p00[k]) * fy;
        B[k] = p01[k] + (p11[k] - p01[k]) * fy;
        du[k] = (p01[k] - p00[k]) * (1.0f - fy) + (p11[k] - p10[k]) * fy;
        dv0[k] = p10[k] - p00[k];
        dv1[k] = p11[k] - p01[k];
    }
    const float fx = gx - (float)x0i;
    P[0] = A[0] + (B[0] - A[0]) * fx;
    P[1] = A[1] + (B[1] - A[1]) * fx;
    P[2] = A[2] + (B[2] - A[2]) * fx;
    const float dv[3] = {dv0[0] + (dv1[0] - dv0[0]) * fx, dv0[1] + (dv1[1] - dv0[1]) * fx,
                         dv0[2] + (dv1[2] - dv0[2]) * fx};
    N[0] = du[1] * dv[2] - du[2] * dv[1];
    N[1] = du[2] * dv[0] - du[0] * dv[2];
    N[2] = du[0] * dv[1] - du[1] * dv[0];
    v3_norm(N);
}

/* ---- LOD choice (src/mc_render.c:2222-2250) ------------------------------ */
int orc_pick_lod(int nlods, float vox_per_pixel) {
    if (nlods <= 1) return 0;
    int L = 0;
    float v = vox_per_pixel;
    while (v >= 2.0f && L < nlods - 1) { v *= 0.5f; L++; }
    return L;
}

float orc_quad_spacing(const float *grid, int gw, int gh) {
    if (!grid || gw < 2 || gh < 1) return 1.0f;
    double sum = 0.0;
    int n = 0;
    const int probes = gh < 32 ? gh : 32;
    for (int i = 0; i < probes; i++) {
        const int gy = (int)(((int64_t)i * (gh - 1)) / (probes > 1 ? probes - 1 : 1));
        const int gx = (int)(((int64_t)i * (gw - 2)) / (probes > 1 ? probes - 1 : 1));
        const float *a = grid + ((size_t)gy * gw + gx) * 3, *b = a + 3;
        if (!qvalid(a) || !qvalid(b)) continue;
        const float dz = b[0] - a[0], dy = b[1] - a[1], dx = b[2] - a[2];
        sum += sqrtf(dz * dz + dy * dy + dx * dx);
        n++;
    }
    return n ? (float)(sum / n) : 1.0f;
}

/* ---- percentile and depth rays (src/mc_render.c:896-933) ------------------ */
/* The value of rank k (0-based, ascending) among a[0..n-1]: a quickselect
 * with Lomuto partitions around the last element (a is permuted). */
static float select_rank(float *a, int n, int k) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const float piv = a[hi];
        int st = lo;
        for (int i = lo; i < hi; ++i)
            if (a[i] < piv) { const float t = a[i]; a[i] = a[st]; a[st] = t; st++; }
        { const float t = a[hi]; a[hi] = a[st]; a[st] = t; }
        if (k == st) return a[st];
        if (k < st) hi = st - 1; else lo = st + 1;
    }
    return a[k];
}

/* PERCENTILE with the default rank 0.9: up to 1024 samples along the ray
 * (every stride-th step for deeper slabs), the value at rank
 * round(0.9 * (n - 1)). */
static uint8_t percentile_ray(const orc_level *L, const float *P, float nz, float ny, float nx, int filter,
                              float t0, float dt, int nsteps) {
    float buf[1024];
    int stride = (nsteps + 1023) / 1024;
    if (stride < 1) stride = 1;
    const float sz_ = dt * nz * stride, sy_ = dt * ny * stride, sx_ = dt * nx * stride;
    float pz = P[0] + t0 * nz, py = P[1] + t0 * ny, px = P[2] + t0 * nx;
    int n = 0;
    for (int it = 0; it < nsteps; it += stride) {
        buf[n++] = sample(L, pz, py, px, filter);
        pz += sz_; py += sy_; px += sx_;
    }
    if (!n) return 0;
    const int k = (int)(0.9f * (float)(n - 1) + 0.5f);
    return to_u8(select_rank(buf, n, k));
}

/* DEPTH with the default threshold 0: the first step whose value is above
 * 0, as 1 + 254 * step / (nsteps - 1) rounded, or 0 when no step is. */
static uint8_t depth_ray(const orc_level *L, const float *P, float nz, float ny, float nx, int filter, float t0,
                         float dt, int nsteps) {
    const float sz_ = dt * nz, sy_ = dt * ny, sx_ = dt * nx;
    float pz = P[0] + t0 * nz, py = P[1] + t0 * ny, px = P[2] + t0 * nx;
    for (int it = 0; it < nsteps; it++) {
        const float v = sample(L, pz, py, px, filter);
        if (v * (1.0f / 255.0f) > 0.0f) {
            const float frac = nsteps > 1 ? (float)it / (float)(nsteps - 1) : 0.0f;
            return (uint8_t)(1.0f + 254.0f * frac + 0.5f);
        }
        pz += sz_; py += sy_; px += sx_;
    }
    return 0;
}

/* ---- one pixel (src/mc_render.c:1150-1281) ----------------------------------- */
static uint8_t render_pixel(const orc_level *L, const float *P, const float *Nn, int filter, int comp, float t0,
                            float dt, int nsteps) {
    if (!pt_valid(P)) return 0;
    if (comp == ORC_COMP_NONE) return to_u8(sample(L, P[0], P[1], P[2], filter));
    float nz = Nn[0], ny = Nn[1], nx = Nn[2];
    const float n2 = nz * nz + ny * ny + nx * nx;
    if (n2 < 1e-12f) return to_u8(sample(L, P[0], P[1], P[2], filter));
    if (n2 < 0.9998f || n2 > 1.0002f) {
        const float nl = 1.0f / sqrtf(n2);
        nz *= nl; ny *= nl; nx *= nl;
    }
    if (comp == ORC_COMP_PERCENTILE) return percentile_ray(L, P, nz, ny, nx, filter, t0, dt, nsteps);
    if (comp == ORC_COMP_DEPTH) return depth_ray(L, P, nz, ny, nx, filter, t0, dt, nsteps);
    const float sz_ = dt * nz, sy_ = dt * ny, sx_ = dt * nx;
    float pz = P[0] + t0 * nz, py = P[1] + t0 * ny, px = P[2] + t0 * nx;
    /* alpha with the reference's defaults for zero alpha_min / alpha_opacity
     * (make_cfg, src/mc_render.c:838-840) */
    const float a_th = 0.0f, a_sc = 1.0f / (1.0f - 0.0f);
    float mn = 255.0f, mx = 0.0f, sum = 0.0f, sum2 = 0.0f, acc = 0.0f, A = 0.0f;
    int it = 0, done = 0;
#define ORC_ALPHA_STEP(v)                                  \
    do {                                                   \
        float a = ((v) * (1.0f / 255.0f) - a_th) * a_sc;   \
        if (a > 0.0f) {                                    \
            if (a > 1.0f) a = 1.0f;                        \
            acc += (1.0f - A) * a * (v);                   \
            A += (1.0f - A) * a;                           \
            if (A >= 0.98f) done = 1;                      \
        }                                                  \
    } while (0)
    if (filter == ORC_FILTER_TRILINEAR) {
        for (; it + 4 <= nsteps && !done; it += 4) {
            float v4[4];
            for (int k = 0; k < 4; k++) {
                v4[k] = trilinear(L, pz, py, px);
                pz += sz_; py += sy_; px += sx_;
            }
            if (comp == ORC_COMP_MIN) {
                for (int k = 0; k < 4; k++) if (v4[k] < mn) mn = v4[k];
            } else if (comp == ORC_COMP_MAX) {
                for (int k = 0; k < 4; k++) if (v4[k] > mx) mx = v4[k];
                if (mx >= 255.0f) done = 1;
            } else if (comp == ORC_COMP_MEAN) {
                sum += v4[0] + v4[1] + v4[2] + v4[3];
            } else if (comp == ORC_COMP_STDDEV) {
                for (int k = 0; k < 4; k++) { sum += v4[k]; sum2 += v4[k] * v4[k]; }
            } else {
                for (int k = 0; k < 4 && !done; k++) ORC_ALPHA_STEP(v4[k]);
            }
        }
    }
    for (; it < nsteps && !done; it++) {
        const float v = sample(L, pz, py, px, filter);
        if (comp == ORC_COMP_MIN) { if (v < mn) mn = v; }
        else if (comp == ORC_COMP_MAX) { if (v > mx) mx = v; }
        else if (comp == ORC_COMP_MEAN) sum += v;
        else if (comp == ORC_COMP_STDDEV) { sum += v; sum2 += v * v; }
        else ORC_ALPHA_STEP(v);
        pz += sz_; py += sy_; px += sx_;
    }
#undef ORC_ALPHA_STEP
    if (comp == ORC_COMP_MIN) return to_u8(mn);
    if (comp == ORC_COMP_MAX) return to_u8(mx);
    if (comp == ORC_COMP_MEAN) return to_u8(nsteps ? sum / (float)nsteps : 0.0f);
    if (comp == ORC_COMP_STDDEV) {
        if (!nsteps) return 0;
        const float m = sum / (float)nsteps;
        const float var = sum2 / (float)nsteps - m * m;
        return to_u8(var > 0.0f ? sqrtf(var) : 0.0f);
    }
    return to_u8(acc);
}

int orc_render_steps(float t0, float t1, float dt) {
    if (!(dt > 0.0f)) dt = 1.0f;
    if (t1 < t0) { const float t = t0; t0 = t1; t1 = t; }
    int n = 0;
    for (float t = t0; t <= t1 + 1e-4f; t += dt) n++;
    return n;
}

/* ---- mc_render_quad_lod (src/mc_render.c:2302-2314) ---------------------- */
int orc_render_quad_lod(int nlods, const orc_level *levels, const float *grid, int gw, int gh, float x0, float y0,
                        float step, int w, int h, int filter, int comp, float t0, float t1, float dt, uint8_t *out) {
    if (nlods < 1 || !levels || !grid || gw < 1 || gh < 1 || !out || w <= 0 || h <= 0) return -1;
    const int L = orc_pick_lod(nlods, step * orc_quad_spacing(grid, gw, gh));
    const float inv = 1.0f / (float)(1 << L);
    const float a = inv, b = 0.5f * inv - 0.5f;
    if (L > 0) { t0 = t0 * inv; t1 = t1 * inv; }
    float cdt = dt > 0.0f ? dt : 1.0f;
    float ct0 = t0, ct1 = t1;
    if (ct1 < ct0) { const float t = ct0; ct0 = ct1; ct1 = t; }
    const int nsteps = orc_render_steps(ct0, ct1, cdt);
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) {
            float P[3], N[3];
            orc_quad_point(grid, gw, gh, x0, y0, step, i, j, P, N);
            if (L > 0 && pt_valid(P))
                for (int k = 0; k < 3; k++) {
                    const float v = P[k] * a + b;
                    P[k] = v < 0.0f ? 0.0f : v;
                }
            out[(size_t)i * w + j] = render_pixel(&levels[L], P, N, filter, comp, ct0, cdt, nsteps);
        }
    return L;
}
