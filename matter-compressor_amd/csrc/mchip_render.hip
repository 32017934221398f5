// mchip_render.hip -- quad-surface render and LOD decimation on CDNA4.
//
// Render (src/mc_render.c:2302-2314, mc_render_quad_lod): one thread per
// pixel.  The thread generates the pixel's point and unit normal from the
// quad's control grid (mc_quad_gen, :1981-2052), remaps it to the chosen
// level (lod_rowgen, :2262-2274), walks the ray along the normal and reduces
// the samples (render_pixel, :1150-1281: trilinear or nearest; none, min,
// mean, max, alpha, stddev, percentile or depth; the trilinear walk takes
// groups of four samples -- the mean adds each group as ((v0+v1)+v2)+v3 --
// then single samples) and writes to_u8 (:793-795).  Every
// floating-point operation is the reference's, in its order: the library is
// built without contraction and with correctly rounded division and square
// root, so each pixel equals the reference renderer's bit for bit.
//
// A level is read from a dense raster in HBM, or from decoded 16^3 blocks in
// the serving arena through a per-block slot table.  Before a render reads
// arena blocks, the touch pass runs the same pixel walk with a source that
// records, instead of reading, every block a sample touches, so the serving
// loop can decode exactly those blocks first.
//
// Decimation (src/mc_archive.c:61-72): one thread per output voxel, the mean
// of the cell's nonzero voxels rounded half up.
#include <hip/hip_runtime.h>

#include <cmath>

#include "mchip_internal.hpp"

namespace mchip {

namespace {

__device__ __forceinline__ uint8_t to_u8(float v) {
    return (uint8_t)(v < 0.0f ? 0 : v > 255.0f ? 255 : (int)(v + 0.5f));
}
__device__ __forceinline__ bool isnan_bits(float f) { return (__float_as_uint(f) & 0x7FFFFFFFu) > 0x7F800000u; }
__device__ __forceinline__ bool pt_valid(const float *p) {
    if (p[0] != p[0] || p[1] != p[1] || p[2] != p[2]) return false;
    return p[0] >= 0.0f && p[1] >= 0.0f && p[2] >= 0.0f;
}
__device__ __forceinline__ bool qvalid(const float *p) {
    return !(p[0] == -1.0f && p[1] == -1.0f && p[2] == -1.0f) && p[0] == p[0] && p[1] == p[1] && p[2] == p[2];
}

__device__ __forceinline__ float lerp3(float c000, float c001, float c010, float c011, float c100, float c101,
                                       float c110, float c111, float dz, float dy, float dx) {
    const float c00 = c000 + (c001 - c000) * dx;
    const float c01 = c010 + (c011 - c010) * dx;
    const float c10 = c100 + (c101 - c100) * dx;
    const float c11 = c110 + (c111 - c110) * dx;
    const float c0 = c00 + (c01 - c00) * dy;
    const float c1 = c10 + (c11 - c10) * dy;
    return c0 + (c1 - c0) * dz;
}

// ---- voxel sources ------------------------------------------------------
struct DenseSrc {
    static constexpr bool kTouch = false;
    const uint8_t *v;
    int nz, ny, nx;
    __device__ float voxel(int z, int y, int x) const {
        if ((unsigned)z >= (unsigned)nz || (unsigned)y >= (unsigned)ny || (unsigned)x >= (unsigned)nx) return 0.0f;
        return (float)v[((size_t)z * ny + y) * nx + x];
    }
    __device__ float trilinear(int z0, int y0, int x0, float dz, float dy, float dx) const {
        if ((unsigned)z0 < (unsigned)(nz - 1) && (unsigned)y0 < (unsigned)(ny - 1) &&
            (unsigned)x0 < (unsigned)(nx - 1)) {
            const size_t sy = (size_t)nx, sz = (size_t)ny * nx;
            const uint8_t *p = v + (size_t)z0 * sz + (size_t)y0 * sy + x0;
            return lerp3(p[0], p[1], p[sy], p[sy + 1], p[sz], p[sz + 1], p[sz + sy], p[sz + sy + 1], dz, dy, dx);
        }
        return lerp3(voxel(z0, y0, x0), voxel(z0, y0, x0 + 1), voxel(z0, y0 + 1, x0), voxel(z0, y0 + 1, x0 + 1),
                     voxel(z0 + 1, y0, x0), voxel(z0 + 1, y0, x0 + 1), voxel(z0 + 1, y0 + 1, x0),
                     voxel(z0 + 1, y0 + 1, x0 + 1), dz, dy, dx);
    }
};

struct TableSrc {
    static constexpr bool kTouch = false;
    TableLevel t;
    int nz, ny, nx;
    __device__ const uint8_t *block(int bz, int by, int bx) const {
        bz -= t.obz; by -= t.oby; bx -= t.obx;
        if ((unsigned)bz >= (unsigned)t.gz || (unsigned)by >= (unsigned)t.gy || (unsigned)bx >= (unsigned)t.gx)
            return nullptr;
        const int s = t.table[((size_t)bz * t.gy + by) * t.gx + bx];
        return s < 0 ? nullptr : t.arena + (size_t)s * kNVox;
    }
    __device__ float voxel(int z, int y, int x) const {
        if ((unsigned)z >= (unsigned)nz || (unsigned)y >= (unsigned)ny || (unsigned)x >= (unsigned)nx) return 0.0f;
        const uint8_t *b = block(z >> 4, y >> 4, x >> 4);
        return b ? (float)b[((z & 15) << 8) | ((y & 15) << 4) | (x & 15)] : 0.0f;
    }
    __device__ float trilinear(int z0, int y0, int x0, float dz, float dy, float dx) const {
        if ((unsigned)z0 < (unsigned)(nz - 1) && (unsigned)y0 < (unsigned)(ny - 1) &&
            (unsigned)x0 < (unsigned)(nx - 1) && (z0 & 15) != 15 && (y0 & 15) != 15 && (x0 & 15) != 15) {
            const uint8_t *b = block(z0 >> 4, y0 >> 4, x0 >> 4);         // all corners in one block
            if (!b) return lerp3(0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, dz, dy, dx);
            const uint8_t *p = b + (((z0 & 15) << 8) | ((y0 & 15) << 4) | (x0 & 15));
            return lerp3(p[0], p[1], p[16], p[17], p[256], p[257], p[272], p[273], dz, dy, dx);
        }
        return lerp3(voxel(z0, y0, x0), voxel(z0, y0, x0 + 1), voxel(z0, y0 + 1, x0), voxel(z0, y0 + 1, x0 + 1),
                     voxel(z0 + 1, y0, x0), voxel(z0 + 1, y0, x0 + 1), voxel(z0 + 1, y0 + 1, x0),
                     voxel(z0 + 1, y0 + 1, x0 + 1), dz, dy, dx);
    }
};

// Records the blocks a sample reads (bit per block of the table's grid);
// every sample returns 0.  A workgroup's threads first deduplicate their
// blocks in an LDS hash set (kTouchSet entries, linear probing), so the
// bitmap in global memory takes one atomic per distinct block per workgroup.
constexpr int kTouchSet = 1024;
struct TouchSrc {
    static constexpr bool kTouch = true;
    uint32_t *bits;
    TableLevel t;
    int nz, ny, nx;
    mutable int last;
    int *set;                  // LDS, kTouchSet entries, -1 = empty
    __device__ void mark(int bz, int by, int bx) const {
        bz -= t.obz; by -= t.oby; bx -= t.obx;
        if ((unsigned)bz >= (unsigned)t.gz || (unsigned)by >= (unsigned)t.gy || (unsigned)bx >= (unsigned)t.gx) return;
        const int id = (bz * t.gy + by) * t.gx + bx;
        if (id == last) return;
        last = id;
        uint32_t h = ((uint32_t)id * 2654435761u) >> 22;                  // 10 bits
        for (int k = 0; k < 16; ++k, h = (h + 1) & (kTouchSet - 1)) {
            const int v = atomicCAS(set + h, -1, id);
            if (v == id) return;                      // another thread of the workgroup marked it
            if (v == -1) break;                       // first in this workgroup: mark globally
            if (k == 15) break;                       // set crowded: mark globally
        }
        atomicOr(bits + (id >> 5), 1u << (id & 31));
    }
    __device__ float voxel(int z, int y, int x) const {
        if ((unsigned)z >= (unsigned)nz || (unsigned)y >= (unsigned)ny || (unsigned)x >= (unsigned)nx) return 0.0f;
        mark(z >> 4, y >> 4, x >> 4);
        return 0.0f;
    }
    __device__ float trilinear(int z0, int y0, int x0, float, float, float) const {
        // corners z0, z0+1 (and y, x) inside the level; their distinct blocks
        const int za = max(z0, 0), zb = min(z0 + 1, nz - 1);
        const int ya = max(y0, 0), yb = min(y0 + 1, ny - 1);
        const int xa = max(x0, 0), xb = min(x0 + 1, nx - 1);
        if (za > zb || ya > yb || xa > xb) return 0.0f;
        for (int bz = za >> 4; bz <= zb >> 4; ++bz)
            for (int by = ya >> 4; by <= yb >> 4; ++by)
                for (int bx = xa >> 4; bx <= xb >> 4; ++bx) mark(bz, by, bx);
        return 0.0f;
    }
};

// ---- sampling (src/mc_render.c:58-89, :176-217) ---------------------------
template <class Src>
__device__ __forceinline__ float trilinear(const Src &s, float z, float y, float x) {
    const float zf = floorf(z), yf = floorf(y), xf = floorf(x);
    return s.trilinear((int)zf, (int)yf, (int)xf, z - zf, y - yf, x - xf);
}
template <class Src>
__device__ __forceinline__ float sample(const Src &s, float z, float y, float x, int filter) {
    if (isnan_bits(z) || isnan_bits(y) || isnan_bits(x)) return 0.0f;
    if (filter == 0) return s.voxel((int)floorf(z + 0.5f), (int)floorf(y + 0.5f), (int)floorf(x + 0.5f));
    return trilinear(s, z, y, x);
}

// ---- quad point and normal (src/mc_render.c:1981-2052) --------------------
__device__ void v3_norm(float *v) {
    const float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 1e-12f) { v[0] /= l; v[1] /= l; v[2] /= l; }
}

// The reference generates row i as mc_quad_gen(q, x0, y0 + i*step, step, w, 1)
// and caches per-cell values along the row; they depend only on the cell and
// the row, so each pixel recomputes them.
__device__ void quad_point(const RenderDesc &d, int row, int j, float *P, float *N) {
    P[0] = P[1] = P[2] = -1.0f;
    N[0] = N[1] = N[2] = 0.0f;
    const float yrow = d.y0 + (float)row * d.step;
    const float gy = yrow + (float)0 * d.step;
    if (!(gy >= 0.0f) || gy > (float)(d.gh - 1)) return;
    int y0i = (int)gy;
    if (y0i > d.gh - 2) y0i = d.gh - 2;
    if (y0i < 0) y0i = 0;
    const float fy = gy - (float)y0i;
    const float *r0 = d.grid + (size_t)y0i * d.gw * 3;
    const float *r1 = d.grid + (size_t)(y0i + (d.gh > 1)) * d.gw * 3;
    const float gx = d.x0 + (float)j * d.step;
    if (!(gx >= 0.0f) || gx > (float)(d.gw - 1)) return;
    int x0i = (int)gx;
    if (x0i > d.gw - 2) x0i = d.gw - 2;
    if (x0i < 0) x0i = 0;
    const float *p00 = r0 + (size_t)x0i * 3, *p01 = r0 + (size_t)(x0i + (d.gw > 1)) * 3;
    const float *p10 = r1 + (size_t)x0i * 3, *p11 = r1 + (size_t)(x0i + (d.gw > 1)) * 3;
    if (!(qvalid(p00) && qvalid(p01) && qvalid(p10) && qvalid(p11))) return;
    float A[3], B[3], du[3], dv0[3], dv1[3];
    #pragma unroll
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

// ---- one pixel (src/mc_render.c:1150-1281) --------------------------------
// The value of rank k (0-based, ascending) among a[0..n-1]: a quickselect
// with Lomuto partitions around the last element (a is permuted).
__device__ inline float select_rank(float *a, int n, int k) {
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

// PERCENTILE with the default rank 0.9 (pct_ray, src/mc_render.c:896-914):
// up to 1024 samples along the ray (every stride-th step for deeper slabs),
// the value at rank round(0.9 * (n - 1)).  kStore false walks the same
// samples without keeping them (the touch pass).
template <bool kStore, class Src>
__device__ uint8_t percentile_ray(const RenderDesc &d, const Src &s, const float *P, float nz, float ny, float nx) {
    float buf[kStore ? 1024 : 1];
    int stride = (d.nsteps + 1023) / 1024;
    if (stride < 1) stride = 1;
    const float sz_ = d.dt * nz * stride, sy_ = d.dt * ny * stride, sx_ = d.dt * nx * stride;
    float pz = P[0] + d.t0 * nz, py = P[1] + d.t0 * ny, px = P[2] + d.t0 * nx;
    int n = 0;
    for (int it = 0; it < d.nsteps; it += stride) {
        const float v = sample(s, pz, py, px, d.filter);
        if constexpr (kStore) buf[n] = v;
        n++;
        pz += sz_; py += sy_; px += sx_;
    }
    if constexpr (!kStore) return 0;
    if (!n) return 0;
    const int k = (int)(0.9f * (float)(n - 1) + 0.5f);
    return to_u8(select_rank(buf, n, k));
}

// DEPTH with the default threshold 0 (depth_ray, src/mc_render.c:917-933):
// the first step whose value is above 0, as 1 + 254 * step / (nsteps - 1)
// rounded, or 0 when no step is.
template <class Src>
__device__ uint8_t depth_ray(const RenderDesc &d, const Src &s, const float *P, float nz, float ny, float nx) {
    const float sz_ = d.dt * nz, sy_ = d.dt * ny, sx_ = d.dt * nx;
    float pz = P[0] + d.t0 * nz, py = P[1] + d.t0 * ny, px = P[2] + d.t0 * nx;
    for (int it = 0; it < d.nsteps; it++) {
        const float v = sample(s, pz, py, px, d.filter);
        if (v * (1.0f / 255.0f) > 0.0f) {
            const float frac = d.nsteps > 1 ? (float)it / (float)(d.nsteps - 1) : 0.0f;
            return (uint8_t)(1.0f + 254.0f * frac + 0.5f);
        }
        pz += sz_; py += sy_; px += sx_;
    }
    return 0;
}

// kPct: the instantiation that renders PERCENTILE (it keeps the ray's
// samples in private memory; the other composites need none).
template <class Src, bool kPct = false>
__device__ uint8_t render_pixel(const RenderDesc &d, const Src &s, int row, int j) {
    float P[3], N[3];
    quad_point(d, row, j, P, N);
    if (d.lod > 0 && pt_valid(P)) {
        #pragma unroll
        for (int k = 0; k < 3; k++) {
            const float v = P[k] * d.a + d.b;
            P[k] = v < 0.0f ? 0.0f : v;
        }
    }
    if (!pt_valid(P)) return 0;
    if (d.comp == 0) return to_u8(sample(s, P[0], P[1], P[2], d.filter));
    float nz = N[0], ny = N[1], nx = N[2];
    const float n2 = nz * nz + ny * ny + nx * nx;
    if (n2 < 1e-12f) return to_u8(sample(s, P[0], P[1], P[2], d.filter));
    if (n2 < 0.9998f || n2 > 1.0002f) {
        const float nl = 1.0f / sqrtf(n2);
        nz *= nl; ny *= nl; nx *= nl;
    }
    if (d.comp == 7) {
        if constexpr (Src::kTouch) return percentile_ray<false>(d, s, P, nz, ny, nx);
        else if constexpr (kPct) return percentile_ray<true>(d, s, P, nz, ny, nx);
        else return 0;                          // rendered by the kPct instantiation only
    }
    if (d.comp == 8) return depth_ray(d, s, P, nz, ny, nx);
    const float sz_ = d.dt * nz, sy_ = d.dt * ny, sx_ = d.dt * nx;
    float pz = P[0] + d.t0 * nz, py = P[1] + d.t0 * ny, px = P[2] + d.t0 * nx;
    float mn = 255.0f, mx = 0.0f, sum = 0.0f, sum2 = 0.0f, acc = 0.0f, A = 0.0f;
    int it = 0;
    bool done = false;
    // front-to-back alpha (comp 4): opacity from the value above a_th,
    // stopping at 98 % accumulated opacity
    auto alpha_step = [&](float v) {
        float a = (v * (1.0f / 255.0f) - d.a_th) * d.a_sc;
        if (a > 0.0f) {
            if (a > 1.0f) a = 1.0f;
            acc += (1.0f - A) * a * v;
            A += (1.0f - A) * a;
            if (A >= 0.98f) done = true;
        }
    };
    if (d.filter == 1) {
        for (; it + 4 <= d.nsteps && !done; it += 4) {
            float v4[4];
            #pragma unroll
            for (int k = 0; k < 4; k++) {
                v4[k] = trilinear(s, pz, py, px);
                pz += sz_; py += sy_; px += sx_;
            }
            if (d.comp == 1) {
                #pragma unroll
                for (int k = 0; k < 4; k++) if (v4[k] < mn) mn = v4[k];
            } else if (d.comp == 3) {
                #pragma unroll
                for (int k = 0; k < 4; k++) if (v4[k] > mx) mx = v4[k];
                if (mx >= 255.0f) done = true;
            } else if (d.comp == 2) {
                sum += v4[0] + v4[1] + v4[2] + v4[3];
            } else if (d.comp == 5) {
                #pragma unroll
                for (int k = 0; k < 4; k++) { sum += v4[k]; sum2 += v4[k] * v4[k]; }
            } else {
                for (int k = 0; k < 4 && !done; k++) alpha_step(v4[k]);
            }
        }
    }
    for (; it < d.nsteps && !done; it++) {
        const float v = sample(s, pz, py, px, d.filter);
        if (d.comp == 1) { if (v < mn) mn = v; }
        else if (d.comp == 3) { if (v > mx) mx = v; }
        else if (d.comp == 2) sum += v;
        else if (d.comp == 5) { sum += v; sum2 += v * v; }
        else alpha_step(v);
        pz += sz_; py += sy_; px += sx_;
    }
    if (d.comp == 1) return to_u8(mn);
    if (d.comp == 3) return to_u8(mx);
    if (d.comp == 2) return to_u8(d.nsteps ? sum / (float)d.nsteps : 0.0f);
    if (d.comp == 5) {
        if (!d.nsteps) return 0;
        const float m = sum / (float)d.nsteps;
        const float var = sum2 / (float)d.nsteps - m * m;
        return to_u8(var > 0.0f ? sqrtf(var) : 0.0f);
    }
    return to_u8(acc);
}

template <class Src, bool kPct>
__global__ __launch_bounds__(256) void k_render(RenderDesc d, Src s, uint8_t *__restrict__ out) {
    const int64_t pix = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (pix >= (int64_t)d.w * d.h) return;
    const int row = (int)(pix / d.w), j = (int)(pix % d.w);
    out[pix] = render_pixel<Src, kPct>(d, s, row, j);
}

__global__ __launch_bounds__(256) void k_render_touch(RenderDesc d, TouchSrc s) {
    __shared__ int set[kTouchSet];
    for (int i = threadIdx.x; i < kTouchSet; i += 256) set[i] = -1;
    __syncthreads();
    const int64_t pix = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (pix >= (int64_t)d.w * d.h) return;
    s.last = -1;
    s.set = set;
    (void)render_pixel(d, s, (int)(pix / d.w), (int)(pix % d.w));
}

__global__ __launch_bounds__(256) void k_bits_list(const uint32_t *__restrict__ bits, uint32_t nbits,
                                                    uint32_t *__restrict__ list, uint32_t *__restrict__ count) {
    const uint32_t w = blockIdx.x * 256 + threadIdx.x;
    if ((uint64_t)w * 32 >= nbits) return;
    uint32_t m = bits[w];
    if (!m) return;
    uint32_t at = atomicAdd(count, (uint32_t)__popc(m));
    while (m) {
        const int b = __ffs(m) - 1;
        m &= m - 1;
        list[at++] = w * 32 + b;
    }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *__restrict__ p, uint32_t v, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_table_set(int32_t *__restrict__ table, const int2 *__restrict__ upd, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) table[upd[i].x] = upd[i].y;
}

// ---- decimation (src/mc_archive.c:61-72) ----------------------------------
__device__ __forceinline__ uint8_t mean_nonzero(const uint32_t *v, int n) {
    int s = 0, c = 0;
    for (int k = 0; k < n; ++k)
        if (v[k]) { s += v[k]; c++; }
    return c ? (uint8_t)((s + c / 2) / c) : 0;
}

__global__ __launch_bounds__(256) void k_decimate(const uint8_t *__restrict__ src, int dz, int dy, int dx,
                                                   uint8_t *__restrict__ out) {
    const int hz = (dz + 1) / 2, hy = (dy + 1) / 2, hx = (dx + 1) / 2;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)hz * hy * hx) return;
    const int x = (int)(i % hx), y = (int)((i / hx) % hy), z = (int)(i / ((int64_t)hx * hy));
    uint32_t v[8];
    int n = 0;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int e = 0; e < 2; ++e) {
                const int sz = 2 * z + a, sy = 2 * y + b, sx = 2 * x + e;
                if (sz >= dz || sy >= dy || sx >= dx) continue;
                v[n++] = src[((size_t)sz * dy + sy) * dx + sx];
            }
    out[i] = mean_nonzero(v, n);
}

// One thread per 4 output voxels along x: each reads two 8-byte rows of
// each of the four (z, y) source rows of its cells.
__global__ __launch_bounds__(256) void k_decimate_group(ChunkGroup g, uint8_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;         // 256*256*64 threads
    const int x4 = i & 63, y = (i >> 6) & 255, z = i >> 14;
    const int sz0 = 2 * z, sy0 = 2 * y, sx0 = 8 * x4;            // source coordinates in the 512^3 group
    const uint8_t *c = g.c[((sz0 >> 8) * 2 + (sy0 >> 8)) * 2 + (sx0 >> 8)];
    uint32_t res = 0;
    if (c) {
        uint2 r[4];
        #pragma unroll
        for (int a = 0; a < 2; ++a)
            #pragma unroll
            for (int b = 0; b < 2; ++b)
                r[a * 2 + b] = *(const uint2 *)(c + (((size_t)((sz0 + a) & 255) << 16) | ((size_t)((sy0 + b) & 255) << 8) |
                                                     (size_t)(sx0 & 255)));
        #pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint32_t v[8];
            #pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t w = q < 2 ? r[k].x : r[k].y;
                const int sh = 16 * (q & 1);
                v[2 * k] = (w >> sh) & 255;
                v[2 * k + 1] = (w >> (sh + 8)) & 255;
            }
            res |= (uint32_t)mean_nonzero(v, 8) << (8 * q);
        }
    }
    *(uint32_t *)(out + ((size_t)z << 16) + ((size_t)y << 8) + 4 * x4) = res;
}

}  // namespace

// ---- host side -------------------------------------------------------------
static bool h_qvalid(const float *p) {
    return !(p[0] == -1.0f && p[1] == -1.0f && p[2] == -1.0f) && p[0] == p[0] && p[1] == p[1] && p[2] == p[2];
}

// mc_quad_spacing (src/mc_render.c:2232-2250)
static float quad_spacing(const float *grid, int gw, int gh) {
    if (!grid || gw < 2 || gh < 1) return 1.0f;
    double sum = 0.0;
    int n = 0;
    const int probes = gh < 32 ? gh : 32;
    for (int i = 0; i < probes; i++) {
        const int gy = (int)(((int64_t)i * (gh - 1)) / (probes > 1 ? probes - 1 : 1));
        const int gx = (int)(((int64_t)i * (gw - 2)) / (probes > 1 ? probes - 1 : 1));
        const float *a = grid + ((size_t)gy * gw + gx) * 3, *b = a + 3;
        if (!h_qvalid(a) || !h_qvalid(b)) continue;
        const float dz = b[0] - a[0], dy = b[1] - a[1], dx = b[2] - a[2];
        sum += sqrtf(dz * dz + dy * dy + dx * dx);
        n++;
    }
    return n ? (float)(sum / n) : 1.0f;
}

// mc_render_pick_lod (src/mc_render.c:2222-2230); every level is present
static int pick_lod(int nlods, float vpp) {
    if (nlods <= 1) return 0;
    int L = 0;
    float v = vpp;
    while (v >= 2.0f && L < nlods - 1) { v *= 0.5f; L++; }
    return L;
}

int render_describe(RenderDesc &d, int nlods, const int32_t *dims, const float *grid_host, const float *grid_dev,
                    int gw, int gh, float x0, float y0, float step, int w, int h, int filter, int comp, float t0,
                    float t1, float dt) {
    if ((filter != 0 && filter != 1) || comp < 0 || comp > 8 || comp == 6) return -1;   // no SHADED
    const int L = pick_lod(nlods, step * quad_spacing(grid_host, gw, gh));
    const float inv = 1.0f / (float)(1 << L);
    if (L > 0) { t0 = t0 * inv; t1 = t1 * inv; }       // render_lod (src/mc_render.c:2276-2288)
    // make_cfg (src/mc_render.c:827-835)
    d.dt = dt > 0.0f ? dt : 1.0f;
    if (t1 < t0) { const float tt = t0; t0 = t1; t1 = tt; }
    d.t0 = t0;
    d.nsteps = 0;
    for (float t = t0; t <= t1 + 1e-4f; t += d.dt) d.nsteps++;
    d.grid = grid_dev;
    d.gw = gw; d.gh = gh;
    d.x0 = x0; d.y0 = y0; d.step = step;
    d.w = w; d.h = h;
    d.filter = filter; d.comp = comp;
    // alpha parameters at the reference's defaults (alpha_min and
    // alpha_opacity zero: threshold 0, scale 1 / (1 - 0))
    d.a_th = 0.0f;
    d.a_sc = 1.0f / (1.0f - 0.0f);
    d.lod = L;
    d.a = inv;
    d.b = 0.5f * inv - 0.5f;
    d.nz = dims[3 * L]; d.ny = dims[3 * L + 1]; d.nx = dims[3 * L + 2];
    return L;
}

static unsigned grid_of(int64_t n) { return (unsigned)((n + 255) / 256); }

void launch_render_dense(const RenderDesc &d, const uint8_t *vol, uint8_t *out, hipStream_t s) {
    DenseSrc src{vol, d.nz, d.ny, d.nx};
    if (d.comp == 7)
        hipLaunchKernelGGL((k_render<DenseSrc, true>), dim3(grid_of((int64_t)d.w * d.h)), dim3(256), 0, s, d, src, out);
    else
        hipLaunchKernelGGL((k_render<DenseSrc, false>), dim3(grid_of((int64_t)d.w * d.h)), dim3(256), 0, s, d, src,
                           out);
}
void launch_render_table(const RenderDesc &d, const TableLevel &t, uint8_t *out, hipStream_t s) {
    TableSrc src{t, d.nz, d.ny, d.nx};
    if (d.comp == 7)
        hipLaunchKernelGGL((k_render<TableSrc, true>), dim3(grid_of((int64_t)d.w * d.h)), dim3(256), 0, s, d, src, out);
    else
        hipLaunchKernelGGL((k_render<TableSrc, false>), dim3(grid_of((int64_t)d.w * d.h)), dim3(256), 0, s, d, src,
                           out);
}
void launch_render_touch(const RenderDesc &d, const TableLevel &t, uint32_t *bits, hipStream_t s) {
    TouchSrc src{bits, t, d.nz, d.ny, d.nx, -1, nullptr};
    hipLaunchKernelGGL(k_render_touch, dim3(grid_of((int64_t)d.w * d.h)), dim3(256), 0, s, d, src);
}
void launch_bits_list(const uint32_t *bits, uint32_t nbits, uint32_t *list, uint32_t *count, hipStream_t s) {
    const uint32_t words = (nbits + 31) / 32;
    if (words) hipLaunchKernelGGL(k_bits_list, dim3(grid_of(words)), dim3(256), 0, s, bits, nbits, list, count);
}
void launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_u32, dim3(grid_of((int64_t)n)), dim3(256), 0, s, p, v, n);
}
void launch_table_set(int32_t *table, const int2 *upd, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_table_set, dim3(grid_of(n)), dim3(256), 0, s, table, upd, n);
}
void launch_decimate(const uint8_t *src, int dz, int dy, int dx, uint8_t *out, hipStream_t s) {
    const int64_t n = (int64_t)((dz + 1) / 2) * ((dy + 1) / 2) * ((dx + 1) / 2);
    if (n) hipLaunchKernelGGL(k_decimate, dim3(grid_of(n)), dim3(256), 0, s, src, dz, dy, dx, out);
}
void launch_decimate_group(const ChunkGroup &g, uint8_t *out, hipStream_t s) {
    hipLaunchKernelGGL(k_decimate_group, dim3(256 * 256 * 64 / 256), dim3(256), 0, s, g, out);
}

}  // namespace mchip
