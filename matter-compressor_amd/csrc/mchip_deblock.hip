// mchip_deblock.hip -- decode-side deblocking filter of a decoded volume
// (mc_deblock, SuperOptimizer/matter-compressor src/mc_codec.c:561-595).
//
// The reference filters every 16-aligned block face of a z,y,x raster in
// place, axis after axis (all X faces, then all Y faces, then all Z faces).
// A face at coordinate c along an axis reads the voxels at c-2, c-1, c, c+1
// of a line and writes c-1 and c.  Faces of one axis are 16 apart, so no
// face reads or writes a voxel another face of the same axis writes: one
// pass per axis, one thread per (face, line), with the stream ordering the
// passes.  Addresses are formed as the reference forms them (flat offsets
// from the face voxel), so a face at the last coordinate of an axis reads
// its c+1 neighbour from the next row or plane, as the reference does; a
// neighbour past the end of the volume reads as 0 (the reference reads
// outside its buffer there).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mchip {

// One face pair (src/mc_codec.c:565-575): air on either side, or a step of at
// least beta, leaves the pair alone; otherwise both sides move by the clamped
// delta, kept within [1, 255].
__device__ __forceinline__ void deblock_pair(uint8_t *v, int64_t n, int64_t at, int64_t step, int beta, int tc) {
    const int a = v[at - step], b = v[at];
    if (!a || !b) return;
    const int d = b - a;
    if (d < 0 ? -d >= beta : d >= beta) return;
    const int p1 = v[at - 2 * step];
    const int q1 = at + step < n ? v[at + step] : 0;
    const int pp = p1 ? p1 : a, qq = q1 ? q1 : b;
    int delta = (4 * (b - a) + (pp - qq) + 4) >> 3;
    delta = delta > tc ? tc : (delta < -tc ? -tc : delta);
    int na = a + delta, nb = b - delta;
    na = na < 1 ? 1 : (na > 255 ? 255 : na);
    nb = nb < 1 ? 1 : (nb > 255 ? 255 : nb);
    v[at - step] = (uint8_t)na;
    v[at] = (uint8_t)nb;
}

// axis 0: X faces (item = (z*ny + y)*nfx + f), 1: Y faces (item = (z*nfy + f)*nx + x),
// 2: Z faces (item = (f*ny + y)*nx + x); face f sits at coordinate 16*(f+1).
__global__ __launch_bounds__(256) void k_deblock(uint8_t *__restrict__ v, int nz, int ny, int nx, int axis,
                                                 int64_t items, int beta, int tc) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= items) return;
    const int64_t n = (int64_t)nz * ny * nx, sy = nx, sz = (int64_t)ny * nx;
    int64_t at, step;
    if (axis == 0) {
        const int64_t nf = (nx - 1) / 16;
        const int64_t f = i % nf, row = i / nf;
        at = row * sy + 16 * (f + 1);
        step = 1;
    } else if (axis == 1) {
        const int64_t nf = (ny - 1) / 16;
        const int64_t x = i % nx, f = (i / nx) % nf, z = i / (nx * nf);
        at = z * sz + 16 * (f + 1) * sy + x;
        step = sy;
    } else {
        const int64_t yx = i % sz, f = i / sz;
        at = 16 * (f + 1) * sz + yx;
        step = sz;
    }
    deblock_pair(v, n, at, step, beta, tc);
}

void launch_deblock(uint8_t *v, int nz, int ny, int nx, float quality, hipStream_t s) {
    const int beta = (int)(3.0f * quality + 6.0f);     // flatness gate
    const int tc = (int)(0.5f * quality + 1.0f);       // largest correction
    const int64_t nfx = nx > 16 ? (nx - 1) / 16 : 0, nfy = ny > 16 ? (ny - 1) / 16 : 0,
                  nfz = nz > 16 ? (nz - 1) / 16 : 0;
    const int64_t items[3] = {(int64_t)nz * ny * nfx, (int64_t)nz * nfy * nx, nfz * ny * nx};
    for (int axis = 0; axis < 3; ++axis)
        if (items[axis] > 0)
            hipLaunchKernelGGL(k_deblock, dim3((unsigned)((items[axis] + 255) / 256)), dim3(256), 0, s, v, nz, ny, nx,
                               axis, items[axis], beta, tc);
}

}  // namespace mchip
