// mchip_decode.hip -- CDNA4 decode path of the 16^3-block codec.
//
// A chunk blob (format v7, SuperOptimizer/matter-compressor src/mc_archive.h:
// 147-164) is decoded by three kernels:
//
//   k_index   one workgroup per chunk: bitmap popcounts and the u16 length
//             table become per-block payload offsets through two workgroup
//             prefix scans (replaces mc_chunk_index, src/mc_archive.h:175-198);
//   k_entropy one lane per present block: the adaptive binary range decoder
//             with the trained priors walks header, EOB, significance,
//             magnitude, sign and correction bins (src/mc_codec.c:739-783) and
//             writes a compact (raster index, level) list;
//   k_recon   one workgroup per block: dequantisation into a 16^3 f32 tile in
//             LDS, the separable inverse DCT (x, y, z passes), round half to
//             even, + dc, clamp, sparse corrections, and 16-byte row stores
//             into the 256^3 raster.
//
// Blocks whose list would not fit the per-block slot (possible only for
// corrupt payloads) are decoded inside k_recon by a single lane straight into
// the LDS tile, so every payload decodes exactly like the reference.
#include "mchip_codec.hpp"
#include "mchip_lane_asm.hpp"
#include "mchip_device.hpp"
#include "mchip_internal.hpp"

namespace mchip {

__constant__ float c_cmf[256];           // cmf[k*16+n], host-built (double cos -> f32)
__constant__ uint16_t c_band_start[9];   // first scan position of each band, [8] = 4096
__device__ uint16_t g_scan[kNVox];       // scan position -> raster index

// ---------------------------------------------------------------------------
// k_index
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t blob_byte(const uint8_t *b, uint64_t len, uint64_t at) {
    return at < len ? (uint32_t)b[at] : 0u;
}

// Exclusive scan over a 256-thread workgroup; returns the prefix and sets total.
__device__ __forceinline__ uint32_t wg_exclusive_scan(uint32_t v, uint32_t *lds4, uint32_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
    #pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds4[wave] = x;
    __syncthreads();
    uint32_t w0 = lds4[0], w1 = lds4[1], w2 = lds4[2], w3 = lds4[3];
    uint32_t base = (wave > 0 ? w0 : 0) + (wave > 1 ? w1 : 0) + (wave > 2 ? w2 : 0);
    total = w0 + w1 + w2 + w3;
    __syncthreads();
    return base + x - v;
}

// Lane order of the entropy kernel: present blocks by decreasing payload
// length in 8-byte buckets, so each wave holds blocks of similar length and
// the longest chains start first.
constexpr uint32_t kLenBuckets = 1024;
__device__ __forceinline__ uint32_t len_bucket(uint32_t len, uint32_t shift) {
    return kLenBuckets - 1u - min(len >> shift, kLenBuckets - 1u);
}

__global__ __launch_bounds__(256) void k_index(const uint8_t *__restrict__ blobs,
                                               const DescRef cd,
                                               BlkRec *__restrict__ rec,
                                               uint16_t *__restrict__ plist,
                                               ChunkInfo *__restrict__ cinfo, uint32_t len_shift) {
    __shared__ uint32_t lds4[4];
    __shared__ uint32_t hist[kLenBuckets];
    const int c = blockIdx.x, t = threadIdx.x;
    for (uint32_t i = t; i < kLenBuckets; i += 256) hist[i] = 0;
    const uint8_t *b = blobs + cd[c].blob_off;
    const uint64_t L = cd[c].blob_len;
    bool ok = L >= 14;
    uint32_t fml = ok ? (blob_byte(b, L, 12) | (blob_byte(b, L, 13) << 8)) : 0u;
    uint64_t bm = 14ull + fml;
    ok = ok && bm + 512 <= L;
    uint32_t bits = ok ? (blob_byte(b, L, bm + 2 * t) | (blob_byte(b, L, bm + 2 * t + 1) << 8)) : 0u;
    uint32_t np;
    uint32_t slot0 = wg_exclusive_scan(__popc(bits), lds4, np);   // (its barriers also order the hist clear)
    uint64_t lens = bm + 512;
    if (!ok || lens + 2ull * np > L) { bits = 0; np = 0; }
    uint32_t sum = 0;
    for (int j = 0, s = (int)slot0; j < 16; ++j)
        if ((bits >> j) & 1) {
            uint64_t a = lens + 2ull * s++;
            sum += b[a] | ((uint32_t)b[a + 1] << 8);
        }
    uint32_t ptot;
    uint32_t pbase = wg_exclusive_scan(sum, lds4, ptot);
    uint64_t pay0 = lens + 2ull * np;
    uint64_t off = pay0 + pbase;
    // counting sort of the present blocks by length bucket: a block's place is
    // its bucket's base plus its rank among the bucket's blocks (the order of
    // the LDS atomics, which decides only which lane decodes the block)
    uint32_t place[16];
    #pragma unroll
    for (int j = 0; j < 16; ++j) place[j] = 0;
    #pragma unroll
    for (int j = 0, s = (int)slot0; j < 16; ++j) {
        int bi = t * 16 + j;
        BlkRec r = {0, 0, 0, 0};
        if ((bits >> j) & 1) {
            uint64_t a = lens + 2ull * s;
            uint32_t len = b[a] | ((uint32_t)b[a + 1] << 8);
            if (off + len <= L) { r.off = (uint32_t)off; r.len = len; r.slot = (uint32_t)s; }
            const uint32_t k = len_bucket(len, len_shift);
            place[j] = (k << 16) | atomicAdd(&hist[k], 1u);
            off += len;
            s++;
        }
        rec[(size_t)c * kNVox + bi] = r;
    }
    __syncthreads();
    // exclusive scan of the bucket counts, four buckets per thread
    const uint32_t h0 = hist[4 * t], h1 = hist[4 * t + 1], h2 = hist[4 * t + 2], h3 = hist[4 * t + 3];
    uint32_t htot;
    const uint32_t hb = wg_exclusive_scan(h0 + h1 + h2 + h3, lds4, htot);
    hist[4 * t] = hb;
    hist[4 * t + 1] = hb + h0;
    hist[4 * t + 2] = hb + h0 + h1;
    hist[4 * t + 3] = hb + h0 + h1 + h2;
    __syncthreads();
    #pragma unroll
    for (int j = 0; j < 16; ++j)
        if ((bits >> j) & 1) {
            const uint32_t at = hist[place[j] >> 16] + (place[j] & 0xFFFFu);
            plist[(size_t)c * kNVox + at] = (uint16_t)(t * 16 + j);
        }
    if (t == 0) { cinfo[c].np = np; cinfo[c].pay0 = (uint32_t)pay0; }
}

// ---------------------------------------------------------------------------
// k_entropy: one lane per present block of one chunk (blockIdx.y = chunk).
// ---------------------------------------------------------------------------
// Decoder variants (a tuning knob; all give identical results):
//   kDecPosition  one loop iteration per scan position (mchip_codec.hpp
//                 decode_coefs), payload read through a prefetched bit buffer;
//   kDecLane      one bin per iteration, branch-free, payload staged in a
//                 per-lane LDS window (LaneCoder).  In SIMT every branch body
//                 taken by any lane costs the whole wave, so the branch-free
//                 loop pays one bin body per bin instead of the union of the
//                 lanes' ladders;
//   kDecAsm       the lane decoder with its bin loop in assembly and event
//                 lanes paused and served together (mchip_lane_asm.hpp); the
//                 default.
constexpr int kDecPosition = 0, kDecLane = 2, kDecAsm = 3;

// Payload window words per lane of the lane decoders: 64 for the launches of
// one chunk (fewer window refills on the longest waves, which set a chunk's
// time); the persistent batched launch takes 16, 32 (the default) or 64
// (tuning "persist_win"), where a smaller LDS footprint lets more waves share
// each CU (batched plain64 207 -> 214 GB/s and mixed-quality 97 -> 114 GB/s
// at 32; one chunk alone is 2 % slower at 32).
constexpr int kWinGrid = kWin;

template <int kMode, int W = kWinGrid>
__device__ __forceinline__ uint2 entropy_lane(const uint8_t *pay, uint32_t len, const QTab *qt, const uint16_t *band,
                                              uint32_t *w, uint32_t cap, uint32_t *stage, uint32_t *ctx,
                                              uint32_t *win) {
    if constexpr (kMode == kDecAsm)
        return entropy_to_list_asm<W>(pay, len, qt, band, w, cap, ctx, win);
    else if constexpr (kMode == kDecLane)
        return entropy_to_list_lane<W>(pay, len, qt, band, w, cap, ctx, win);
    else
        return entropy_to_list(pay, len, qt, band, w, cap, stage, 64);
}

template <int kMode>
__global__ __launch_bounds__(64) void k_entropy(const uint8_t *__restrict__ blobs,
                                                const DescRef cd,
                                                const QTab *__restrict__ qtabs,
                                                const BlkRec *__restrict__ rec,
                                                const uint16_t *__restrict__ plist,
                                                const ChunkInfo *__restrict__ cinfo,
                                                uint32_t *__restrict__ scratch,
                                                uint2 *__restrict__ binfo) {
    constexpr bool kLaneMode = kMode == kDecLane || kMode == kDecAsm;
    __shared__ uint32_t stage[kLaneMode ? 1 : kStage * 64];
    __shared__ uint32_t sctx[kLaneMode ? kCtxSlots * 64 : 1];
    __shared__ uint32_t pwin[kLaneMode ? kWinAlloc * 64 : 1];
    __shared__ QTab sq;
    __shared__ uint16_t sband[9];
    const int c = blockIdx.y;
    const uint32_t np = cinfo[c].np;
    if (blockIdx.x * 64 >= np) return;
    {   // the chunk's priors and band starts into LDS (off the global-memory path)
        const uint32_t *src = (const uint32_t *)(qtabs + cd[c].qtab);
        uint32_t *dst = (uint32_t *)&sq;
        for (uint32_t i = threadIdx.x; i < sizeof(QTab) / 4; i += 64) dst[i] = src[i];
        if (threadIdx.x < 9) sband[threadIdx.x] = c_band_start[threadIdx.x];
    }
    __syncthreads();
    const uint32_t slot = blockIdx.x * 64 + threadIdx.x;
    if (slot >= np) return;
    const size_t cb = (size_t)c * kNVox;
    const int bi = plist[cb + slot];
    const BlkRec r = rec[cb + bi];
    if (r.off == 0 || r.len == 0) {            // invalid or empty: fallback path
        binfo[cb + bi] = make_uint2(kStatusFallback << 16, 0);
        return;
    }
    const uint8_t *pay = blobs + cd[c].blob_off + r.off;
    uint64_t wofs = cd[c].scratch_off + 8ull * (r.off - cinfo[c].pay0) + 32ull * r.slot;
    binfo[cb + bi] = entropy_lane<kMode>(pay, r.len, &sq, sband, scratch + wofs, 8u * r.len + 32u,
                                         stage + threadIdx.x, sctx + threadIdx.x, pwin + threadIdx.x);
}

// ---------------------------------------------------------------------------
// Persistent entropy decode over a batch of chunks of uneven cost (mixed
// qualities): the 64-block groups of all chunks (a chunk's group k holds its
// blocks of sorted rank 64k .. 64k+63, longest payloads first) are ordered by
// the payload length of their first block, longest first, and a machine-sized
// grid of workgroups takes them in that order through an atomic counter, so
// the heaviest groups start first and no workgroup idles while work remains.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_group_order(const BlkRec *__restrict__ rec,
                                                      const uint16_t *__restrict__ plist,
                                                      const ChunkInfo *__restrict__ cinfo, int n,
                                                      uint32_t *__restrict__ order, uint32_t *__restrict__ pctl,
                                                      uint32_t group_shift) {
    constexpr int kBuckets = 256;             // (1 << group_shift)-byte payload-length buckets, longest first
    __shared__ uint32_t hist[kBuckets], base[kBuckets];
    const int t = threadIdx.x;
    for (int i = t; i < kBuckets; i += 1024) hist[i] = 0;
    __syncthreads();
    const int ng = n * 64;
    auto key = [&](int g, bool &valid) -> uint32_t {
        const int c = g >> 6, k = g & 63;
        const uint32_t np = cinfo[c].np;
        valid = (uint32_t)k * 64 < np;
        if (!valid) return 0;
        const uint32_t len = rec[(size_t)c * kNVox + plist[(size_t)c * kNVox + k * 64]].len;
        const uint32_t b = len >> group_shift;
        return kBuckets - 1 - (b < kBuckets - 1 ? b : kBuckets - 1);
    };
    for (int g = t; g < ng; g += 1024) {
        bool v;
        const uint32_t b = key(g, v);
        if (v) atomicAdd(&hist[b], 1u);
    }
    __syncthreads();
    if (t == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < kBuckets; ++i) { base[i] = acc; acc += hist[i]; }
        pctl[0] = 0;                         // work counter (reset here, so a graph replay needs no memset)
        pctl[1] = acc;                       // groups
    }
    __syncthreads();
    for (int g = t; g < ng; g += 1024) {
        bool v;
        const uint32_t b = key(g, v);
        if (v) order[atomicAdd(&base[b], 1u)] = (uint32_t)g;
    }
}

template <int kMode, int W>
__global__ __launch_bounds__(64) void k_entropy_persist(const uint8_t *__restrict__ blobs, const DescRef cd,
                                                        const QTab *__restrict__ qtabs,
                                                        const BlkRec *__restrict__ rec,
                                                        const uint16_t *__restrict__ plist,
                                                        const ChunkInfo *__restrict__ cinfo,
                                                        uint32_t *__restrict__ scratch, uint2 *__restrict__ binfo,
                                                        const uint32_t *__restrict__ order,
                                                        uint32_t *__restrict__ pctl) {
    static_assert(kMode == kDecLane || kMode == kDecAsm, "persistent launch takes a lane decoder");
    __shared__ uint32_t sctx[kCtxSlots * 64];
    __shared__ uint32_t pwin[(W + 1) * 64];
    __shared__ QTab sq;
    __shared__ uint16_t sband[9];
    if (threadIdx.x < 9) sband[threadIdx.x] = c_band_start[threadIdx.x];
    const uint32_t ngroups = pctl[1];
    int cur = -1;                                // chunk whose quality table is in sq
    for (;;) {
        uint32_t w = 0;
        if (threadIdx.x == 0) w = atomicAdd(&pctl[0], 1u);
        w = __shfl(w, 0);                        // one wave per workgroup
        if (w >= ngroups) break;
        const uint32_t g = order[w];
        const int c = (int)(g >> 6), k = (int)(g & 63);
        __syncthreads();                         // the previous group is done with sq
        if (c != cur) {
            const uint32_t *src = (const uint32_t *)(qtabs + cd[c].qtab);
            uint32_t *dst = (uint32_t *)&sq;
            for (uint32_t i = threadIdx.x; i < sizeof(QTab) / 4; i += 64) dst[i] = src[i];
            cur = c;
        }
        __syncthreads();
        const uint32_t slot = (uint32_t)k * 64 + threadIdx.x;
        if (slot < cinfo[c].np) {
            const size_t cb = (size_t)c * kNVox;
            const int bi = plist[cb + slot];
            const BlkRec r = rec[cb + bi];
            if (r.off == 0 || r.len == 0) {
                binfo[cb + bi] = make_uint2(kStatusFallback << 16, 0);
            } else {
                const uint8_t *pay = blobs + cd[c].blob_off + r.off;
                const uint64_t wofs = cd[c].scratch_off + 8ull * (r.off - cinfo[c].pay0) + 32ull * r.slot;
                binfo[cb + bi] = entropy_lane<kMode, W>(pay, r.len, &sq, sband, scratch + wofs,
                                                                  8u * r.len + 32u, nullptr, sctx + threadIdx.x,
                                                                  pwin + threadIdx.x);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_recon: one 256-thread workgroup per block.
// ---------------------------------------------------------------------------
constexpr int kPad = 17;                  // padded LDS line stride (floats)

struct TileSink {                         // single-lane fallback: levels -> LDS
    float *tile;
    const QTab *qt;
    __device__ __forceinline__ void coef(uint32_t p, uint16_t lv) {
        int idx = g_scan[p];
        int l1 = (idx >> 8) + ((idx >> 4) & 15) + (idx & 15);
        tile[(idx >> 4) * kPad + (idx & 15)] = dequant((int)(int16_t)lv, qt->step[l1]);
    }
};

__device__ __forceinline__ uint8_t recon_voxel(float x, int dc) {
    int v = (int)((uint32_t)lrint_low32(x) + (uint32_t)dc);
    v = v < 0 ? 0 : (v > 255 ? 255 : v);
    return (uint8_t)v;
}

__device__ __forceinline__ void apply_correction(uint8_t *u, uint32_t pos, uint32_t delta) {
    int v = (int)((uint32_t)u[pos] + delta);
    v = v < 0 ? 0 : (v > 255 ? 255 : v);
    u[pos] = (uint8_t)v;
}

// One block decoded by lane 0 of the workgroup straight into the LDS tile,
// then transformed by all 256 lanes; leaves the 4096 voxels in u8t.
__device__ void recon_single_lane(const uint8_t *payload, uint32_t len, const QTab *__restrict__ qt,
                                  float *tile, uint8_t *u8t) {
    __shared__ int s_dc, s_flags, s_nsig;
    const int t = threadIdx.x;
    const float *cmf = c_cmf;
    for (int i = t; i < 256 * kPad; i += 256) tile[i] = 0.0f;
    __syncthreads();
    RcDec d;
    if (t == 0) {
        d.init(payload, 0, len);
        int dc, flags;
        decode_header(d, qt, dc, flags);
        TileSink sink{tile, qt};
        s_nsig = (int)decode_coefs(d, qt, c_band_start, sink);
        s_dc = dc; s_flags = flags;
    }
    __syncthreads();
    const int dc = s_dc, flags = s_flags;
    if (s_nsig == 0) {
        ((uint4 *)u8t)[t] = make_uint4(dc * 0x01010101u, dc * 0x01010101u, dc * 0x01010101u, dc * 0x01010101u);
    } else {
        // as in recon_from_list: zero lines stay zero, and inputs that no
        // line of the wave has nonzero are skipped
        float v[16];
        uint64_t wk[16];
        {
            float *ln = tile + t * kPad;
            #pragma unroll
            for (int k = 0; k < 16; ++k) v[k] = ln[k];
            if (line_masks(v, wk)) idct16_line_sparse(v, cmf, wk);
            #pragma unroll
            for (int k = 0; k < 16; ++k) ln[k] = v[k];
        }
        __syncthreads();
        {
            float *ln = tile + (t >> 4) * 16 * kPad + (t & 15);
            #pragma unroll
            for (int k = 0; k < 16; ++k) v[k] = ln[k * kPad];
            if (line_masks(v, wk)) idct16_line_sparse(v, cmf, wk);
            #pragma unroll
            for (int k = 0; k < 16; ++k) ln[k * kPad] = v[k];
        }
        __syncthreads();
        {
            float *ln = tile + (t >> 4) * kPad + (t & 15);
            #pragma unroll
            for (int k = 0; k < 16; ++k) v[k] = ln[k * 16 * kPad];
            if (line_masks(v, wk)) idct16_line_sparse(v, cmf, wk);
            #pragma unroll
            for (int k = 0; k < 16; ++k) u8t[k * 256 + t] = recon_voxel(v[k], dc);
        }
    }
    __syncthreads();
    if (t == 0 && (flags & 2)) {
        uint32_t n = d.expgolomb() + 1, pos = 0;
        if (n > (uint32_t)kNVox) n = kNVox;
        for (uint32_t i = 0; i < n; ++i) {
            pos += d.expgolomb();
            int neg = d.bypass();
            uint32_t m = d.expgolomb() + 1;
            if (pos >= (uint32_t)kNVox) break;
            apply_correction(u8t, pos, neg ? 0u - m : m);
        }
    }
    __syncthreads();
}

// Dequantise a level list into the LDS tile, inverse transform, round, add
// dc, clamp and apply the corrections; leaves the voxels in u8t.
__device__ __forceinline__ void recon_from_list(const uint2 inf, const uint32_t *__restrict__ w,
                                                const QTab *__restrict__ qt, float *tile, uint8_t *u8t) {
    const int t = threadIdx.x;
    const float *cmf = c_cmf;
    const int dc = inf.x & 255, flags = (inf.x >> 8) & 255;
    const uint32_t nw = inf.y & 0xFFFFu, nc = inf.y >> 16;
    if (nw == 0) {
        ((uint4 *)u8t)[t] = make_uint4(dc * 0x01010101u, dc * 0x01010101u, dc * 0x01010101u, dc * 0x01010101u);
    } else {
        for (int i = t; i < 256 * kPad; i += 256) tile[i] = 0.0f;
        __syncthreads();
        for (uint32_t i = t; i < nw; i += 256) {
            uint32_t e = w[i];
            int idx = g_scan[e & 0xFFFF];
            int l1 = (idx >> 8) + ((idx >> 4) & 15) + (idx & 15);
            tile[(idx >> 4) * kPad + (idx & 15)] = dequant((int)(int16_t)(e >> 16), qt->step[l1]);
        }
        __syncthreads();
        // A line whose 16 inputs are all zero transforms to zeros (every sum
        // starts at +0 and adding a signed zero never changes it), so its
        // transform is skipped: the tile already holds zeros there, and in
        // the last pass its voxels are dc.  Within the lines that do run, an
        // input that is zero in every line of the wave is skipped too
        // (idct16_line_sparse).  At q=6 that leaves about 62 % of the x and y
        // passes' products and 86 % of the z pass's.
        float v[16];
        uint64_t wk[16];
        {   // x pass: line (z,y) = t
            float *ln = tile + t * kPad;
            #pragma unroll
            for (int k = 0; k < 16; ++k) v[k] = ln[k];
            if (line_masks(v, wk)) {
                idct16_line_sparse(v, cmf, wk);
                #pragma unroll
                for (int k = 0; k < 16; ++k) ln[k] = v[k];
            }
        }
        __syncthreads();
        {   // y pass: line (z,x)
            float *ln = tile + (t >> 4) * 16 * kPad + (t & 15);
            #pragma unroll
            for (int k = 0; k < 16; ++k) v[k] = ln[k * kPad];
            if (line_masks(v, wk)) {
                idct16_line_sparse(v, cmf, wk);
                #pragma unroll
                for (int k = 0; k < 16; ++k) ln[k * kPad] = v[k];
            }
        }
        __syncthreads();
        {   // z pass: line (y,x); round, + dc, clamp into the byte tile
            float *ln = tile + (t >> 4) * kPad + (t & 15);
            #pragma unroll
            for (int k = 0; k < 16; ++k) v[k] = ln[k * 16 * kPad];
            if (line_masks(v, wk)) {
                idct16_line_sparse(v, cmf, wk);
                #pragma unroll
                for (int k = 0; k < 16; ++k) u8t[k * 256 + t] = recon_voxel(v[k], dc);
            } else {
                #pragma unroll
                for (int k = 0; k < 16; ++k) u8t[k * 256 + t] = recon_voxel(0.0f, dc);
            }
        }
    }
    if (flags & 2) {
        // Corrections in list order (src/mc_codec.c:770-781).  Their
        // positions are a running sum of gaps; when they strictly increase
        // (as in every stream the encoder writes) no voxel is corrected
        // twice, so the threads apply them in parallel; otherwise thread 0
        // applies them one after another.
        int dup = 0;
        for (uint32_t i = t; i + 1 < nc; i += 256) dup |= w[nw + 2 * i] >= w[nw + 2 * i + 2];
        dup = __syncthreads_or(dup);
        if (!dup) {
            for (uint32_t i = t; i < nc; i += 256) apply_correction(u8t, w[nw + 2 * i], w[nw + 2 * i + 1]);
        } else if (t == 0) {
            for (uint32_t i = 0; i < nc; ++i) apply_correction(u8t, w[nw + 2 * i], w[nw + 2 * i + 1]);
        }
    }
    __syncthreads();
}

// Map the linear workgroup id so that consecutive blocks (which share 128-byte
// output lines) run on the same XCD: hardware places workgroup i on XCD i % 8.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t i, uint32_t n) {
    if (n % 8u) return i;
    return (i % 8u) * (n / 8u) + i / 8u;
}

// kB blocks per workgroup (a tuning knob): workgroup g reconstructs blocks
// kB*g .. kB*g + kB-1 of the XCD-remapped order one after another.
template <int kB>
__global__ __launch_bounds__(256) void k_recon(const uint8_t *__restrict__ blobs,
                                               const DescRef cd,
                                               const QTab *__restrict__ qtabs,
                                               const BlkRec *__restrict__ rec,
                                               const ChunkInfo *__restrict__ cinfo,
                                               const uint32_t *__restrict__ scratch,
                                               const uint2 *__restrict__ binfo,
                                               uint8_t *__restrict__ out, int nchunks) {
    __shared__ float tile[256 * kPad];
    __shared__ __attribute__((aligned(16))) uint8_t u8t[kNVox];
    const uint32_t g0 = xcd_remap(blockIdx.x, gridDim.x) * kB;
    const int t = threadIdx.x;
    for (int k = 0; k < kB; ++k) {
        const uint32_t g = g0 + k;
        const int c = (int)(g >> 12), bi = (int)(g & 4095);
        if (c >= nchunks) return;
        const size_t cb = (size_t)c * kNVox;
        const BlkRec r = rec[cb + bi];
        uint8_t *ob = out + cd[c].out_off + ((size_t)(bi >> 8) * 16 * 65536) +
                      ((size_t)((bi >> 4) & 15) * 16 * 256) + (size_t)(bi & 15) * 16;
        uint4 *orow = (uint4 *)(ob + (size_t)(t >> 4) * 65536 + (size_t)(t & 15) * 256);
        if (r.off == 0) { *orow = make_uint4(0, 0, 0, 0); continue; }
        const uint2 inf = binfo[cb + bi];
        const QTab *qt = qtabs + cd[c].qtab;
        if ((inf.x >> 16) == kStatusOk) {
            const uint32_t *w = scratch + cd[c].scratch_off + 8ull * (r.off - cinfo[c].pay0) + 32ull * r.slot;
            recon_from_list(inf, w, qt, tile, u8t);
        } else {
            // single-lane fallback: thread 0 decodes the stream straight into LDS
            recon_single_lane(blobs + cd[c].blob_off + r.off, r.len, qt, tile, u8t);
        }
        *orow = ((const uint4 *)u8t)[t];
    }
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
hipError_t upload_constants(const float *cmf, const uint16_t *band_start, const uint16_t *scan) {
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_cmf), cmf, sizeof(float) * 256);
    if (e != hipSuccess) return e;
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_band_start), band_start, sizeof(uint16_t) * 9);
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(g_scan), scan, sizeof(uint16_t) * kNVox);
}

void launch_index(const DecodeArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_index, dim3(a.n), dim3(256), 0, s, a.blobs, a.cd, a.rec, a.plist, a.cinfo,
                       (uint32_t)a.len_shift);
}
template <int kMode>
static hipError_t occupancy_persist_mode(int win, int *per_cu) {
    if (win == 16) return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_entropy_persist<kMode, 16>, 64, 0);
    if (win == 64) return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_entropy_persist<kMode, 64>, 64, 0);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_entropy_persist<kMode, 32>, 64, 0);
}
hipError_t occupancy_persist(int decoder, int win, int *per_cu) {
    return decoder == kDecLane ? occupancy_persist_mode<kDecLane>(win, per_cu)
                               : occupancy_persist_mode<kDecAsm>(win, per_cu);
}

template <int kMode>
static void launch_persist(const DecodeArgs &a, hipStream_t s) {
    if (a.persist_win == 16)
        hipLaunchKernelGGL((k_entropy_persist<kMode, 16>), dim3(a.nwg), dim3(64), 0, s, a.blobs, a.cd, a.qtabs,
                           a.rec, a.plist, a.cinfo, a.scratch, a.binfo, a.order, a.pctl);
    else if (a.persist_win == 64)
        hipLaunchKernelGGL((k_entropy_persist<kMode, 64>), dim3(a.nwg), dim3(64), 0, s, a.blobs, a.cd, a.qtabs,
                           a.rec, a.plist, a.cinfo, a.scratch, a.binfo, a.order, a.pctl);
    else
        hipLaunchKernelGGL((k_entropy_persist<kMode, 32>), dim3(a.nwg), dim3(64), 0, s, a.blobs, a.cd, a.qtabs,
                           a.rec, a.plist, a.cinfo, a.scratch, a.binfo, a.order, a.pctl);
}

void launch_entropy(const DecodeArgs &a, hipStream_t s) {
    if (a.persistent && (a.decoder == kDecLane || a.decoder == kDecAsm)) {
        hipLaunchKernelGGL(k_group_order, dim3(1), dim3(1024), 0, s, a.rec, a.plist, a.cinfo, a.n, a.order, a.pctl,
                           (uint32_t)a.group_shift);
        if (a.decoder == kDecAsm) launch_persist<kDecAsm>(a, s);
        else launch_persist<kDecLane>(a, s);
        return;
    }
    if (a.decoder == kDecAsm)
        hipLaunchKernelGGL(k_entropy<kDecAsm>, dim3(kNVox / 64, a.n), dim3(64), 0, s, a.blobs, a.cd, a.qtabs, a.rec,
                           a.plist, a.cinfo, a.scratch, a.binfo);
    else if (a.decoder == kDecLane)
        hipLaunchKernelGGL(k_entropy<kDecLane>, dim3(kNVox / 64, a.n), dim3(64), 0, s, a.blobs, a.cd, a.qtabs, a.rec,
                           a.plist, a.cinfo, a.scratch, a.binfo);
    else
        hipLaunchKernelGGL(k_entropy<kDecPosition>, dim3(kNVox / 64, a.n), dim3(64), 0, s, a.blobs, a.cd, a.qtabs,
                           a.rec, a.plist, a.cinfo, a.scratch, a.binfo);
}
void launch_recon(const DecodeArgs &a, hipStream_t s) {
    const unsigned nb = (unsigned)a.n * kNVox;
    if (a.recon_bpw == 4)
        hipLaunchKernelGGL(k_recon<4>, dim3(nb / 4), dim3(256), 0, s, a.blobs, a.cd, a.qtabs, a.rec, a.cinfo,
                           a.scratch, a.binfo, a.out, a.n);
    else if (a.recon_bpw == 2)
        hipLaunchKernelGGL(k_recon<2>, dim3(nb / 2), dim3(256), 0, s, a.blobs, a.cd, a.qtabs, a.rec, a.cinfo,
                           a.scratch, a.binfo, a.out, a.n);
    else
        hipLaunchKernelGGL(k_recon<1>, dim3(nb), dim3(256), 0, s, a.blobs, a.cd, a.qtabs, a.rec, a.cinfo,
                           a.scratch, a.binfo, a.out, a.n);
}

// ---------------------------------------------------------------------------
// Block jobs: decode an arbitrary list of payloads into 4096-byte arena slots
// (region reads and the decoded-block arena).
// ---------------------------------------------------------------------------
template <int kMode>
__global__ __launch_bounds__(64) void k_entropy_jobs(const uint8_t *__restrict__ base,
                                                     const BlockJob *__restrict__ jobs, int n,
                                                     const QTab *__restrict__ qtabs,
                                                     uint32_t *__restrict__ scratch, uint2 *__restrict__ binfo) {
    constexpr bool kLaneMode = kMode == kDecLane || kMode == kDecAsm;
    __shared__ uint32_t stage[kLaneMode ? 1 : kStage * 64];
    __shared__ uint32_t sctx[kLaneMode ? kCtxSlots * 64 : 1];
    __shared__ uint32_t pwin[kLaneMode ? kWinAlloc * 64 : 1];
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    const BlockJob j = jobs[i];
    if (j.len == 0) { binfo[i] = make_uint2(kStatusFallback << 16, 0); return; }
    binfo[i] = entropy_lane<kMode>(base + j.pay_off, j.len, qtabs + j.qtab, c_band_start, scratch + j.scratch_off,
                                   8u * j.len + 32u, stage + threadIdx.x, sctx + threadIdx.x, pwin + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_recon_jobs(const uint8_t *__restrict__ base,
                                                    const BlockJob *__restrict__ jobs,
                                                    const QTab *__restrict__ qtabs,
                                                    const uint32_t *__restrict__ scratch,
                                                    const uint2 *__restrict__ binfo, uint8_t *__restrict__ arena) {
    __shared__ float tile[256 * kPad];
    __shared__ __attribute__((aligned(16))) uint8_t u8t[kNVox];
    const int i = blockIdx.x, t = threadIdx.x;
    const BlockJob j = jobs[i];
    const QTab *qt = qtabs + j.qtab;
    const uint2 inf = binfo[i];
    if ((inf.x >> 16) == kStatusOk) recon_from_list(inf, scratch + j.scratch_off, qt, tile, u8t);
    else recon_single_lane(base + j.pay_off, j.len, qt, tile, u8t);
    ((uint4 *)(arena + (size_t)j.slot * kNVox))[t] = ((const uint4 *)u8t)[t];
}

void launch_jobs(const uint8_t *base, const BlockJob *jobs, int n, const QTab *qtabs, uint32_t *scratch,
                 uint2 *binfo, uint8_t *arena, int decoder, hipStream_t s) {
    if (decoder == kDecAsm)
        hipLaunchKernelGGL(k_entropy_jobs<kDecAsm>, dim3((n + 63) / 64), dim3(64), 0, s, base, jobs, n, qtabs,
                           scratch, binfo);
    else if (decoder == kDecLane)
        hipLaunchKernelGGL(k_entropy_jobs<kDecLane>, dim3((n + 63) / 64), dim3(64), 0, s, base, jobs, n, qtabs,
                           scratch, binfo);
    else
        hipLaunchKernelGGL(k_entropy_jobs<kDecPosition>, dim3((n + 63) / 64), dim3(64), 0, s, base, jobs, n, qtabs,
                           scratch, binfo);
    hipLaunchKernelGGL(k_recon_jobs, dim3(n), dim3(256), 0, s, base, jobs, qtabs, scratch, binfo, arena);
}

// Assemble n boxes of dz x dy x dx voxels from arena blocks.  Box b's touched
// block grid starts at block (gz0,gy0,gx0) = box origin >> 4 and has
// gdim = (gz,gy,gx) cells in bmap[b * gz*gy*gx ...] (arena slot or -1 = 0).
// Each thread writes 16 consecutive voxels of one output row: they come from
// at most two 16-byte block rows (one 16-byte load each), aligned with byte
// funnel shifts.
__device__ __forceinline__ uint32_t pick4(uint32_t q, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    return q == 0 ? a : (q == 1 ? b : (q == 2 ? c : d));
}

__global__ __launch_bounds__(256) void k_crop(const uint8_t *__restrict__ arena, const int32_t *__restrict__ bmap,
                                              const int64_t *__restrict__ origin, int gz, int gy, int gx,
                                              int64_t dz, int64_t dy, int64_t dx, uint8_t *__restrict__ out,
                                              uint64_t batch_stride) {
    const int b = blockIdx.y;
    const int64_t segs = (dx + 15) / 16;
    const int64_t item = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (item >= dz * dy * segs) return;
    const int64_t x = (item % segs) * 16, y = (item / segs) % dy, z = item / (segs * dy);
    const int64_t z0 = origin[3 * b], y0 = origin[3 * b + 1], x0 = origin[3 * b + 2];
    const int64_t Z = z0 + z, Y = y0 + y, X = x0 + x;
    const int32_t *bm = bmap + (size_t)b * gz * gy * gx;
    const int cz = (int)((Z >> 4) - (z0 >> 4)), cy = (int)((Y >> 4) - (y0 >> 4));
    const int cxa = (int)((X >> 4) - (x0 >> 4));
    const uint32_t row = (uint32_t)((Z & 15) * 256 + (Y & 15) * 16);
    const uint32_t sh = (uint32_t)(X & 15);
    const int32_t *cell = bm + (cz * gy + cy) * gx;
    const int32_t sa = cxa < gx ? cell[cxa] : -1;
    const int32_t sb = (sh && cxa + 1 < gx) ? cell[cxa + 1] : -1;
    const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
    const uint4 L = sa >= 0 ? *(const uint4 *)(arena + (size_t)sa * kNVox + row) : zero;
    const uint4 R = sb >= 0 ? *(const uint4 *)(arena + (size_t)sb * kNVox + row) : zero;
    // bytes sh .. sh+15 of the 32-byte concatenation L|R
    const uint32_t q = sh >> 2, r = sh & 3;
    const uint32_t w0 = pick4(q, L.x, L.y, L.z, L.w), w1 = pick4(q, L.y, L.z, L.w, R.x),
                   w2 = pick4(q, L.z, L.w, R.x, R.y), w3 = pick4(q, L.w, R.x, R.y, R.z),
                   w4 = pick4(q, R.x, R.y, R.z, R.w);
    uint4 v;
    v.x = __builtin_amdgcn_alignbyte(w1, w0, r);
    v.y = __builtin_amdgcn_alignbyte(w2, w1, r);
    v.z = __builtin_amdgcn_alignbyte(w3, w2, r);
    v.w = __builtin_amdgcn_alignbyte(w4, w3, r);
    uint8_t *o = out + (size_t)b * batch_stride + (size_t)(z * dy + y) * dx + x;
    const int64_t n = dx - x < 16 ? dx - x : 16;
    if (n == 16 && (((uintptr_t)o) & 15) == 0) {
        *(uint4 *)o = v;
    } else {
        const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
        for (int k = 0; k < n; ++k) o[k] = (uint8_t)(vv[k >> 2] >> (8 * (k & 3)));
    }
}

void launch_crop(const uint8_t *arena, const int32_t *bmap, const int64_t *origin, int n, int gz, int gy, int gx,
                 int64_t dz, int64_t dy, int64_t dx, uint8_t *out, uint64_t batch_stride, hipStream_t s) {
    const int64_t items = dz * dy * ((dx + 15) / 16);
    hipLaunchKernelGGL(k_crop, dim3((unsigned)((items + 255) / 256), n), dim3(256), 0, s, arena, bmap, origin, gz, gy,
                       gx, dz, dy, dx, out, batch_stride);
}


// ---------------------------------------------------------------------------
// Fixed device workspace of the decode entry point that allocates nothing
// (mchip_decode_chunks_static): block records, lane order, block records of
// the entropy stage, the persistent launch's group order, the call's
// quality tables and the coefficient lists all live in this array.  Host
// code takes its device address once per device (static_scratch_address)
// and passes that address to the kernels like any workspace pointer.
// ---------------------------------------------------------------------------
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

hipError_t static_scratch_address(void **p) { return hipGetSymbolAddress(p, HIP_SYMBOL(scratch)); }

// The quality tables of one call travel in the kernel arguments.
struct QTabPack {
    QTab t[kInlineDesc];
};
__global__ __launch_bounds__(128) void k_put_qtabs(const QTabPack pack, int n, QTab *__restrict__ dst) {
    constexpr int kWords = (int)(sizeof(QTab) / 4);
    for (int i = threadIdx.x; i < n * kWords; i += 128)
        ((uint32_t *)dst)[i] = ((const uint32_t *)pack.t)[i];
}
void launch_put_qtabs(const QTab *tabs, int n, QTab *dst, hipStream_t s) {
    QTabPack pk;
    for (int i = 0; i < kInlineDesc; ++i) pk.t[i] = tabs[i < n ? i : 0];
    hipLaunchKernelGGL(k_put_qtabs, dim3(1), dim3(128), 0, s, pk, n, dst);
}

}  // namespace mchip
