// mchip_encode.hip -- CDNA4 encode path of the 16^3-block codec.
//
// A dense 256^3 chunk becomes a v7 chunk blob (SuperOptimizer/
// matter-compressor src/mc_codec.c:809-845, src/mc_archive.c:822-901) in five
// kernels:
//
//   k_enc_transform  one workgroup per block: gather the 16^3 tile, count and
//                    sum (dc, material fraction), forward DCT in LDS (x, y, z
//                    passes), dead-zone quantisation, levels in scan order
//                    and the EOB; with tau > 0 also dequantise, inverse DCT
//                    and list the voxels whose error exceeds tau
//                    (src/mc_codec.c:629-709);
//   k_enc_entropy    one lane per block: the range encoder writes header,
//                    EOB, significance/magnitude/sign and correction bins
//                    into the block's slot (src/mc_codec.c:711-737);
//   k_enc_assemble   one workgroup per chunk: block bitmap, u16 length table
//                    and payload offsets by prefix scans, the fraction map
//                    (one lane, src/mc_codec.c:600-611), blob header;
//   k_enc_copy       one workgroup per block: payload into its final place;
//   k_enc_hash       one wave per chunk: the blob's XXH64 (mchip_hash.hip,
//                    src/mc_archive.c:76-88).
#include "mchip_codec.hpp"
#include "mchip_device.hpp"
#include "mchip_internal.hpp"

namespace mchip {

// c_cmf, c_band_start and g_scan are defined in mchip_decode.hip; both files
// are compiled as one translation unit (mchip_kernels.hip).

constexpr int kPadE = 17;

__device__ __forceinline__ uint32_t xcd_remap_e(uint32_t i, uint32_t n) {
    if (n % 8u) return i;
    return (i % 8u) * (n / 8u) + i / 8u;
}

// sum over a 256-thread workgroup (all threads get the result)
__device__ __forceinline__ uint32_t wg_sum(uint32_t v, uint32_t *lds4) {
    #pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0) lds4[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t r = lds4[0] + lds4[1] + lds4[2] + lds4[3];
    __syncthreads();
    return r;
}
__device__ __forceinline__ uint32_t wg_max(uint32_t v, uint32_t *lds4) {
    #pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, (uint32_t)__shfl_xor(v, d, 64));
    if ((threadIdx.x & 63) == 0) lds4[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t r = max(max(lds4[0], lds4[1]), max(lds4[2], lds4[3]));
    __syncthreads();
    return r;
}
__device__ __forceinline__ uint32_t wg_excl_scan(uint32_t v, uint32_t *lds4, uint32_t &total) {
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

// ---------------------------------------------------------------------------
// k_enc_transform
// ---------------------------------------------------------------------------
// kBlocks = false: chunk c's block bi is gathered from its 256^3 raster;
// kBlocks = true : independent 16^3 blocks, block g at in + 4096*g with its
//                  own descriptor cd[g] (single-block API).
template <bool kBlocks>
__global__ __launch_bounds__(256) void k_enc_transform(const uint8_t *__restrict__ in,
                                                       const EncChunkDesc *__restrict__ cd,
                                                       const QTab *__restrict__ qtabs,
                                                       uint4 *__restrict__ einfo,
                                                       int16_t *__restrict__ levels,
                                                       uint32_t *__restrict__ corr, int nchunks) {
    __shared__ float tile[256 * kPadE];
    __shared__ __attribute__((aligned(16))) uint8_t vox[kNVox];
    __shared__ int16_t lvr[kNVox];
    __shared__ uint32_t lds4[4];
    const uint32_t g = kBlocks ? blockIdx.x : xcd_remap_e(blockIdx.x, gridDim.x);
    const int c = kBlocks ? (int)g : (int)(g >> 12), bi = (int)(g & 4095), t = threadIdx.x;
    if (c >= nchunks) return;
    const size_t gb = kBlocks ? (size_t)g : (size_t)c * kNVox + bi;
    uint4 row;
    if (kBlocks) {
        row = *(const uint4 *)(in + (size_t)g * kNVox + (size_t)t * 16);
    } else {
        const uint8_t *src = in + cd[c].in_off + (size_t)(bi >> 8) * 16 * 65536 +
                             (size_t)((bi >> 4) & 15) * 16 * 256 + (size_t)(bi & 15) * 16;
        row = *(const uint4 *)(src + (size_t)(t >> 4) * 65536 + (size_t)(t & 15) * 256);
    }
    ((uint4 *)vox)[t] = row;
    uint32_t s = 0, n = 0;
    {
        const uint32_t w[4] = {row.x, row.y, row.z, row.w};
        #pragma unroll
        for (int j = 0; j < 4; ++j)
            #pragma unroll
            for (int b = 0; b < 4; ++b) {
                uint32_t v = (w[j] >> (8 * b)) & 255u;
                s += v;
                n += v != 0;
            }
    }
    const uint32_t sum = wg_sum(s, lds4), cnt = wg_sum(n, lds4);
    if (cnt == 0) {
        if (t == 0) einfo[gb] = make_uint4(0, 0, 0, 0);
        return;
    }
    const int dc = (int)((sum + 2048u) >> 12);
    uint32_t frac = (cnt * 15u + 2048u) / 4096u;
    if (!frac) frac = 1;
    const QTab *qt = qtabs + cd[c].qtab;
    const int tau = cd[c].tau;
    const float *cmf = c_cmf;
    float v[16];
    {   // x pass on line (z,y) = t
        #pragma unroll
        for (int x = 0; x < 16; ++x) v[x] = (float)((int)vox[t * 16 + x] - dc);
        fdct16_line(v, cmf);
        float *ln = tile + t * kPadE;
        #pragma unroll
        for (int k = 0; k < 16; ++k) ln[k] = v[k];
    }
    __syncthreads();
    {   // y pass on line (z,x)
        float *ln = tile + (t >> 4) * 16 * kPadE + (t & 15);
        #pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = ln[k * kPadE];
        fdct16_line(v, cmf);
        #pragma unroll
        for (int k = 0; k < 16; ++k) ln[k * kPadE] = v[k];
    }
    __syncthreads();
    int lv[16];
    {   // z pass on line (y,x), then quantise
        const int y = t >> 4, x = t & 15;
        float *ln = tile + y * kPadE + x;
        #pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = ln[k * 16 * kPadE];
        fdct16_line(v, cmf);
        #pragma unroll
        for (int k = 0; k < 16; ++k) {
            lv[k] = quant(v[k], qt->rstep[k + y + x]);
            lvr[k * 256 + t] = (int16_t)lv[k];
        }
    }
    __syncthreads();
    // levels in scan order for p in [16t, 16t+16) and the EOB
    int16_t sp[16];
    uint32_t last = 0, nz = 0;
    #pragma unroll
    for (int j = 0; j < 16; ++j) {
        sp[j] = lvr[g_scan[t * 16 + j]];
        if (sp[j]) { last = t * 16 + j + 1; nz++; }
    }
    const uint32_t eob = wg_max(last, lds4);
    const uint32_t nnz = wg_sum(nz, lds4);
    if ((uint32_t)t * 16 < eob) {
        int16_t *L = levels + gb * kNVox + t * 16;
        uint4 a, b;
        a.x = (uint16_t)sp[0] | ((uint32_t)(uint16_t)sp[1] << 16);
        a.y = (uint16_t)sp[2] | ((uint32_t)(uint16_t)sp[3] << 16);
        a.z = (uint16_t)sp[4] | ((uint32_t)(uint16_t)sp[5] << 16);
        a.w = (uint16_t)sp[6] | ((uint32_t)(uint16_t)sp[7] << 16);
        b.x = (uint16_t)sp[8] | ((uint32_t)(uint16_t)sp[9] << 16);
        b.y = (uint16_t)sp[10] | ((uint32_t)(uint16_t)sp[11] << 16);
        b.z = (uint16_t)sp[12] | ((uint32_t)(uint16_t)sp[13] << 16);
        b.w = (uint16_t)sp[14] | ((uint32_t)(uint16_t)sp[15] << 16);
        ((uint4 *)L)[0] = a;
        ((uint4 *)L)[1] = b;
    }
    uint32_t ncorr = 0;
    if (tau > 0) {
        // local reconstruction exactly as the decoder computes it (an
        // all-zero line transforms to zeros and is left as it is)
        uint64_t wk[16];
        {
            const int y = t >> 4, x = t & 15;
            float *ln = tile + y * kPadE + x;
            #pragma unroll
            for (int k = 0; k < 16; ++k) ln[k * 16 * kPadE] = dequant(lv[k], qt->step[k + y + x]);
        }
        __syncthreads();
        {
            float *ln = tile + t * kPadE;
            #pragma unroll
            for (int k = 0; k < 16; ++k) v[k] = ln[k];
            if (line_masks(v, wk)) idct16_line_sparse(v, cmf, wk);
            #pragma unroll
            for (int k = 0; k < 16; ++k) ln[k] = v[k];
        }
        __syncthreads();
        {
            float *ln = tile + (t >> 4) * 16 * kPadE + (t & 15);
            #pragma unroll
            for (int k = 0; k < 16; ++k) v[k] = ln[k * kPadE];
            if (line_masks(v, wk)) idct16_line_sparse(v, cmf, wk);
            #pragma unroll
            for (int k = 0; k < 16; ++k) ln[k * kPadE] = v[k];
        }
        __syncthreads();
        {
            float *ln = tile + (t >> 4) * kPadE + (t & 15);
            #pragma unroll
            for (int k = 0; k < 16; ++k) v[k] = ln[k * 16 * kPadE];
            if (line_masks(v, wk)) idct16_line_sparse(v, cmf, wk);
            #pragma unroll
            for (int k = 0; k < 16; ++k) {
                int rv = (int)((uint32_t)lrint_low32(v[k]) + (uint32_t)dc);
                rv = rv < 0 ? 0 : (rv > 255 ? 255 : rv);
                int err = (int)vox[k * 256 + t] - rv;
                int ae = err < 0 ? -err : err;
                int d = ae > tau ? (err < 0 ? -(ae - tau) : (ae - tau)) : 0;
                lvr[k * 256 + t] = (int16_t)d;     // reuse: error excess by raster
            }
        }
        __syncthreads();
        // compact the corrections in raster order
        int16_t dv[16];
        uint32_t cntc = 0;
        #pragma unroll
        for (int j = 0; j < 16; ++j) { dv[j] = lvr[t * 16 + j]; cntc += dv[j] != 0; }
        uint32_t base = wg_excl_scan(cntc, lds4, ncorr);
        uint32_t *C = corr + gb * kNVox;
        #pragma unroll
        for (int j = 0; j < 16; ++j)
            if (dv[j]) C[base++] = (uint32_t)(t * 16 + j) | ((uint32_t)(uint16_t)dv[j] << 16);
    }
    if (t == 0) einfo[gb] = make_uint4((uint32_t)dc | (frac << 8) | (1u << 12), nnz, eob, ncorr);
}

// ---------------------------------------------------------------------------
// k_enc_entropy: one lane per block (row blockIdx.y = the order[]-th chunk)
// ---------------------------------------------------------------------------
// Fraction map of one chunk (src/mc_codec.c:600-611): the 64 lanes stage the
// 4096 nibbles in LDS, lane 0 range-codes them.  Output at fm: [u32 length]
// [bytes]; the length is clamped to the buffer like the parallel-append
// stitch does (src/mc_archive.c:885-887).
__device__ void encode_fracmap_wg(const uint4 *__restrict__ ei, uint32_t *__restrict__ fm, uint32_t *lds) {
    uint8_t *frac = (uint8_t *)lds;                 // 4096 bytes of the staging area
    for (int i = threadIdx.x; i < kNVox; i += 64) frac[i] = (uint8_t)((ei[i].x >> 8) & 15u);
    __syncthreads();
    if (threadIdx.x != 0) return;
    RcEnc e;
    e.init(fm + 1, kFmapCap);
    // the three context sets (previous nibble 0, 1-14, 15) stay in registers:
    // the bin picks its set by selects, never by an indexed (scratch) array
    uint32_t c0[4], c1[4], c2[4];
    #pragma unroll
    for (int b = 0; b < 4; ++b) c0[b] = c1[b] = c2[b] = 2048;
    int prev = 0;
    for (int i = 0; i < kNVox; ++i) {
        const int v = frac[i];
        const bool z = prev == 0, f = prev == 15;
        #pragma unroll
        for (int b = 3; b >= 0; --b) {
            uint32_t p = z ? c0[b] : f ? c2[b] : c1[b];
            e.bit(p, (v >> b) & 1);
            c0[b] = z ? p : c0[b];
            c2[b] = f ? p : c2[b];
            c1[b] = (z || f) ? c1[b] : p;
        }
        prev = v;
    }
    e.finish();
    fm[0] = e.n < (uint32_t)kFmapCap ? e.n : (uint32_t)kFmapCap;
}

// k_enc_lane_order: one workgroup per chunk.  The entropy pass codes a
// block per lane, so a wave takes as long as its costliest lane and runs every
// branch any of its lanes takes.  Waves of blocks of like cost waste less: the
// chunk's blocks are counting-sorted by an estimate of their coding work,
// eob + 2 * nonzero levels + 4 * corrections (one scan position per eob, a
// magnitude and a sign per nonzero level, three Exp-Golomb codes per
// correction), costliest first, empty blocks last.  The order within a cost
// bucket is whatever the LDS atomics give; each block's stream goes to its own
// slot, so the blobs do not depend on it.
constexpr int kLaneBuckets = 1024;
__global__ __launch_bounds__(256) void k_enc_lane_order(const uint4 *__restrict__ einfo,
                                                        uint16_t *__restrict__ lanes) {
    __shared__ uint32_t cnt[kLaneBuckets];
    __shared__ uint32_t lds4[4];
    const uint4 *ei = einfo + (size_t)blockIdx.x * kNVox;
    uint16_t *lo = lanes + (size_t)blockIdx.x * kNVox;
    for (int i = threadIdx.x; i < kLaneBuckets; i += 256) cnt[i] = 0;
    __syncthreads();
    constexpr int kPer = kNVox / 256;
    uint32_t bk[kPer], at[kPer];
    #pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint4 inf = ei[threadIdx.x + 256 * k];
        const uint32_t cost = min(inf.z + 2u * inf.y + 4u * inf.w, (uint32_t)(kLaneBuckets - 2) * 16u);
        bk[k] = (inf.x >> 12) & 1u ? (uint32_t)(kLaneBuckets - 2) - (cost >> 4) : (uint32_t)(kLaneBuckets - 1);
        at[k] = atomicAdd(&cnt[bk[k]], 1u);
    }
    __syncthreads();
    // exclusive scan of the bucket counts, four buckets per thread
    uint32_t c4[4], s4 = 0;
    #pragma unroll
    for (int j = 0; j < 4; ++j) { c4[j] = cnt[threadIdx.x * 4 + j]; s4 += c4[j]; }
    uint32_t total;
    uint32_t base = wg_excl_scan(s4, lds4, total);
    #pragma unroll
    for (int j = 0; j < 4; ++j) { cnt[threadIdx.x * 4 + j] = base; base += c4[j]; }
    __syncthreads();
    #pragma unroll
    for (int k = 0; k < kPer; ++k) lo[cnt[bk[k]] + at[k]] = (uint16_t)(threadIdx.x + 256 * k);
}

template <bool kBlocks>
__global__ __launch_bounds__(64) void k_enc_entropy(const EncChunkDesc *__restrict__ cd,
                                                    const QTab *__restrict__ qtabs,
                                                    const uint4 *__restrict__ einfo,
                                                    const int16_t *__restrict__ levels,
                                                    const uint32_t *__restrict__ corr,
                                                    uint32_t *__restrict__ slots,
                                                    uint32_t *__restrict__ plen, uint32_t *__restrict__ fmapbuf,
                                                    int nblocks, const uint32_t *__restrict__ order,
                                                    const uint16_t *__restrict__ lanes) {
    // chunk of this row of workgroups (rows are dispatched in order)
    const uint32_t cy = kBlocks || !order ? blockIdx.y : order[blockIdx.y];
    __shared__ uint32_t stage[kEncStage * 64];
    static_assert(sizeof(stage) >= kNVox, "the fraction map stages its 4096 nibbles in the staging area");
    __shared__ QTab sq;
    __shared__ uint16_t sband[9];
    if (!kBlocks && blockIdx.x == 0) {   // one extra workgroup per chunk, dispatched first: the fraction map
        encode_fracmap_wg(einfo + (size_t)cy * kNVox, fmapbuf + (size_t)cy * (kFmapStride / 4), stage);
        return;
    }
    int bi = (kBlocks ? blockIdx.x : blockIdx.x - 1) * 64 + threadIdx.x;
    if (!kBlocks && lanes) bi = lanes[(size_t)cy * kNVox + bi];
    const QTab *qt;
    if (!kBlocks) {   // one chunk per workgroup: its priors and band starts into LDS
        const uint32_t *src = (const uint32_t *)(qtabs + cd[cy].qtab);
        uint32_t *dst = (uint32_t *)&sq;
        for (uint32_t i = threadIdx.x; i < sizeof(QTab) / 4; i += 64) dst[i] = src[i];
        if (threadIdx.x < 9) sband[threadIdx.x] = c_band_start[threadIdx.x];
        __syncthreads();
    }
    if (kBlocks && bi >= nblocks) return;
    const int c = kBlocks ? bi : (int)cy;
    const size_t gb = kBlocks ? (size_t)bi : (size_t)c * kNVox + bi;
    const uint4 inf = einfo[gb];
    if (!(inf.x & (1u << 12))) { plen[gb] = 0; return; }
    qt = kBlocks ? qtabs + cd[c].qtab : &sq;
    const uint32_t dc = inf.x & 255u, eob = inf.z, ncorr = inf.w;
    RcEnc e;
    e.init(slots + gb * (kSlotBytes / 4), kSlotBytes, stage + threadIdx.x, 64);
    encode_block_stream(e, qt, kBlocks ? c_band_start : sband, dc, eob, levels + gb * kNVox, ncorr,
                        corr + gb * kNVox);
    e.finish();
    plen[gb] = e.n;
}

// ---------------------------------------------------------------------------
// k_enc_assemble: one workgroup per chunk
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_enc_assemble(const EncChunkDesc *__restrict__ cd,
                                                      const uint4 *__restrict__ einfo,
                                                      const uint32_t *__restrict__ plen,
                                                      const uint32_t *__restrict__ fmapbuf,
                                                      uint32_t *__restrict__ dst,
                                                      uint8_t *__restrict__ out,
                                                      EncChunkResult *__restrict__ res) {
    __shared__ uint32_t lds4[4];
    const int c = blockIdx.x, t = threadIdx.x;
    const size_t cb = (size_t)c * kNVox;
    uint32_t bits = 0, lsum = 0, bad = 0;
    uint32_t lj[16];
    #pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int bi = t * 16 + j;
        const uint4 inf = einfo[cb + bi];
        const uint32_t present = (inf.x >> 12) & 1u;
        lj[j] = present ? plen[cb + bi] : 0u;
        bits |= present << j;
        lsum += lj[j];
        bad |= lj[j] > (uint32_t)kSlotBytes;
    }
    uint32_t np;
    const uint32_t slot0 = wg_excl_scan(__popc(bits), lds4, np);
    uint32_t paytot;
    const uint32_t pbase = wg_excl_scan(lsum, lds4, paytot);
    const uint32_t nbad = wg_sum(bad, lds4);
    const uint32_t *fm = fmapbuf + (size_t)c * (kFmapStride / 4);
    const uint32_t fml = fm[0];
    const uint64_t bm_off = 14ull + fml, lens_off = bm_off + 512, pay0 = lens_off + 2ull * np;
    const uint64_t total = np ? pay0 + paytot : 0;
    const bool fits = total <= cd[c].out_cap && nbad == 0;
    if (t == 0) {
        res[c].blob_len = total;
        res[c].status = nbad ? kEncSlotOverflow : (total > cd[c].out_cap ? kEncOutTooSmall : 0u);
        res[c].npresent = np;
        res[c].pay0 = (uint32_t)pay0;
    }
    if (!np || !fits) return;
    uint8_t *o = out + cd[c].out_off;
    if (t == 0) {
        const uint32_t qb = __float_as_uint(cd[c].q_store);
        for (int i = 0; i < 4; ++i) o[i] = (uint8_t)(qb >> (8 * i));
        for (int i = 4; i < 12; ++i) o[i] = 0;
        o[12] = (uint8_t)fml;
        o[13] = (uint8_t)(fml >> 8);
    }
    const uint8_t *fmb = (const uint8_t *)(fm + 1);
    for (uint32_t i = t; i < fml; i += 256) o[14 + i] = fmb[i];
    o[bm_off + 2 * t] = (uint8_t)bits;
    o[bm_off + 2 * t + 1] = (uint8_t)(bits >> 8);
    uint32_t s = slot0, at = pbase;
    #pragma unroll
    for (int j = 0; j < 16; ++j)
        if ((bits >> j) & 1) {
            o[lens_off + 2ull * s] = (uint8_t)lj[j];
            o[lens_off + 2ull * s + 1] = (uint8_t)(lj[j] >> 8);
            dst[cb + t * 16 + j] = (uint32_t)pay0 + at;
            at += lj[j];
            s++;
        }
}

// ---------------------------------------------------------------------------
// k_enc_copy: payload of block (c, bi) from its slot to the blob
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_enc_copy(const EncChunkDesc *__restrict__ cd,
                                                 const uint4 *__restrict__ einfo,
                                                 const uint32_t *__restrict__ plen,
                                                 const uint32_t *__restrict__ dst,
                                                 const uint32_t *__restrict__ slots,
                                                 const EncChunkResult *__restrict__ res,
                                                 uint8_t *__restrict__ out) {
    const uint32_t g = blockIdx.x;
    const int c = (int)(g >> 12), bi = (int)(g & 4095);
    const size_t gb = (size_t)c * kNVox + bi;
    if (res[c].status != 0 || res[c].blob_len == 0) return;
    if (!((einfo[gb].x >> 12) & 1u)) return;
    const uint32_t n = plen[gb];
    const uint8_t *s = (const uint8_t *)(slots + gb * (kSlotBytes / 4));
    uint8_t *o = out + cd[c].out_off + dst[gb];
    for (uint32_t i = threadIdx.x; i < n; i += 64) o[i] = s[i];
}

// ---------------------------------------------------------------------------
void launch_enc_transform(const EncodeArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_enc_transform<false>, dim3((unsigned)a.n * kNVox), dim3(256), 0, s, a.in, a.cd, a.qtabs,
                       a.einfo, a.levels, a.corr, a.n);
}
void launch_enc_entropy(const EncodeArgs &a, hipStream_t s) {
    if (a.lanes) hipLaunchKernelGGL(k_enc_lane_order, dim3((unsigned)a.n), dim3(256), 0, s, a.einfo, a.lanes);
    hipLaunchKernelGGL(k_enc_entropy<false>, dim3(kNVox / 64 + 1, a.n), dim3(64), 0, s, a.cd, a.qtabs, a.einfo,
                       a.levels, a.corr, a.slots, a.plen, a.fmap, 0, a.order, a.lanes);
}
// independent blocks: a.n blocks, descriptor per block
void launch_enc_blocks(const EncodeArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_enc_transform<true>, dim3((unsigned)a.n), dim3(256), 0, s, a.in, a.cd, a.qtabs, a.einfo,
                       a.levels, a.corr, a.n);
    hipLaunchKernelGGL(k_enc_entropy<true>, dim3(((unsigned)a.n + 63) / 64), dim3(64), 0, s, a.cd, a.qtabs,
                       a.einfo, a.levels, a.corr, a.slots, a.plen, nullptr, a.n, nullptr, nullptr);
}
void launch_enc_assemble(const EncodeArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_enc_assemble, dim3(a.n), dim3(256), 0, s, a.cd, a.einfo, a.plen, a.fmap, a.dst, a.out,
                       a.res);
}
void launch_enc_copy(const EncodeArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_enc_copy, dim3((unsigned)a.n * kNVox), dim3(64), 0, s, a.cd, a.einfo, a.plen, a.dst,
                       a.slots, a.res, a.out);
}

}  // namespace mchip
