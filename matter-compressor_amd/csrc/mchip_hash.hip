// mchip_hash.hip -- XXH64 on CDNA4: the chunk-blob hash of
// src/mc_archive.c:76-88 (XXH64 of the fraction map, the block bitmap, the
// u16 length table and the payloads, each seeded with the previous result,
// XORed together) for blobs being encoded and for blobs being verified
// (src/mc_archive.c:1072-1111), and plain XXH64 of byte ranges.
//
// XXH64's stripe phase is four independent accumulator chains, each the
// serial recurrence acc = rotl64(acc + w * P2, 31) * P1 over every fourth
// 8-byte word.  The recurrence does not compose into anything shorter, so a
// range's hash costs (stripes x one round's dependent latency) however many
// lanes help.  One wave hashes one range:
//
//   * all 64 lanes load 256-stripe tiles (8 KiB, coalesced; the next tile is
//     in flight while the current one is consumed) and multiply each word by
//     P2, the part of a round that is off the chain;
//   * the products wait in LDS;
//   * lanes 0..3 run the four chains from LDS, one round per stripe.
//
// The lead-out (merge of the accumulators, the < 32-byte tail, avalanche) is
// done by lane 0, as in the published XXH64 algorithm.

namespace mchip {

#define XP1 0x9E3779B185EBCA87ULL
#define XP2 0xC2B2AE3D27D4EB4FULL
#define XP3 0x165667B19E3779F9ULL
#define XP4 0x85EBCA77C2B2AE63ULL
#define XP5 0x27D4EB2F165667C5ULL

constexpr int kHashTile = 256;                  // stripes per LDS tile
constexpr int kHashWords = kHashTile * 4;       // 8-byte words per tile
constexpr int kHashPerLane = kHashWords / 64;   // words each lane loads

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
__device__ __forceinline__ uint64_t xround(uint64_t acc, uint64_t in) { return rotl64(acc + in * XP2, 31) * XP1; }
// a round whose input product w * P2 is already formed
__device__ __forceinline__ uint64_t xround_pre(uint64_t acc, uint64_t wp2) { return rotl64(acc + wp2, 31) * XP1; }

__device__ __forceinline__ uint64_t ld64u(const uint8_t *p) {
    uint64_t v = 0;
    #pragma unroll
    for (int i = 7; i >= 0; --i) v = (v << 8) | p[i];
    return v;
}
__device__ __forceinline__ uint32_t ld32u(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

// Finish an XXH64 after the stripe phase: p points at the unconsumed tail.
__device__ uint64_t xxh_tail(uint64_t h, const uint8_t *p, const uint8_t *end) {
    for (; p + 8 <= end; p += 8) { h ^= xround(0, ld64u(p)); h = rotl64(h, 27) * XP1 + XP4; }
    if (p + 4 <= end) { h ^= (uint64_t)ld32u(p) * XP1; h = rotl64(h, 23) * XP2 + XP3; p += 4; }
    for (; p < end; ++p) { h ^= (uint64_t)(*p) * XP5; h = rotl64(h, 11) * XP1; }
    h ^= h >> 33; h *= XP2; h ^= h >> 29; h *= XP3; h ^= h >> 32;
    return h;
}

// XXH64 of [p, p+len) with `seed`.  Every lane of the (single-wave)
// workgroup calls it with the same arguments; the result is returned in
// every lane.  tl is the workgroup's LDS tile of kHashWords words.
//
// Words are read as aligned dwords and funnelled (alignbit), so p may have
// any alignment.  A word whose bytes straddle a dword boundary reads the
// aligned dword holding its last byte; that dword lies inside the page of a
// byte of the range, so no read leaves mapped memory.
__device__ __forceinline__ uint64_t xxh64_wave(const uint8_t *p, uint64_t len, uint64_t seed, uint64_t *tl) {
    const int lane = threadIdx.x & 63;
    const uint64_t nstripe = len / 32;
    const uint64_t ntiles = (nstripe + kHashTile - 1) / kHashTile;
    // global address space: through a generic pointer the tile loads are flat
    // loads, which also count against lgkmcnt, so the chains' waits on their
    // LDS reads would wait for the next tile's loads too
    typedef const __attribute__((address_space(1))) uint32_t gword;
    gword *wb = (gword *)((uintptr_t)p & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)((uintptr_t)p & 3) * 8;
    uint64_t v = lane == 0 ? seed + XP1 + XP2 : lane == 1 ? seed + XP2 : lane == 2 ? seed : seed - XP1;

    uint32_t r0[kHashPerLane], r1[kHashPerLane], r2[kHashPerLane];
    auto load_tile = [&](uint64_t t) {
        gword *b = wb + 2 * t * kHashWords + 2 * lane;
        if ((t + 1) * kHashTile <= nstripe) {            // full tile: every word is in range
            #pragma unroll
            for (int i = 0; i < kHashPerLane; ++i) {
                r0[i] = b[128 * i];
                r1[i] = b[128 * i + 1];
            }
            if (sh) {
                #pragma unroll
                for (int i = 0; i < kHashPerLane; ++i) r2[i] = b[128 * i + 2];
            }
        } else {
            const uint32_t nw = (uint32_t)(nstripe - t * kHashTile) * 4;
            #pragma unroll
            for (int i = 0; i < kHashPerLane; ++i) {
                const bool in = (uint32_t)(lane + 64 * i) < nw;
                r0[i] = in ? b[128 * i] : 0;
                r1[i] = in ? b[128 * i + 1] : 0;
                r2[i] = in && sh ? b[128 * i + 2] : 0;
            }
        }
    };
    if (ntiles) load_tile(0);
    for (uint64_t t = 0; t < ntiles; ++t) {
        __syncthreads();                 // the chains of the previous tile are done with tl
        #pragma unroll
        for (int i = 0; i < kHashPerLane; ++i) {
            const uint32_t lo = __builtin_amdgcn_alignbit(r1[i], r0[i], sh);
            const uint32_t hi = __builtin_amdgcn_alignbit(r2[i], r1[i], sh);
            tl[lane + 64 * i] = ((uint64_t)hi << 32 | lo) * XP2;
        }
        __syncthreads();
        if (t + 1 < ntiles) load_tile(t + 1);
        if (lane < 4) {
            const uint32_t ns = (uint32_t)min((uint64_t)kHashTile, nstripe - t * kHashTile);
            const uint64_t *q = tl + lane;
            // groups of 8 rounds; the LDS reads of the next group are issued
            // before the rounds of the current one (held there by scheduling
            // fences)
            auto get = [&](uint64_t (&g)[8], uint32_t s0) {
                #pragma unroll
                for (int u = 0; u < 8; ++u) g[u] = q[4 * (s0 + u)];
            };
            auto rounds = [&](const uint64_t (&g)[8]) {
                #pragma unroll
                for (int u = 0; u < 8; ++u) v = xround_pre(v, g[u]);
            };
            uint32_t s = 0;
            if (ns >= 8) {
                uint64_t a[8], b[8];
                get(a, 0);
                for (; s + 24 <= ns; s += 16) {
                    get(b, s + 8);
                    MCHIP_SCHED_FENCE();
                    rounds(a);
                    get(a, s + 16);
                    MCHIP_SCHED_FENCE();
                    rounds(b);
                }
                if (s + 16 <= ns) {
                    get(b, s + 8);
                    MCHIP_SCHED_FENCE();
                    rounds(a);
                    rounds(b);
                    s += 16;
                } else {
                    rounds(a);
                    s += 8;
                }
            }
            for (; s < ns; ++s) v = xround_pre(v, q[4 * s]);
        }
    }
    const uint64_t v1 = __shfl(v, 0, 64), v2 = __shfl(v, 1, 64), v3 = __shfl(v, 2, 64), v4 = __shfl(v, 3, 64);
    uint64_t h = 0;
    if (lane == 0) {
        if (len >= 32) {
            h = rotl64(v1, 1) + rotl64(v2, 7) + rotl64(v3, 12) + rotl64(v4, 18);
            const uint64_t vv[4] = {v1, v2, v3, v4};
            for (int i = 0; i < 4; ++i) { h ^= xround(0, vv[i]); h = h * XP1 + XP4; }
        } else {
            h = seed + XP5;
        }
        h += len;
        h = xxh_tail(h, p + 32 * nstripe, p + len);
    }
    return __shfl(h, 0, 64);
}

// The chunk-blob hash of src/mc_archive.c:76-88 over blob bytes [o, o+total):
// header 14 bytes, fraction map (fml bytes), 512-byte bitmap, npresent u16
// lengths, payloads, each hashed with the previous result as seed (the
// first with the fixed seed "mcchunk") and XORed in.  The caller has checked
// that the structure fits in `total` (the payload offset is at most total).
__device__ __forceinline__ uint64_t blob_hash_wave(const uint8_t *o, uint64_t total, uint32_t fml, uint32_t np,
                                                   uint64_t *tl) {
    const uint64_t seg_off[4] = {14, 14ull + fml, 14ull + fml + 512, 14ull + fml + 512 + 2ull * np};
    uint64_t h = 0;
    #pragma unroll 1
    for (int k = 0; k < 4; ++k) {
        const uint64_t end = k < 3 ? seg_off[k + 1] : total;
        const uint64_t hk = xxh64_wave(o + seg_off[k], end - seg_off[k], k ? h : 0x6D636368756E6BULL, tl);
        h = k ? h ^ hk : hk;
    }
    return h;
}

// Blocks present in a blob: the popcount of its 512-byte bitmap, summed over
// the wave (every lane gets the count).
__device__ __forceinline__ uint32_t bitmap_count_wave(const uint8_t *bm) {
    const int lane = threadIdx.x & 63;
    uint32_t c = 0;
    #pragma unroll
    for (int k = 0; k < 8; ++k) c += __popc((uint32_t)bm[lane * 8 + k]);
    #pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    return c;
}

// k_enc_hash: one wave per encoded chunk; writes the hash into the blob
// header (bytes 4..11).
__global__ __launch_bounds__(64) void k_enc_hash(const EncChunkDesc *__restrict__ cd,
                                                 const EncChunkResult *__restrict__ res,
                                                 uint8_t *__restrict__ out) {
    __shared__ uint64_t tl[kHashWords];
    const int c = blockIdx.x;
    if (res[c].status != 0 || res[c].blob_len == 0) return;
    uint8_t *o = out + cd[c].out_off;
    const uint32_t fml = (uint32_t)o[12] | ((uint32_t)o[13] << 8);
    const uint64_t h = blob_hash_wave(o, res[c].blob_len, fml, res[c].npresent, tl);
    if (threadIdx.x < 8) o[4 + threadIdx.x] = (uint8_t)(h >> (8 * threadIdx.x));
}

// k_blob_hash: one wave per blob at base + off[i] of len[i] bytes.  out[i]
// receives the computed hash; bad[i] is 1 when the blob's structure does not
// fit in len[i] (then out[i] is 0) or the computed hash differs from the one
// stored in its header, else 0 (src/mc_archive.c:1093-1100).
__global__ __launch_bounds__(64) void k_blob_hash(const uint8_t *__restrict__ base, const uint64_t *__restrict__ off,
                                                  const uint64_t *__restrict__ len, int n, uint64_t *__restrict__ out,
                                                  uint8_t *__restrict__ bad) {
    __shared__ uint64_t tl[kHashWords];
    const int i = blockIdx.x;
    if (i >= n) return;
    const uint8_t *o = base + off[i];
    const uint64_t total = len[i];
    bool ok = total >= 14 + 512;
    uint32_t fml = 0, np = 0;
    if (ok) {
        fml = (uint32_t)o[12] | ((uint32_t)o[13] << 8);
        ok = 14ull + fml + 512 <= total;
    }
    if (ok) {
        np = bitmap_count_wave(o + 14 + fml);
        ok = 14ull + fml + 512 + 2ull * np <= total;
    }
    uint64_t h = 0;
    if (ok) h = blob_hash_wave(o, total, fml, np, tl);
    if (threadIdx.x == 0) {
        uint64_t want = 0;
        if (ok)
            for (int k = 7; k >= 0; --k) want = (want << 8) | o[4 + k];
        out[i] = h;
        bad[i] = (!ok || want != h) ? 1 : 0;
    }
}

// k_xxh64_ranges: plain XXH64 of n byte ranges, one wave each.
__global__ __launch_bounds__(64) void k_xxh64_ranges(const uint8_t *__restrict__ base,
                                                     const uint64_t *__restrict__ off,
                                                     const uint64_t *__restrict__ len,
                                                     const uint64_t *__restrict__ seed, int n,
                                                     uint64_t *__restrict__ out) {
    __shared__ uint64_t tl[kHashWords];
    const int i = blockIdx.x;
    if (i >= n) return;
    const uint64_t h = xxh64_wave(base + off[i], len[i], seed[i], tl);
    if (threadIdx.x == 0) out[i] = h;
}

void launch_enc_hash(const EncodeArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_enc_hash, dim3(a.n), dim3(64), 0, s, a.cd, a.res, a.out);
}
void launch_blob_hash(const uint8_t *base, const uint64_t *off, const uint64_t *len, int n, uint64_t *out,
                      uint8_t *bad, hipStream_t s) {
    hipLaunchKernelGGL(k_blob_hash, dim3(n), dim3(64), 0, s, base, off, len, n, out, bad);
}
void launch_xxh64_ranges(const uint8_t *base, const uint64_t *off, const uint64_t *len, const uint64_t *seed, int n,
                         uint64_t *out, hipStream_t s) {
    hipLaunchKernelGGL(k_xxh64_ranges, dim3(n), dim3(64), 0, s, base, off, len, seed, n, out);
}

}  // namespace mchip
