// mchip_device.hpp -- device-side building blocks of the 16^3-block codec for
// CDNA4 (gfx950): the adaptive binary range decoder/encoder, the coefficient
// grammar, the dead-zone quantiser and the separable f32 DCT-16 line kernels.
//
// Arithmetic follows the reference codec bit for bit (SuperOptimizer/
// matter-compressor src/mc_codec.c, src/mc_codec_float.h).  Every f32
// expression is written as separate IEEE operations in the reference's order;
// the translation unit is compiled with -ffp-contract=off so that no
// multiply-add is fused.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Scheduling fence for the instruction scheduler (no code): keeps the
// serial decoders' long-latency reads issued early in each bin.
#ifndef MCHIP_SCHED_FENCE
#define MCHIP_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
#endif

namespace mchip {

constexpr int kNVox = 4096;
constexpr uint32_t kRcTop = 1u << 24;
constexpr uint32_t kStatusOk = 0;
constexpr uint32_t kStatusFallback = 1;   // block list overflowed: decode it with the single-lane path

// Per-quality tables built on the host (libm powf/log are evaluated there,
// exactly as the reference's context does: src/mc_codec.c:156-167, :541-552).
struct QTab {
    float step[46];          // q * (1+L1)^0.65 for L1 = 0..45
    float rstep[46];         // 1/step
    uint16_t sig[32];        // interpolated priors, class order of src/mc_codec.c:108
    uint16_t mag[12];
    uint16_t eob[14];
    uint16_t flag[2];
    uint16_t dc[8];
};

// Payload reader of the decoder: a 64-bit MSB-first bit buffer refilled one
// aligned 32-bit word at a time, with the following word loaded one refill
// ahead so the serial decode never waits on memory.  Bytes outside the
// payload read as 0, as in the reference decoder (src/mc_codec.c:197-205).
// Number of aligned payload words of a payload of n bytes at address a.
__device__ __forceinline__ int32_t payload_words(const uint8_t *p, uint32_t n) {
    const uint32_t lead = (uint32_t)((uintptr_t)p & 3);
    return n ? (int32_t)((lead + n - 1) >> 2) + 1 : 1;
}

struct RcDec {
    const uint32_t *wp;      // aligned word holding payload byte 4*k - lead
    int32_t k;               // index of the word in nextw
    int32_t last;            // index of the last word that holds payload bytes
    uint32_t lead;           // payload start offset inside word 0
    uint32_t len;
    uint32_t nextw;          // prefetched word k (masked, byte-swapped)
    uint64_t nb;             // upcoming payload bits, MSB first
    int32_t nbits;
    uint32_t code, range;

    // word i with bytes outside [0, len) cleared, first payload byte on top
    __device__ __forceinline__ uint32_t fetch(int32_t i) const {
        const int32_t ii = i > last ? last : i;
        uint32_t w = wp[ii];
        const int32_t b0 = 4 * i - (int32_t)lead;           // payload index of byte 0 of word i
        uint32_t keep = 0xFFFFFFFFu;
        if (b0 < 0) keep &= 0xFFFFFFFFu << (8 * (-b0));
        const int32_t room = (int32_t)len - b0;               // payload bytes from byte 0 on
        if (room < 4) keep = room <= 0 ? 0u : keep & (0xFFFFFFFFu >> (8 * (4 - room)));
        return __builtin_bswap32(w & keep);
    }
    __device__ __forceinline__ void refill() {
        nb |= (uint64_t)nextw << (32 - nbits);
        nbits += 32;
        k++;
        nextw = fetch(k);
    }
    __device__ __forceinline__ void init(const uint8_t *base, uint64_t off, uint32_t n) {
        const uint64_t a = (uint64_t)(uintptr_t)(base + off);
        wp = (const uint32_t *)(uintptr_t)(a & ~3ull);
        lead = (uint32_t)(a & 3);
        len = n;
        last = n ? (int32_t)((lead + n - 1) >> 2) : 0;
        nb = (uint64_t)fetch(0) << (32 + 8 * lead);           // drop the bytes before the payload
        nbits = 32 - 8 * (int32_t)lead;
        k = 1;
        nextw = fetch(1);
        refill();                                             // >= 40 bits buffered
        code = (uint32_t)(nb >> 32) << 8;                     // 5 bytes; the first one is shifted out
        code |= (uint32_t)(nb >> 24) & 0xFFu;
        nb <<= 40;
        nbits -= 40;
        range = 0xFFFFFFFFu;
        if (nbits < 32) refill();
    }
    // Shift in as many bytes as the range needs (clz(range)/8 of them).
    __device__ __forceinline__ void norm() {
        uint32_t kb = (uint32_t)__clz(range) >> 3;            // 0..3 for range >= 1
        kb = kb > 3 ? 3 : kb;
        const uint32_t sh = 8 * kb;
        const uint32_t top = (uint32_t)((nb >> (63 - sh)) >> 1);
        code = (code << sh) | top;
        range <<= sh;
        nb <<= sh;
        nbits -= (int32_t)sh;
        if (nbits < 32) refill();
    }
    // adaptive binary decision with P(0) = p0/4096 (src/mc_codec.c:198-204)
    __device__ __forceinline__ int bit(uint32_t &p0) {
        uint32_t r0 = __umul24(range >> 12, p0);
        int b = code >= r0;
        code = b ? code - r0 : code;
        range = b ? range - r0 : r0;
        p0 = b ? p0 - (p0 >> 4) : p0 + ((4096u - p0) >> 4);
        norm();
        return b;
    }
    // decision with a fresh context initialised to p0 (used-once contexts)
    __device__ __forceinline__ int bit_once(uint32_t p0) { uint32_t p = p0; return bit(p); }
    __device__ __forceinline__ int bypass() {
        range >>= 1;
        int b = code >= range;
        code = b ? code - range : code;
        norm();
        return b;
    }
    __device__ __forceinline__ uint32_t bypass_k16(int kk) {   // 1 <= kk <= 16
        range >>= kk;
        uint32_t q = code / range, m = (1u << kk) - 1u;
        if (q > m) q = m;
        code -= q * range;
        norm();
        return q;
    }
    // batched bypass of kk bits (src/mc_codec.c:216-226)
    __device__ __forceinline__ uint32_t bypass_n(int kk) {
        uint32_t v = 0;
        while (kk > 16) { v = (v << 16) | bypass_k16(16); kk -= 16; }
        if (kk == 0) return v;
        return (v << kk) | bypass_k16(kk);
    }
    // order-0 Exp-Golomb in bypass bits (src/mc_codec.c:234-238).  A valid
    // stream never has more than 32 prefix ones; the cap only bounds the work
    // spent on corrupt input.
    __device__ __forceinline__ uint32_t expgolomb() {
        uint32_t nb1 = 0;
        while (nb1 < 64 && bypass()) nb1++;
        if (!nb1) return 0;
        return ((1u << (nb1 & 31)) | bypass_n((int)nb1)) - 1u;
    }
};

// Dequantised value of level lv at step (src/mc_codec_float.h:185-190).
__device__ __forceinline__ float dequant(int lv, float step) {
    if (!lv) return 0.0f;
    float a = (float)(lv < 0 ? -lv : lv);
    float t0 = (a - 1.0f) * step;
    float t1 = 0.80f * step;
    float t2 = 0.40f * step;
    float r = (t0 + t1) + t2;
    return lv < 0 ? -r : r;
}

// Quantised level of coefficient c (src/mc_codec.c:686-692).
__device__ __forceinline__ int quant(float c, float rstep) {
    float t = fabsf(c) * rstep;
    t = t + (1.0f - 0.80f);
    t = t > 0.0f ? t : 0.0f;
    t = t < 32767.0f ? t : 32767.0f;
    int v = (int)t;
    return c < 0.0f ? -v : v;
}

// (int)lrintf(x) as the reference's x86-64 build evaluates it: round half to
// even, 64-bit conversion (out of range -> INT64_MIN), low 32 bits kept.
__device__ __forceinline__ int lrint_low32(float x) {
    float r = __builtin_rintf(x);
    float ar = fabsf(r);
    if (ar < 2147483648.0f) return (int)r;
    if (!(ar < 9223372036854775808.0f)) return 0;
    return (int)(long long)r;
}

// 1-D inverse DCT-III of one 16-point line, in place, with the reference's
// even/odd split and in-order accumulation from +0 (src/mc_codec_float.h:98-105).
// Skipping zero inputs (as the reference does) yields the same bits as the
// dense loop below: adding a signed zero never changes a sum that starts at +0.
__device__ __forceinline__ void idct16_line(float v[16], const float *cm /* [k*16+n] */) {
    float e[8], o[8];
    #pragma unroll
    for (int n = 0; n < 8; ++n) { e[n] = 0.0f; o[n] = 0.0f; }
    #pragma unroll
    for (int k = 0; k < 16; k += 2) {
        float c = v[k];
        #pragma unroll
        for (int n = 0; n < 8; ++n) { float p = cm[k * 16 + n] * c; e[n] = e[n] + p; }
    }
    #pragma unroll
    for (int k = 1; k < 16; k += 2) {
        float c = v[k];
        #pragma unroll
        for (int n = 0; n < 8; ++n) { float p = cm[k * 16 + n] * c; o[n] = o[n] + p; }
    }
    #pragma unroll
    for (int n = 0; n < 8; ++n) { v[n] = e[n] + o[n]; v[15 - n] = e[n] - o[n]; }
}

// The same transform for a wave's lines at once, skipping input k wherever no
// line of the wave has it nonzero (bit l of wk[k] is set when lane l's input
// k is nonzero; a uniform branch per input).  The reference skips each zero
// input of its line (src/mc_codec_float.h:109-126); a lane that runs an input
// of value zero because another lane needs it adds a signed zero to a sum that
// is never -0 (it starts at +0, its products are nonzero, and a sum that
// cancels rounds to +0), which leaves the sum unchanged.  Bit-identical to
// idct16_line.
__device__ __forceinline__ void idct16_line_sparse(float v[16], const float *cm, const uint64_t (&wk)[16]) {
    float e[8], o[8];
    #pragma unroll
    for (int n = 0; n < 8; ++n) { e[n] = 0.0f; o[n] = 0.0f; }
    #pragma unroll
    for (int k = 0; k < 16; k += 2) {
        if (wk[k]) {
            const float c = v[k];
            #pragma unroll
            for (int n = 0; n < 8; ++n) { float p = cm[k * 16 + n] * c; e[n] = e[n] + p; }
        }
    }
    #pragma unroll
    for (int k = 1; k < 16; k += 2) {
        if (wk[k]) {
            const float c = v[k];
            #pragma unroll
            for (int n = 0; n < 8; ++n) { float p = cm[k * 16 + n] * c; o[n] = o[n] + p; }
        }
    }
    #pragma unroll
    for (int n = 0; n < 8; ++n) { v[n] = e[n] + o[n]; v[15 - n] = e[n] - o[n]; }
}

// Per-input wave masks of a line (see idct16_line_sparse); returns whether
// this lane's line has any nonzero input (either sign of zero counts as zero).
__device__ __forceinline__ bool line_masks(const float (&v)[16], uint64_t (&wk)[16]) {
    bool nz = false;
    #pragma unroll
    for (int k = 0; k < 16; ++k) {
        const bool b = v[k] != 0.0f;
        wk[k] = __builtin_amdgcn_ballot_w64(b);
        nz |= b;
    }
    return nz;
}

// 1-D forward DCT-II of one line, in place (src/mc_codec_float.h:83-95).
// An all-zero line maps to zeros either way.
__device__ __forceinline__ void fdct16_line(float v[16], const float *cm) {
    float s[8], d[8];
    #pragma unroll
    for (int n = 0; n < 8; ++n) { s[n] = v[n] + v[15 - n]; d[n] = v[n] - v[15 - n]; }
    #pragma unroll
    for (int k = 0; k < 16; ++k) {
        float a = 0.0f;
        #pragma unroll
        for (int n = 0; n < 8; ++n) { float p = cm[k * 16 + n] * ((k & 1) ? d[n] : s[n]); a = a + p; }
        v[k] = a;
    }
}

}  // namespace mchip

namespace mchip {

// Range encoder (src/mc_codec.c:178-195, :209-233) writing bytes into a
// 16-byte aligned buffer through a word accumulator.  With a staging area
// (LDS on the GPU) completed words are staged and flushed in 16-byte stores
// every kEncStage words, keeping stores off the serial coding chain; without
// one (stage == nullptr) words go straight to out.  Bytes beyond cap are
// counted but not stored; n is the full payload length.  16 words per lane
// (4 KiB per 64-lane workgroup) rather than 32: the encoder's entropy pass
// then keeps 6 instead of 4.5 waves per SIMD (mixed-quality entropy pass
// 42.4 -> 37.3 ms for 256 chunks).
constexpr uint32_t kEncStage = 16;
struct RcEnc {
    uint32_t *out;
    uint32_t *st;
    uint32_t stride, base, nst;
    uint32_t cap, n, wbuf;
    uint64_t low;
    uint32_t range, cache, pending;

    __device__ __forceinline__ void init(uint32_t *o, uint32_t capacity, uint32_t *stage = nullptr,
                                         uint32_t stage_stride = 1) {
        out = o; st = stage; stride = stage_stride; base = 0; nst = 0;
        cap = capacity; n = 0; wbuf = 0;
        low = 0; range = 0xFFFFFFFFu; cache = 0; pending = 1;
    }
    __device__ __forceinline__ void flush_stage() {
        uint32_t i = 0;
        for (; i + 4 <= nst; i += 4)
            *(uint4 *)(out + base + i) = make_uint4(st[i * stride], st[(i + 1) * stride], st[(i + 2) * stride],
                                                    st[(i + 3) * stride]);
        for (; i < nst; ++i) out[base + i] = st[i * stride];
        base += nst;
        nst = 0;
    }
    __device__ __forceinline__ void put_word(uint32_t w) {
        if (!st) { out[base++] = w; return; }
        st[nst * stride] = w;
        if (++nst == kEncStage) flush_stage();
    }
    __device__ __forceinline__ void put(uint32_t b) {
        if (n < cap) {
            wbuf |= (b & 0xFFu) << (8 * (n & 3));
            if ((n & 3) == 3) { put_word(wbuf); wbuf = 0; }
        }
        n++;
    }
    __device__ __forceinline__ void shift_low() {
        uint32_t top = (uint32_t)(low >> 32);
        if (top != 0 || low < 0xFF000000ull) {
            put(cache + top);
            while (--pending) put(0xFFu + top);
            cache = (uint32_t)(low >> 24) & 0xFFu;
        }
        pending++;
        low = (low << 8) & 0xFFFFFFFFull;
    }
    __device__ __forceinline__ void norm() {
        #pragma unroll 1
        for (int i = 0; i < 4 && range < kRcTop; ++i) { shift_low(); range <<= 8; }
    }
    __device__ __forceinline__ void bit(uint32_t &p0, int b) {
        uint32_t r0 = __umul24(range >> 12, p0);
        low += b ? (uint64_t)r0 : 0ull;
        range = b ? range - r0 : r0;
        p0 = b ? p0 - (p0 >> 4) : p0 + ((4096u - p0) >> 4);
        norm();
    }
    __device__ __forceinline__ void bit_once(uint32_t p0, int b) { uint32_t p = p0; bit(p, b); }
    __device__ __forceinline__ void bypass(int b) {
        range >>= 1;
        if (b) low += range;
        norm();
    }
    __device__ __forceinline__ void bypass_k16(uint32_t v, int k) {   // 1 <= k <= 16
        range >>= k;
        low += (uint64_t)(v & ((1u << k) - 1u)) * range;
        norm();
    }
    __device__ __forceinline__ void bypass_n(uint32_t v, int k) {
        while (k > 16) { bypass_k16((v >> (k - 16)) & 0xFFFFu, 16); k -= 16; }
        if (k > 0) bypass_k16(v, k);
    }
    __device__ __forceinline__ void expgolomb(uint32_t v) {
        int nb = 31 - __clz(v + 1u);            // floor(log2(v+1))
        for (int i = 0; i < nb; ++i) bypass(1);
        bypass(0);
        if (nb) bypass_n((v + 1u) & ((1u << nb) - 1u), nb);
    }
    __device__ __forceinline__ void finish() {
        #pragma unroll 1
        for (int i = 0; i < 5; ++i) shift_low();
        if ((n & 3) && n < cap) put_word(wbuf);
        if (st) flush_stage();
    }
};

}  // namespace mchip
