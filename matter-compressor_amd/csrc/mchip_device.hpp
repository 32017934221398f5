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

namespace mchip {

constexpr int kNVox = 4096;
constexpr uint32_t kRcTop = 1u << 24;

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

// Byte source for the decoder: reads payload bytes through aligned 32-bit
// words and returns 0 past the end of the payload (src/mc_codec.c:197-205).
struct ByteSrc {
    const uint32_t *words;   // 4-byte aligned base of the containing buffer
    uint32_t start;          // byte offset of the payload in that buffer
    uint32_t len;            // payload length
    uint32_t pos;            // bytes consumed
    uint32_t cur_w;          // index of the cached word (0xFFFFFFFF = none)
    uint32_t cur;            // cached word

    __device__ __forceinline__ void init(const uint8_t *base, uint64_t off, uint32_t n) {
        uint64_t a = (uint64_t)(uintptr_t)(base + off);
        words = (const uint32_t *)(uintptr_t)(a & ~3ull);
        start = (uint32_t)(a & 3);
        len = n; pos = 0; cur_w = 0xFFFFFFFFu; cur = 0;
    }
    __device__ __forceinline__ uint32_t next() {
        if (pos >= len) return 0;
        uint32_t b = start + pos;
        uint32_t w = b >> 2;
        if (w != cur_w) { cur = words[w]; cur_w = w; }
        pos++;
        return (cur >> ((b & 3) * 8)) & 0xFFu;
    }
};

struct RcDec {
    ByteSrc src;
    uint32_t code, range;

    __device__ __forceinline__ void init(const uint8_t *base, uint64_t off, uint32_t n) {
        src.init(base, off, n);
        code = 0; range = 0xFFFFFFFFu;
        #pragma unroll
        for (int i = 0; i < 5; ++i) code = (code << 8) | src.next();
    }
    // Renormalise.  Any reachable state has range >= 1, which needs at most
    // three byte shifts; the bound keeps a zero range (only possible with a
    // corrupt prior override) from looping.
    __device__ __forceinline__ void norm() {
        #pragma unroll 1
        for (int i = 0; i < 4 && range < kRcTop; ++i) { code = (code << 8) | src.next(); range <<= 8; }
    }
    // adaptive binary decision with P(0) = p0/4096 (src/mc_codec.c:198-204)
    __device__ __forceinline__ int bit(uint32_t &p0) {
        uint32_t r0 = __umul24(range >> 12, p0);
        int b = code >= r0;
        if (b) { code -= r0; range -= r0; p0 -= p0 >> 4; }
        else   { range = r0; p0 += (4096u - p0) >> 4; }
        norm();
        return b;
    }
    // decision with a fresh context initialised to p0 (used-once contexts)
    __device__ __forceinline__ int bit_once(uint32_t p0) { uint32_t p = p0; return bit(p); }
    __device__ __forceinline__ int bypass() {
        range >>= 1;
        int b = code >= range;
        if (b) code -= range;
        norm();
        return b;
    }
    __device__ __forceinline__ uint32_t bypass_k16(int k) {   // 1 <= k <= 16
        range >>= k;
        uint32_t q = code / range, m = (1u << k) - 1u;
        if (q > m) q = m;
        code -= q * range;
        norm();
        return q;
    }
    // batched bypass of k bits (src/mc_codec.c:216-226)
    __device__ __forceinline__ uint32_t bypass_n(int k) {
        uint32_t v = 0;
        while (k > 16) { v = (v << 16) | bypass_k16(16); k -= 16; }
        if (k == 0) return v;
        return (v << k) | bypass_k16(k);
    }
    // order-0 Exp-Golomb in bypass bits (src/mc_codec.c:234-238).  A valid
    // stream never has more than 32 prefix ones; the cap only bounds the work
    // spent on corrupt input.
    __device__ __forceinline__ uint32_t expgolomb() {
        uint32_t nb = 0;
        while (nb < 64 && bypass()) nb++;
        if (!nb) return 0;
        return ((1u << (nb & 31)) | bypass_n((int)nb)) - 1u;
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

// Range encoder writing bytes into a 4-byte aligned buffer (global or LDS)
// through a word accumulator (src/mc_codec.c:178-195, :209-233).  Bytes
// beyond cap are counted but not stored; n is the full payload length.
struct RcEnc {
    uint32_t *out;
    uint32_t cap, n, wbuf;
    uint64_t low;
    uint32_t range, cache, pending;

    __device__ __forceinline__ void init(uint32_t *o, uint32_t capacity) {
        out = o; cap = capacity; n = 0; wbuf = 0;
        low = 0; range = 0xFFFFFFFFu; cache = 0; pending = 1;
    }
    __device__ __forceinline__ void put(uint32_t b) {
        if (n < cap) {
            wbuf |= (b & 0xFFu) << (8 * (n & 3));
            if ((n & 3) == 3) { out[n >> 2] = wbuf; wbuf = 0; }
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
        if (!b) { range = r0; p0 += (4096u - p0) >> 4; }
        else    { low += r0; range -= r0; p0 -= p0 >> 4; }
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
        if ((n & 3) && n < cap) out[n >> 2] = wbuf;
    }
};

}  // namespace mchip
