// mchip_codec.hpp -- the block bitstream grammar on top of the range coder:
// header bins, EOB, significance / magnitude / sign bins and max-error
// corrections, for decoding (any sink) and encoding.  Shared by the HIP
// kernels and, through a host shim, by the CPU test harness of the device
// logic.  Follows SuperOptimizer/matter-compressor src/mc_codec.c:240-375,
// :711-783.
#pragma once
#include "mchip_device.hpp"

namespace mchip {

// ---------------------------------------------------------------------------
// Coefficient grammar (src/mc_codec.c:321-375), shared by the list decoder
// and the single-lane fallback.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t decode_magnitude(RcDec &d, uint32_t (&m)[12]) {
    uint32_t v = 0;
    bool done = false;
    #pragma unroll
    for (int k = 0; k < 11; ++k) {
        if (!done) {
            if (d.bit(m[k])) v++;
            else done = true;
        }
    }
    if (done) return v + 1;
    if (!d.bit(m[11])) return v + 1;
    return v + d.expgolomb() + 1;
}

template <class Sink>
__device__ __forceinline__ uint32_t decode_coefs(RcDec &d, const QTab *__restrict__ qt, const uint16_t *band_start,
                                                 Sink &sink) {
    int k = 0;
    #pragma unroll
    for (int i = 0; i < 13; ++i)
        if (k == i && d.bit_once(qt->eob[i])) k = i + 1;
    uint32_t eob = k == 0 ? 0u : (k == 1 ? 1u : ((1u << (k - 1)) | d.bypass_n(k - 1)));
    if (eob > (uint32_t)kNVox) eob = kNVox;
    uint32_t m[12];
    #pragma unroll
    for (int i = 0; i < 12; ++i) m[i] = qt->mag[i];
    uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int band = -1;
    uint32_t next_band = 0, hist = 0, nsig = 0;
    for (uint32_t p = 0; p < eob; ++p) {
        if (p >= next_band) {
            band++;
            s0 = qt->sig[band * 4 + 0]; s1 = qt->sig[band * 4 + 1];
            s2 = qt->sig[band * 4 + 2]; s3 = qt->sig[band * 4 + 3];
            next_band = band_start[band + 1];
        }
        int sig = 1;
        if (p + 1 != eob) {
            uint32_t dens = __popc(hist & 0xFFFFu);
            uint32_t p0 = dens == 0 ? s0 : dens == 1 ? s1 : dens == 2 ? s2 : s3;
            sig = d.bit(p0);
            s0 = dens == 0 ? p0 : s0;
            s1 = dens == 1 ? p0 : s1;
            s2 = dens == 2 ? p0 : s2;
            s3 = dens >= 3 ? p0 : s3;
        }
        hist = (hist << 1) | (uint32_t)sig;
        if (sig) {
            uint32_t mag = decode_magnitude(d, m);
            int neg = d.bypass();
            uint32_t lv = neg ? 0u - mag : mag;
            sink.coef(p, (uint16_t)lv);
            nsig++;
        }
    }
    return nsig;
}

// Word list of one block (coefficients as scan position | level << 16, then
// corrections as pos, delta word pairs).  Words go through a small staging
// area (LDS on the GPU, lane-interleaved) and reach global memory in 16-byte
// stores every kStage words, so the serial decode never waits on its own
// stores: on CDNA the vector-memory counter also counts stores, and a store
// still in flight would stall the next payload-word wait.
constexpr uint32_t kStage = 64;
struct ListSink {
    uint32_t *w;          // global destination, 16-byte aligned
    uint32_t *st;         // staging slot 0 of this lane
    uint32_t stride;      // words between consecutive staging slots
    uint32_t n, cap;      // words offered (may exceed cap) and capacity
    uint32_t base, nst;   // first unflushed index and staged count
    __device__ __forceinline__ void init(uint32_t *dst, uint32_t capacity, uint32_t *stage, uint32_t stage_stride) {
        w = dst; cap = capacity; st = stage; stride = stage_stride; n = 0; base = 0; nst = 0;
    }
    __device__ __forceinline__ void flush() {
        uint32_t i = 0;
        for (; i + 4 <= nst; i += 4)
            *(uint4 *)(w + base + i) = make_uint4(st[i * stride], st[(i + 1) * stride], st[(i + 2) * stride],
                                                  st[(i + 3) * stride]);
        for (; i < nst; ++i) w[base + i] = st[i * stride];
        base += nst;
        nst = 0;
    }
    __device__ __forceinline__ void put(uint32_t v) {
        if (n < cap) {
            st[nst * stride] = v;
            if (++nst == kStage) flush();
        }
        n++;
    }
    __device__ __forceinline__ void coef(uint32_t p, uint16_t lv) { put(p | ((uint32_t)lv << 16)); }
};

__device__ __forceinline__ void decode_header(RcDec &d, const QTab *__restrict__ qt, int &dc, int &flags) {
    int f = 0, v = 0;
    f |= d.bit_once(qt->flag[0]) ? 1 : 0;
    f |= d.bit_once(qt->flag[1]) ? 2 : 0;
    #pragma unroll
    for (int b = 0; b < 8; ++b) v = (v << 1) | d.bit_once(qt->dc[b]);
    dc = v; flags = f;
}

// ---------------------------------------------------------------------------
// Lane decoder: one bin per loop iteration for every lane, with the bin's
// arithmetic, context adaptation, renormalisation and state transition
// written as selects, so a wave pays one bin body per iteration instead of
// the union of its lanes' branch bodies.  Only rare events branch (band
// change, last position, Exp-Golomb suffix, window refill, list flush).
// Same bins, same order, same results as decode_coefs.
//
// The payload is read from a per-lane window of kWin words (LDS on the GPU,
// word j at win[j * kLaneStride]; one paired read with offsets j + 1 and j
// yields the two words as the low and high halves of a 64-bit value) holding
// the payload's aligned words byte-swapped to MSB-first, zero past the
// payload end (the reference reads zeros there, src/mc_codec.c:197).  rd_bits
// is the bit position of the next unread byte inside the window.
// ---------------------------------------------------------------------------
#ifndef MCHIP_LANE_WIN
#define MCHIP_LANE_WIN 64
#endif
constexpr int32_t kWin = MCHIP_LANE_WIN;   // default window words per lane (multiple of 16)
constexpr int32_t kWinAlloc = kWin + 1;   // words allocated per lane (see LaneCoder::norm_fast)

#ifndef MCHIP_LANE_STRIDE
#define MCHIP_LANE_STRIDE 64
#endif
constexpr uint32_t kLaneStride = MCHIP_LANE_STRIDE;   // lanes sharing one LDS array

// Fill win[0, W) with the payload's aligned words first .. first + W.
// Out of line: it runs once per W words.
template <int W>
__device__ __noinline__ void window_fill(const uint8_t *p, uint32_t len, int32_t first, uint32_t *win) {
    const uint32_t lead = (uint32_t)((uintptr_t)p & 3);
    const uint32_t *aw = (const uint32_t *)((uintptr_t)p - lead);
    const int32_t nwords = (int32_t)((lead + len + 3) >> 2);   // aligned words holding payload bytes
    #pragma unroll 1
    for (int32_t j0 = 0; j0 < W; j0 += 16) {
        if (first + j0 >= nwords) {                            // rest is past the payload
            #pragma unroll 1
            for (int32_t j = j0; j < W; ++j) win[j * kLaneStride] = 0u;
            break;
        }
        uint32_t a[16];
        #pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int32_t i = first + j0 + u;
            a[u] = i < nwords ? aw[i] : 0u;
        }
        #pragma unroll
        for (int u = 0; u < 16; ++u) {
            // aligned word i holds payload bytes 4i - lead .. 4i - lead + 3 (little-endian)
            const int64_t room = (int64_t)len + lead - 4ll * (first + j0 + u);
            const uint32_t m = room >= 4 ? 0xFFFFFFFFu : (room <= 0 ? 0u : 0xFFFFFFFFu >> (8 * (4 - (int32_t)room)));
            win[(j0 + u) * kLaneStride] = __builtin_bswap32(a[u] & m);
        }
    }
}

// A window of W words per lane (W + 1 allocated).
template <int W = kWin>
struct LaneCoder {
    static_assert(W % 16 == 0 && W >= 16, "whole window fills");
    const uint8_t *pay;
    uint32_t len;
    uint32_t *win;
    int32_t first;           // aligned payload word held in win[0]
    uint32_t rd_bits;        // bit position of the next unread byte in the window
    uint32_t w_hi, w_lo;     // window words rd_bits / 32 and the one after it
    uint32_t code, range;

    // Issue the window reads for the current position; they are consumed by
    // the next renormalisation, so their latency overlaps a whole bin.
    __device__ __forceinline__ void fetch() {
        const uint32_t *q = win + (rd_bits >> 5) * kLaneStride;
        w_hi = q[0];                      // word j
        w_lo = q[kLaneStride];            // word j + 1
    }
    __device__ __forceinline__ uint32_t next32() const {    // 32 bits (MSB first) at rd_bits
        return (uint32_t)(((((uint64_t)w_hi) << 32) | w_lo) << (rd_bits & 24) >> 32);
    }
    __device__ __forceinline__ bool need_slide() const { return rd_bits >= 32u * (W - 1); }
    __device__ __forceinline__ void slide() {
        first += (int32_t)(rd_bits >> 5);
        rd_bits &= 31u;
        window_fill<W>(pay, len, first, win);
        fetch();
    }
    __device__ __forceinline__ void init(const uint8_t *p, uint32_t n, uint32_t *w) {
        pay = p; len = n; win = w; first = 0;
        window_fill<W>(p, n, 0, win);
        // code = payload bytes 1..4 (byte 0 is the encoder's leading cache byte)
        rd_bits = 8u * (uint32_t)((uintptr_t)p & 3) + 8u;
        fetch();
        code = next32();
        rd_bits += 32;
        fetch();
        range = 0xFFFFFFFFu;
    }
    // Renormalise: shift in 0..3 bytes until range >= 2^24.  The _fast form
    // leaves a window slide to the caller: with rd_bits below the slide
    // threshold on entry it reads at most word W (W + 1 words are
    // allocated), and the words are valid again after slide().
    __device__ __forceinline__ void norm_fast() {
        const uint32_t sh = (uint32_t)__builtin_clz(range) & 24u;   // range >= 1
        code = (uint32_t)(((((uint64_t)code) << 32) | next32()) << sh >> 32);
        range <<= sh;
        rd_bits += sh;
        fetch();
    }
    __device__ __forceinline__ void norm() {
        norm_fast();
        if (need_slide()) slide();
    }
    __device__ __forceinline__ uint32_t bit(uint32_t p0) {      // fixed-probability bin
        const uint32_t r0 = __umul24(range >> 12, p0);
        const uint32_t b = code >= r0;
        code = b ? code - r0 : code;
        range = b ? range - r0 : r0;
        norm();
        return b;
    }
    __device__ __forceinline__ uint32_t bypass() {
        range >>= 1;
        const uint32_t b = code >= range;
        code = b ? code - range : code;
        norm();
        return b;
    }
    __device__ __forceinline__ uint32_t bypass_k16(int kk) {
        range >>= kk;
        uint32_t q = code / range, m = (1u << kk) - 1u;
        if (q > m) q = m;
        code -= q * range;
        norm();
        return q;
    }
    __device__ __forceinline__ uint32_t bypass_n(int kk) {
        uint32_t v = 0;
        while (kk > 16) { v = (v << 16) | bypass_k16(16); kk -= 16; }
        if (kk == 0) return v;
        return (v << kk) | bypass_k16(kk);
    }
    __device__ __forceinline__ uint32_t expgolomb() {
        uint32_t n1 = 0;
        while (n1 < 64 && bypass()) n1++;
        if (!n1) return 0;
        return ((1u << (n1 & 31)) | bypass_n((int)n1)) - 1u;
    }
};

// List sink of the lane decoder: every loop iteration stores its candidate
// entry at w[n] and advances n only when the entry is kept, so the serial
// loop has no branch and no LDS staging (the decode loop itself issues no
// global loads, so nothing ever waits on these stores).  Stores clip at
// the capacity; overflow is reported through count().
struct LaneSink {
    uint32_t *w;
    uint32_t cap, n;
    __device__ __forceinline__ void init(uint32_t *dst, uint32_t capacity) { w = dst; cap = capacity; n = 0; }
    __device__ __forceinline__ void emit(uint32_t v, bool keep) {
        w[n < cap ? n : cap - 1u] = v;
        n += keep ? 1u : 0u;
    }
    __device__ __forceinline__ uint32_t count() const { return n; }
};

// Coefficients of one block; ctx[s * kLaneStride] holds slot s: 0-3 the
// significance contexts of the current band, 4-15 the magnitude rungs, 16 a
// dummy written by bypass bins.
template <int W>
__device__ __forceinline__ uint32_t decode_coefs_lane(LaneCoder<W> &d, const QTab *__restrict__ qt,
                                                      const uint16_t *band_start, LaneSink &sink, uint32_t *ctx) {
    uint32_t k0 = 0;
    for (uint32_t i = 0; i < 13; ++i) {
        if (!d.bit(qt->eob[i])) break;
        k0 = i + 1;
    }
    uint32_t eob = k0 == 0 ? 0u : (k0 == 1 ? 1u : ((1u << (k0 - 1)) | d.bypass_n((int)k0 - 1)));
    if (eob > (uint32_t)kNVox) eob = kNVox;
    if (eob == 0) return 0;
    for (int i = 0; i < 12; ++i) ctx[(4 + i) * kLaneStride] = qt->mag[i];
    for (int i = 0; i < 4; ++i) ctx[i * kLaneStride] = qt->sig[i];
    uint32_t band = 0, next_band = band_start[1];
    // Bin state st: 0 significance, 1..12 magnitude rung st-1 (context slot
    // st + 3), 13 Exp-Golomb prefix (nb ones so far), 14 sign.  The context
    // slot of a bin is dens for st = 0, else min(st + 3, 16) (16: the dummy
    // slot of bypass bins).
    constexpr uint32_t kEgp = 13, kSgn = 14;
    uint32_t st = 0, p = 0, hist = 0, nb = 0, mag = 0;
    uint32_t evp = next_band < eob - 1 ? next_band : eob - 1;   // next position needing the slow path
    if (eob == 1) { hist = 1; st = 1; }                          // position 0 is the last: implied
    uint32_t slot = st == 0 ? 0u : 4u;                           // context slot of the coming bin
    uint32_t p0 = ctx[slot * kLaneStride];                       // its value, read one bin ahead
    while (p < eob) {
        const bool byp = st >= kEgp;
        const uint32_t rb = d.range >> 1, rm = __umul24(d.range >> 12, p0);
        const uint32_t r0 = byp ? rb : rm;
        const bool b = d.code >= r0;
        d.code -= b ? r0 : 0u;
        d.range = (b && !byp) ? d.range - r0 : r0;
        // adapt this bin's context before the next read, so that read sees it:
        // p0 - p0/16 after a 1, p0 + (4096 - p0)/16 after a 0 (both floored),
        // and p0 + floor((4096 - p0)/16) = p0 - floor((p0 - 4081)/16)
        ctx[slot * kLaneStride] = p0 - (uint32_t)((int32_t)(p0 - (b ? 0u : 4081u)) >> 4);
        d.norm_fast();     // issues the window read for the next bin
        // state transition
        const bool isSig = st == 0, isEgp = st == kEgp, isSgn = st == kSgn;
        const uint32_t entry = p | ((b ? 0u - mag : mag) << 16);
        hist = isSig ? (hist << 1) | (b ? 1u : 0u) : hist;
        p += ((isSig && !b) || isSgn) ? 1u : 0u;
        mag = isSgn ? mag : (st < 12u ? st : 12u);
        nb = isEgp ? nb + (b ? 1u : 0u) : 0u;
        const uint32_t nx = st + 1u < kEgp ? st + 1u : kEgp;
        const uint32_t nst = isSig ? (b ? 1u : 0u) : (isSgn ? 0u : (b ? nx : kSgn));
        // the next bin's context, read now so its latency overlaps the rest
        uint32_t dens = (uint32_t)__popc(hist & 0xFFFFu);
        dens = dens < 3u ? dens : 3u;
        const uint32_t sl = nst + 3u < 16u ? nst + 3u : 16u;
        slot = nst == 0 ? dens : sl;
        p0 = ctx[slot * kLaneStride];
        st = nst;
        MCHIP_SCHED_FENCE();
        sink.emit(entry, isSgn);
        // One test for every rare event: window slide, an Exp-Golomb prefix
        // bin (the suffix may follow), a band change or the last (implied)
        // position.
        const bool evt = nst == 0 && p >= evp;
        if (d.need_slide() || isEgp || evt) {
            if (d.need_slide()) d.slide();
            if (isEgp && (b ? nb == 64u : nb != 0u)) {           // Exp-Golomb suffix
                mag = 12u + ((1u << (nb & 31)) | d.bypass_n((int)nb)) - 1u;
                st = kSgn;
            }
            if (evt) {
                if (p >= next_band && band < 7) {
                    band++;
                    next_band = band_start[band + 1];
                    for (int i = 0; i < 4; ++i) ctx[i * kLaneStride] = qt->sig[band * 4 + i];
                }
                if (p + 1 == eob) { hist = (hist << 1) | 1u; st = 1; }
                evp = next_band < eob - 1 ? next_band : eob - 1;
            }
            uint32_t dn = (uint32_t)__popc(hist & 0xFFFFu);
            dn = dn < 3u ? dn : 3u;
            slot = st == 0 ? dn : (st + 3u < 16u ? st + 3u : 16u);
            // a distinct (volatile) read keeps the compiler from merging it
            // with the hot path's read and sinking both to the loop top; the
            // contexts are in LDS (every caller passes its __shared__ array),
            // and naming that address space makes the read a ds_read, where a
            // volatile generic pointer became a flat load waiting on vmcnt
            p0 = *(volatile __attribute__((address_space(3))) uint32_t *)(ctx + slot * kLaneStride);
        }
    }
    return sink.count();
}

// Lane decoder entry: header, coefficients and corrections into the list at
// w (cap words).  Returns the block record like entropy_to_list.  Coefs is the
// coefficient decoder (decode_coefs_lane, or the assembly loop of
// mchip_lane_asm.hpp on the GPU); both read the same bins in the same order.
// W is the window size in words (win holds W + 1 words per lane).
template <int W>
using LaneCoefFn = uint32_t (*)(LaneCoder<W> &, const QTab *__restrict__, const uint16_t *, LaneSink &, uint32_t *);
template <int W, LaneCoefFn<W> Coefs>
__device__ __forceinline__ uint2 entropy_to_list_lane_with(const uint8_t *payload, uint32_t len,
                                                           const QTab *__restrict__ qt, const uint16_t *band_start,
                                                           uint32_t *w, uint32_t cap, uint32_t *ctx, uint32_t *win) {
    LaneCoder<W> d;
    d.init(payload, len, win);
    uint32_t f = 0, dcv = 0;
    f |= d.bit(qt->flag[0]) ? 1u : 0u;
    f |= d.bit(qt->flag[1]) ? 2u : 0u;
    for (int b = 0; b < 8; ++b) dcv = (dcv << 1) | d.bit(qt->dc[b]);
    LaneSink sink;
    sink.init(w, cap);
    const uint32_t nsig = Coefs(d, qt, band_start, sink, ctx);
    uint32_t nc = 0;
    if (f & 2) {
        uint32_t n = d.expgolomb() + 1, pos = 0;
        if (n > (uint32_t)kNVox) n = kNVox;
        for (uint32_t i = 0; i < n; ++i) {
            pos += d.expgolomb();
            const uint32_t neg = d.bypass();
            const uint32_t m = d.expgolomb() + 1;
            if (pos >= (uint32_t)kNVox) break;
            sink.emit(pos, true);
            sink.emit(neg ? 0u - m : m, true);
            nc++;
        }
    }
    const uint32_t total = sink.count();
    const uint32_t status = total > cap ? kStatusFallback : kStatusOk;
    return make_uint2(dcv | (f << 8) | (status << 16), nsig | (nc << 16));
}

template <int W = kWin>
__device__ __forceinline__ uint2 entropy_to_list_lane(const uint8_t *payload, uint32_t len, const QTab *__restrict__ qt,
                                                      const uint16_t *band_start, uint32_t *w, uint32_t cap,
                                                      uint32_t *ctx, uint32_t *win) {
    return entropy_to_list_lane_with<W, decode_coefs_lane<W>>(payload, len, qt, band_start, w, cap, ctx, win);
}

// Entropy-decode one payload into a (raster index, level) list at w (cap
// words), corrections appended as (pos, delta) word pairs.  Returns the
// block record: x = dc | flags << 8 | status << 16, y = nsig | ncorr << 16.
__device__ __forceinline__ uint2 entropy_to_list(const uint8_t *payload, uint32_t len, const QTab *__restrict__ qt,
                                                 const uint16_t *band_start, uint32_t *w, uint32_t cap,
                                                 uint32_t *stage, uint32_t stage_stride) {
    RcDec d;
    d.init(payload, 0, len);
    int dc, flags;
    decode_header(d, qt, dc, flags);
    ListSink sink;
    sink.init(w, cap, stage, stage_stride);
    const uint32_t nsig = decode_coefs(d, qt, band_start, sink);
    uint32_t nc = 0;
    if (flags & 2) {
        uint32_t n = d.expgolomb() + 1, pos = 0;
        if (n > (uint32_t)kNVox) n = kNVox;
        for (uint32_t i = 0; i < n; ++i) {
            pos += d.expgolomb();
            int neg = d.bypass();
            uint32_t m = d.expgolomb() + 1;
            if (pos >= (uint32_t)kNVox) break;
            sink.put(pos);
            sink.put(neg ? 0u - m : m);
            nc++;
        }
    }
    sink.flush();
    uint32_t status = sink.n > sink.cap ? kStatusFallback : kStatusOk;
    return make_uint2((uint32_t)dc | ((uint32_t)flags << 8) | (status << 16), nsig | (nc << 16));
}

__device__ __forceinline__ void encode_magnitude(RcEnc &e, uint32_t (&m)[12], uint32_t mag) {
    uint32_t rest = mag - 1;
    if (rest == 0) { e.bit(m[0], 0); return; }
    bool done = false;
    #pragma unroll
    for (int k = 0; k < 11; ++k) {
        if (!done) {
            e.bit(m[k], 1);
            rest--;
            if (rest == 0) { e.bit(m[k + 1], 0); done = true; }
        }
    }
    if (!done) { e.bit(m[11], 1); e.expgolomb(rest); }
}

// Sequential reader of a 16-byte aligned u32 array that keeps the next
// 16 bytes in flight (the coding chain never waits on a load it just issued).
struct Prefetch4 {
    const uint4 *p;
    uint4 cur, nxt;
    uint32_t grp, ngrp;
    __device__ __forceinline__ void init(const uint32_t *base, uint32_t nwords) {
        p = (const uint4 *)base;
        ngrp = (nwords + 3) / 4;
        grp = 0;
        cur = ngrp > 0 ? p[0] : make_uint4(0, 0, 0, 0);
        nxt = ngrp > 1 ? p[1] : make_uint4(0, 0, 0, 0);
    }
    // word i; each call's i is the previous call's i or the one after it.
    // The group only ever steps to nxt: a "nxt or p[g]" choice lets the
    // compiler load through a selected pointer, which puts the whole reader
    // in scratch memory and a scratch round trip on every coded bin.
    __device__ __forceinline__ uint32_t get(uint32_t i) {
        if ((i >> 2) != grp) {
            cur = nxt;
            grp++;
            nxt = make_uint4(0, 0, 0, 0);
            if (grp + 1 < ngrp) nxt = p[grp + 1];
        }
        const uint32_t lo = (i & 1) ? cur.y : cur.x, hi = (i & 1) ? cur.w : cur.z;
        return (i & 2) ? hi : lo;
    }
};

// Stream of one block: header bins, EOB, coefficient bins over the
// scan-ordered levels L[0..eob), corrections C[0..ncorr) as (pos | delta<<16)
// (src/mc_codec.c:711-729); the caller flushes.
__device__ __forceinline__ void encode_block_stream(RcEnc &e, const QTab *__restrict__ qt, const uint16_t *band_start,
                                                    uint32_t dc, uint32_t eob, const int16_t *__restrict__ L,
                                                    uint32_t ncorr, const uint32_t *__restrict__ C) {
    e.bit_once(qt->flag[0], 0);
    e.bit_once(qt->flag[1], ncorr > 0);
    #pragma unroll
    for (int b = 7; b >= 0; --b) e.bit_once(qt->dc[7 - b], (dc >> b) & 1);
    // EOB: bit-length prefix (no terminator at 13), then the bits below the MSB
    {
        const int k = eob ? 32 - __clz(eob) : 0;
        #pragma unroll
        for (int i = 0; i < 13; ++i) {
            if (i < k) e.bit_once(qt->eob[i], 1);
            else if (i == k) e.bit_once(qt->eob[i], 0);
        }
        if (k > 1) e.bypass_n(eob & ((1u << (k - 1)) - 1u), k - 1);
    }
    uint32_t m[12];
    #pragma unroll
    for (int i = 0; i < 12; ++i) m[i] = qt->mag[i];
    uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int band = -1;
    uint32_t next_band = 0, hist = 0;
    Prefetch4 lr;
    lr.init((const uint32_t *)L, (eob + 1) / 2);
    for (uint32_t p = 0; p < eob; ++p) {
        if (p >= next_band) {
            band++;
            s0 = qt->sig[band * 4 + 0]; s1 = qt->sig[band * 4 + 1];
            s2 = qt->sig[band * 4 + 2]; s3 = qt->sig[band * 4 + 3];
            next_band = band_start[band + 1];
        }
        const int v = (int)(int16_t)(lr.get(p >> 1) >> (16 * (p & 1)));
        const int nzb = v != 0;
        if (p + 1 != eob) {
            uint32_t dens = __popc(hist & 0xFFFFu);
            uint32_t p0 = dens == 0 ? s0 : dens == 1 ? s1 : dens == 2 ? s2 : s3;
            e.bit(p0, nzb);
            s0 = dens == 0 ? p0 : s0;
            s1 = dens == 1 ? p0 : s1;
            s2 = dens == 2 ? p0 : s2;
            s3 = dens >= 3 ? p0 : s3;
        }
        hist = (hist << 1) | (uint32_t)nzb;
        if (nzb) {
            encode_magnitude(e, m, (uint32_t)(v < 0 ? -v : v));
            e.bypass(v < 0);
        }
    }
    if (ncorr) {
        e.expgolomb(ncorr - 1);
        uint32_t prev = 0;
        Prefetch4 cr;
        cr.init(C, ncorr);
        for (uint32_t i = 0; i < ncorr; ++i) {
            const uint32_t w = cr.get(i);
            const uint32_t pos = w & 0xFFFFu;
            const int d = (int)(int16_t)(w >> 16);
            e.expgolomb(pos - prev);
            prev = pos;
            e.bypass(d < 0);
            e.expgolomb((uint32_t)(d < 0 ? -d : d) - 1u);
        }
    }
}

}  // namespace mchip
