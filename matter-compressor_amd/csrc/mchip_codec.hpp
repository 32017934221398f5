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

// Coefficient decode with one bin per loop iteration.  Each lane walks its
// own state machine (significance, magnitude ladder, Exp-Golomb prefix and
// suffix, sign), so a wave spends max(bins) iterations instead of the union
// of its lanes' per-coefficient work.  The 16 adaptive contexts of the
// block (4 significance contexts of the current band + 12 magnitude rungs)
// live at ctx[i * cstride] (LDS on the GPU).  Same bins, same order, same
// results as decode_coefs (src/mc_codec.c:260-375).
enum : uint32_t { kPhSig = 0, kPhMag = 1, kPhEgPrefix = 2, kPhEgSuffix = 3, kPhSign = 4, kPhDone = 5 };

template <class Sink>
__device__ __forceinline__ uint32_t decode_coefs_sm(RcDec &d, const QTab *__restrict__ qt, const uint16_t *band_start,
                                                    Sink &sink, uint32_t *ctx, uint32_t cstride) {
    int k0 = 0;
    #pragma unroll
    for (int i = 0; i < 13; ++i)
        if (k0 == i && d.bit_once(qt->eob[i])) k0 = i + 1;
    uint32_t eob = k0 == 0 ? 0u : (k0 == 1 ? 1u : ((1u << (k0 - 1)) | d.bypass_n(k0 - 1)));
    if (eob > (uint32_t)kNVox) eob = kNVox;
    for (int i = 0; i < 12; ++i) ctx[(4 + i) * cstride] = qt->mag[i];
    uint32_t band = 0, next_band = band_start[1];
    for (int i = 0; i < 4; ++i) ctx[i * cstride] = qt->sig[i];
    uint32_t ph = eob ? kPhSig : kPhDone;
    uint32_t p = 0, hist = 0, k = 0, v = 0, nb = 0, mag = 0, nsig = 0;
    while (ph != kPhDone) {
        if (ph == kPhSig) {
            if (p >= next_band) {
                band++;
                next_band = band_start[band + 1];
                for (int i = 0; i < 4; ++i) ctx[i * cstride] = qt->sig[band * 4 + i];
            }
            if (p + 1 == eob) { hist = (hist << 1) | 1u; ph = kPhMag; k = 0; v = 0; }   // implied significant
        }
        if (ph == kPhEgSuffix) {
            const uint32_t x = ((1u << (nb & 31)) | d.bypass_n((int)nb)) - 1u;
            mag = v + x + 1;
            ph = kPhSign;
            continue;
        }
        const bool adapt = ph <= kPhMag;
        uint32_t dens = (uint32_t)__popc(hist & 0xFFFFu);
        dens = dens < 3u ? dens : 3u;
        const uint32_t ci = ph == kPhSig ? dens : 4u + k;
        uint32_t p0 = adapt ? ctx[ci * cstride] : 0u;
        const uint32_t r0 = adapt ? __umul24(d.range >> 12, p0) : d.range >> 1;
        const uint32_t b = d.code >= r0;
        d.code = b ? d.code - r0 : d.code;
        d.range = (b && adapt) ? d.range - r0 : r0;
        d.norm();
        if (adapt) {
            p0 = b ? p0 - (p0 >> 4) : p0 + ((4096u - p0) >> 4);
            ctx[ci * cstride] = p0;
        }
        if (ph == kPhSig) {
            hist = (hist << 1) | b;
            if (b) { ph = kPhMag; k = 0; v = 0; }
            else if (++p == eob) ph = kPhDone;
        } else if (ph == kPhMag) {
            if (k < 11) {
                if (b) { v++; k++; }
                else { mag = v + 1; ph = kPhSign; }
            } else if (!b) {
                mag = v + 1; ph = kPhSign;
            } else {
                ph = kPhEgPrefix; nb = 0;
            }
        } else if (ph == kPhEgPrefix) {
            if (b) {
                if (++nb == 64) ph = kPhEgSuffix;     // corrupt-input bound, as decode_coefs
            } else if (nb == 0) {
                mag = v + 1; ph = kPhSign;
            } else {
                ph = kPhEgSuffix;
            }
        } else {                                      // sign
            sink.coef(p, (uint16_t)(b ? 0u - mag : mag));
            nsig++;
            ph = ++p == eob ? kPhDone : kPhSig;
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

// Entropy-decode one payload into a (raster index, level) list at w (cap
// words), corrections appended as (pos, delta) word pairs.  Returns the
// block record: x = dc | flags << 8 | status << 16, y = nsig | ncorr << 16.
template <bool kSM>
__device__ __forceinline__ uint2 entropy_to_list(const uint8_t *payload, uint32_t len, const QTab *__restrict__ qt,
                                                 const uint16_t *band_start, uint32_t *w, uint32_t cap,
                                                 uint32_t *stage, uint32_t stage_stride, uint32_t *ctx,
                                                 const uint32_t *win = nullptr, int32_t nwin = 0) {
    RcDec d;
    d.init(payload, 0, len, win, nwin, (int32_t)stage_stride);
    int dc, flags;
    decode_header(d, qt, dc, flags);
    ListSink sink;
    sink.init(w, cap, stage, stage_stride);
    const uint32_t nsig = kSM ? decode_coefs_sm(d, qt, band_start, sink, ctx, stage_stride)
                              : decode_coefs(d, qt, band_start, sink);
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
    // word i; calls must come in non-decreasing i
    __device__ __forceinline__ uint32_t get(uint32_t i) {
        const uint32_t g = i >> 2;
        if (g != grp) {
            cur = g == grp + 1 ? nxt : p[g];
            grp = g;
            nxt = g + 1 < ngrp ? p[g + 1] : make_uint4(0, 0, 0, 0);
        }
        const uint32_t k = i & 3;
        return k == 0 ? cur.x : k == 1 ? cur.y : k == 2 ? cur.z : cur.w;
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
