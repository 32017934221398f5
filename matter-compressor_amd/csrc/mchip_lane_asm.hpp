// mchip_lane_asm.hpp -- the lane decoder with its coefficient loop written in
// CDNA4 (gfx950) assembly.  Device code only; included by mchip_decode.hip.
//
// Same bins, same order, same list as decode_coefs_lane (mchip_codec.hpp),
// which restates the reference's coefficient grammar (SuperOptimizer/
// matter-compressor src/mc_codec.c:321-375) with the range coder of
// src/mc_codec.c:178-237.
//
// Why assembly: a wave decodes 64 blocks in lockstep, one bin per loop
// iteration, so a chunk's decode time is its longest blocks' bin count times
// the wave's cost per bin.  On gfx950 that cost is set by the instruction
// count and form: VOP2/VOPC forms that use VCC issue in about 3 SIMD cycles,
// VOP3 forms with an SGPR mask or three operands in about 4.5, and a lone
// wave waits about 9 cycles on a dependent result.  The compiled loop spent
// about 95 instructions per bin; this one spends 67, with the context read
// of the next bin issued right after the decision so its LDS latency is
// covered by the renormalisation, the list store and the state update.
//
// Which bins run where:
//   - the loop: significance bins (state 0), magnitude-rung bins (states 1..12)
//     and sign bins (state 14, bypass);
//   - C++ (the handler after the loop): Exp-Golomb prefix and suffix of
//     magnitudes above 12 (state 13), window slides, band changes, the implied
//     last position and the end of the block.
// A lane that reaches one of those events leaves EXEC (it pauses) and the
// loop goes on with the others; the handler runs once some lane has waited
// kHoldBins bins, or when no lane is left running, and serves every paused
// lane in one pass.  One pass per event instead of one per lane-event is what
// makes the handler cheap: a wave meets about 1 600 lane-events per chunk.
//
// Context slots (LDS, kLaneStride lanes per slot): 0-3 significance (current
// band), 4-15 magnitude rungs 0-11, 17 the sign bin's (bypass, value unused).
#pragma once
#include "mchip_codec.hpp"

#ifndef MCHIP_HOLD_BINS
#define MCHIP_HOLD_BINS 16
#endif

namespace mchip {

constexpr uint32_t kHoldBins = MCHIP_HOLD_BINS;        // bins a paused lane waits at most
constexpr uint32_t kCtxSlots = 18;                      // context slots per lane
static_assert(kLaneStride == 64, "slot stride of 256 bytes");
static_assert(kHoldBins >= 1 && kHoldBins <= 64, "hold window");

// LDS byte offset of a pointer into a __shared__ array.
__device__ __forceinline__ uint32_t lds_off(const void *p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}

__device__ __forceinline__ uint32_t lds_read_volatile(uint32_t off) {
    return *(volatile __attribute__((address_space(3))) uint32_t *)(uintptr_t)off;
}

// Coefficients of one block into the list (as decode_coefs_lane); ctx holds
// kCtxSlots slots of this lane, d.win its payload window of W words.
template <int W>
__device__ __forceinline__ uint32_t decode_coefs_asm(LaneCoder<W> &d, const QTab *__restrict__ qt,
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
    uint32_t st = 0, p = 0, hist = 0, mag = 0;
    // evp: the position at which the loop must stop the lane (next band
    // start, the last position, whose significance is implied, or eob)
    uint32_t evp = next_band < eob - 1 ? next_band : eob - 1;
    if (eob == 1) { hist = 1; st = 1; evp = eob; }              // position 0 is the last: implied
    const uint32_t ctxb = lds_off(ctx), winb = lds_off(d.win);
    constexpr uint32_t kSlotBytes = 4u * kLaneStride;
    uint32_t cp = ctxb + (st == 0 ? 0u : 4u) * kSlotBytes;      // context address of the coming bin
    uint32_t p0 = lds_read_volatile(cp);                        // its value
    uint32_t code = d.code, range = d.range, rd = d.rd_bits, n = sink.n;
    uint64_t wp = ((uint64_t)d.w_hi << 32) | d.w_lo;            // window words rd/32 (high) and +1 (low)
    const uint64_t lst = (uint64_t)(uintptr_t)sink.w;
    const uint32_t capm1 = sink.cap - 1u;
    uint32_t done = 0;
    do {
        uint32_t t0, t1, t2, t3, t4, t5, t6, t7, t8, r0, bi, sh, age;
        uint64_t sg, ss, si, sj, sr, la, held, sx;
        // Registers: v101 holds code during the loop, v100 the next 32
        // payload bits beside it (a 64-bit shift renormalises code);
        // v[102:103] the window pair shifted to the read position; v104 the
        // list index beside v105 = 0 (64-bit address arithmetic).
        asm volatile(
            "v_mov_b32 v101, %[code]\n"
            "v_mov_b32 v105, 0\n"
            "s_mov_b64 %[held], 0\n"
            "s_mov_b32 %[age], %[hold]\n"
            "L_bin_%=:\n"
            "s_waitcnt lgkmcnt(0)\n"
            // decision: r0 = sign bin ? range >> 1 : (range >> 12) * p0; b = code >= r0
            "v_lshrrev_b32 %[t0], 12, %[range]\n"
            "v_cmp_eq_u32_e64 %[sg], 14, %[st]\n"
            "v_mul_u32_u24 %[t0], %[t0], %[p0]\n"
            "v_lshrrev_b32 %[t1], 1, %[range]\n"
            "v_add_u32 %[t4], 0xfffff00f, %[p0]\n"
            "v_and_b32 %[t5], 24, %[rd]\n"
            "v_cndmask_b32_e64 %[r0], %[t0], %[t1], %[sg]\n"
            "v_cmp_eq_u32_e64 %[ss], 0, %[st]\n"
            "v_sub_u32 %[t3], v101, %[r0]\n"
            "v_cmp_ge_u32 vcc, v101, %[r0]\n"
            "v_sub_u32 %[t2], %[range], %[r0]\n"
            "v_add_u32 %[t7], 1, %[st]\n"
            "v_lshlrev_b64 v[102:103], %[t5], %[wp]\n"
            // context adaptation (p0 - p0/16 after a 1, p0 + (4096 - p0)/16
            // after a 0, both floored), written before the next read
            "v_cndmask_b32 %[t4], %[t4], %[p0], vcc\n"
            "v_cndmask_b32_e64 %[bi], 0, 1, vcc\n"
            "v_ashrrev_i32 %[t4], 4, %[t4]\n"
            "v_sub_u32 %[t4], %[p0], %[t4]\n"
            "ds_write_b32 %[cp], %[t4]\n"
            // next state: significance 0 -> 0 (p+1) or rung 1; rung k -> k+1
            // or sign; sign -> significance (p+1); its context read goes now
            "v_lshl_or_b32 %[t6], %[hist], 1, %[bi]\n"
            "s_andn2_b64 %[si], %[ss], vcc\n"
            "v_cndmask_b32 %[t7], 14, %[t7], vcc\n"
            "s_or_b64 %[si], %[si], %[sg]\n"
            "v_cndmask_b32_e64 %[hist], %[hist], %[t6], %[ss]\n"
            "v_cndmask_b32_e64 %[t7], %[t7], 0, %[si]\n"
            "v_and_b32 %[t8], 0xffff, %[hist]\n"
            "v_bcnt_u32_b32 %[t8], %[t8], 0\n"
            "v_cmp_eq_u32_e64 %[sr], 0, %[t7]\n"
            "v_min_u32 %[t8], 3, %[t8]\n"
            "v_add_u32 %[t1], 3, %[t7]\n"
            "v_cndmask_b32_e64 %[t1], %[t1], %[t8], %[sr]\n"
            "v_lshl_add_u32 %[cp], %[t1], 8, %[ctxb]\n"
            "ds_read_b32 %[p0], %[cp]\n"
            // code and range after the decision, renormalised by whole bytes
            "v_cndmask_b32_e64 %[t2], %[t2], %[r0], %[sg]\n"
            "v_mov_b32 v100, v103\n"
            "v_cndmask_b32 v101, v101, %[t3], vcc\n"
            "v_cndmask_b32 %[range], %[r0], %[t2], vcc\n"
            "v_ffbh_u32 %[sh], %[range]\n"
            "v_and_b32 %[sh], 24, %[sh]\n"
            "v_add_u32 %[rd], %[rd], %[sh]\n"
            "v_and_b32 %[t5], -32, %[rd]\n"
            "v_lshl_add_u32 %[t5], %[t5], 3, %[winb]\n"
            "ds_read2st64_b32 %[wp], %[t5] offset0:1 offset1:0\n"
            "v_lshlrev_b64 v[100:101], %[sh], v[100:101]\n"
            "v_lshlrev_b32 %[range], %[sh], %[range]\n"
            // list entry p | level << 16 (kept at sign bins), magnitude, position
            "v_sub_u32 %[t6], 0, %[mag]\n"
            "v_cndmask_b32 %[t6], %[mag], %[t6], vcc\n"
            "v_lshl_or_b32 %[t6], %[t6], 16, %[p]\n"
            "v_min_u32 v104, %[capm1], %[n]\n"
            "v_lshl_add_u64 %[la], v[104:105], 2, %[lst]\n"
            // only the sign bins' lanes store: a 64-lane store of scattered
            // words on every bin occupied the memory pipeline once several
            // waves share a CU (batched decode 140 -> 206 GB/s)
            "s_and_saveexec_b64 %[sx], %[sg]\n"
            "global_store_dword %[la], %[t6], off\n"
            "s_mov_b64 exec, %[sx]\n"
            "v_cndmask_b32_e64 %[mag], %[st], %[mag], %[sg]\n"
            "v_addc_co_u32_e64 %[n], %[sj], 0, %[n], %[sg]\n"
            "v_addc_co_u32_e64 %[p], %[sj], 0, %[p], %[si]\n"
            "v_mov_b32 %[st], %[t7]\n"
            // events: window slide, Exp-Golomb prefix next, event position
            "v_cmp_le_u32_e64 %[sr], %[evp], %[p]\n"
            "v_cmp_eq_u32_e64 %[sj], 13, %[st]\n"
            "v_cmp_lt_u32 vcc, %[thr], %[rd]\n"
            "s_or_b64 vcc, vcc, %[sr]\n"
            "s_or_b64 vcc, vcc, %[sj]\n"
            // lanes with an event leave EXEC; the loop stops once a lane has
            // waited kHoldBins bins (age borrows) or no lane is left running
            "s_or_b64 %[held], %[held], vcc\n"
            "s_subb_u32 %[age], %[age], 0\n"
            "s_cbranch_scc1 L_out_%=\n"
            "s_andn2_b64 exec, exec, vcc\n"
            "s_cbranch_scc1 L_bin_%=\n"
            "L_out_%=:\n"
            "s_or_b64 exec, exec, %[held]\n"
            "s_waitcnt lgkmcnt(0)\n"
            "v_mov_b32 %[code], v101\n"
            : [code] "+v"(code), [range] "+v"(range), [p0] "+v"(p0), [cp] "+v"(cp), [st] "+v"(st),
              [hist] "+v"(hist), [p] "+v"(p), [mag] "+v"(mag), [rd] "+v"(rd), [wp] "+v"(wp), [n] "+v"(n),
              [t0] "=&v"(t0), [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3), [t4] "=&v"(t4), [t5] "=&v"(t5),
              [t6] "=&v"(t6), [t7] "=&v"(t7), [t8] "=&v"(t8), [r0] "=&v"(r0), [bi] "=&v"(bi), [sh] "=&v"(sh),
              [la] "=&v"(la), [sg] "=&s"(sg), [ss] "=&s"(ss), [si] "=&s"(si), [sj] "=&s"(sj), [sr] "=&s"(sr),
              [held] "=&s"(held), [age] "=&s"(age), [sx] "=&s"(sx)
            : [ctxb] "v"(ctxb), [winb] "v"(winb), [evp] "v"(evp), [capm1] "v"(capm1), [lst] "v"(lst),
              [hold] "n"(kHoldBins - 1), [thr] "n"(32 * (W - 1) - 1)
            : "vcc", "v100", "v101", "v102", "v103", "v104", "v105", "memory");
        // the paused lanes' events, in the order the bins need them
        d.code = code; d.range = range; d.rd_bits = rd;
        d.w_lo = (uint32_t)wp; d.w_hi = (uint32_t)(wp >> 32);
        if (d.need_slide() || st == 13u || p >= evp) {
            if (d.need_slide()) d.slide();
            if (st == 13u) { mag = 12u + d.expgolomb(); st = 14u; }    // magnitude above 12: EG0 tail
            if (p >= evp) {
                if (p >= eob) {
                    done = 1;
                } else {
                    if (p >= next_band && band < 7) {
                        band++;
                        next_band = band_start[band + 1];
                        for (int i = 0; i < 4; ++i) ctx[i * kLaneStride] = qt->sig[band * 4 + i];
                    }
                    evp = next_band < eob - 1 ? next_band : eob - 1;
                    if (p + 1 == eob) { hist = (hist << 1) | 1u; st = 1; evp = eob; }
                }
            }
            code = d.code; range = d.range; rd = d.rd_bits;
            wp = ((uint64_t)d.w_hi << 32) | d.w_lo;
            uint32_t dn = (uint32_t)__popc(hist & 0xFFFFu);
            dn = dn < 3u ? dn : 3u;
            cp = ctxb + (st == 0 ? dn : st + 3u) * kSlotBytes;
            p0 = lds_read_volatile(cp);
        }
    } while (!done);
    sink.n = n;
    return n;
}

// Header, coefficients and corrections of one block into the list at w (cap
// words), as entropy_to_list_lane; ctx has kCtxSlots slots per lane, win a
// window of W words (W + 1 allocated) per lane.
template <int W = kWin>
__device__ __forceinline__ uint2 entropy_to_list_asm(const uint8_t *payload, uint32_t len, const QTab *__restrict__ qt,
                                                     const uint16_t *band_start, uint32_t *w, uint32_t cap,
                                                     uint32_t *ctx, uint32_t *win) {
    return entropy_to_list_lane_with<W, decode_coefs_asm<W>>(payload, len, qt, band_start, w, cap, ctx, win);
}

}  // namespace mchip
