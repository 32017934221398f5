// CPU harness of the device coder logic: the same mchip_codec.hpp the HIP
// kernels use, compiled for the host through the shim in hip/hip_runtime.h.
// Test infrastructure only.
#define MCHIP_LANE_STRIDE 1
#include "mchip_codec.hpp"

using namespace mchip;

extern "C" {

// Entropy-decode a payload with the device decoder; expands the list into
// raster levels (via scan) and returns corrections.  Returns the status.
// Variant 2: the lane decoder (one bin per step); 0: one step per position.
static int variant = 2;
void dl_set_variant(int v) { variant = v; }

int dl_decode(const QTab *qt, const uint16_t *band_start, const uint16_t *scan, const uint8_t *payload, uint32_t len,
              int16_t *levels, int *dc, int *flags, uint32_t *ncorr, uint32_t *cpos, int32_t *cdel) {
    static thread_local uint32_t w[8 * 17408 + 64 + 3 * 4096 * 2];
    uint32_t cap = 8u * len + 32u;
    uint32_t stage[kStage], ctx[17];
    static thread_local uint32_t lwin[kWinAlloc];
    uint2 r = variant == 2 ? entropy_to_list_lane(payload, len, qt, band_start, w, cap, ctx, lwin)
                           : entropy_to_list(payload, len, qt, band_start, w, cap, stage, 1);
    memset(levels, 0, 4096 * sizeof(int16_t));
    *dc = r.x & 255;
    *flags = (r.x >> 8) & 255;
    uint32_t nw = r.y & 0xFFFF, nc = r.y >> 16;
    int status = (int)(r.x >> 16);
    if (status != 0) { *ncorr = 0; return status; }
    for (uint32_t i = 0; i < nw; ++i) levels[scan[w[i] & 0xFFFF]] = (int16_t)(w[i] >> 16);
    for (uint32_t i = 0; i < nc; ++i) { cpos[i] = w[nw + 2 * i]; cdel[i] = (int32_t)w[nw + 2 * i + 1]; }
    *ncorr = nc;
    return 0;
}

// Full single-lane decode into a per-block level array (no capacity limit).
struct ArraySink {
    int16_t *lv;
    const uint16_t *scan;
    void coef(uint32_t p, uint16_t v) { lv[scan[p]] = (int16_t)v; }
};
uint32_t dl_decode_levels_full(const QTab *qt, const uint16_t *band_start, const uint16_t *scan,
                               const uint8_t *payload, uint32_t len, int16_t *levels, int *dc, int *flags) {
    RcDec d;
    d.init(payload, 0, len);
    decode_header(d, qt, *dc, *flags);
    memset(levels, 0, 4096 * sizeof(int16_t));
    ArraySink s{levels, scan};
    return decode_coefs(d, qt, band_start, s);
}

// Encode a block stream from scan-ordered levels and corrections.
uint32_t dl_encode(const QTab *qt, const uint16_t *band_start, uint32_t dc, uint32_t eob, const int16_t *L,
                   uint32_t ncorr, const uint32_t *C, uint8_t *out, uint32_t cap) {
    RcEnc e;
    uint32_t stage[kEncStage];
    e.init((uint32_t *)out, cap, stage, 1);
    encode_block_stream(e, qt, band_start, dc, eob, L, ncorr, C);
    e.finish();
    return e.n;
}

// A block payload with no coefficients (eob 0) and the given corrections:
// gap[i] (position step, 0 repeats the position), neg[i], mag[i] >= 1, in the
// bin order of the reference's encoder.  Lets tests build correction lists
// the encoder itself never writes (repeated positions).
uint32_t dl_encode_corrections(const QTab *qt, uint32_t dc, uint32_t n, const uint32_t *gap, const int *neg,
                               const uint32_t *mag, uint8_t *out, uint32_t cap) {
    RcEnc e;
    uint32_t stage[kEncStage];
    e.init((uint32_t *)out, cap, stage, 1);
    e.bit_once(qt->flag[0], 0);
    e.bit_once(qt->flag[1], 1);
    for (int b = 7; b >= 0; --b) e.bit_once(qt->dc[7 - b], (dc >> b) & 1);
    e.bit_once(qt->eob[0], 0);
    e.expgolomb(n - 1);
    for (uint32_t i = 0; i < n; ++i) {
        e.expgolomb(gap[i]);
        e.bypass(neg[i]);
        e.expgolomb(mag[i] - 1);
    }
    e.finish();
    return e.n;
}

// IDCT of one 16-point line with the device kernel's arithmetic.
void dl_idct_line(float *v, const float *cmf) { idct16_line(v, cmf); }
// The wave-sparse inverse with the caller's input masks (bit 0 of wk[k] set:
// the wave runs input k); returns line_masks' verdict on the line.
int dl_idct_line_sparse(float *v, const float *cmf, const uint64_t *wk) {
    uint64_t own[16], m[16];
    const bool nz = line_masks(*(const float(*)[16])v, own);
    for (int k = 0; k < 16; ++k) m[k] = wk[k];
    if (nz) idct16_line_sparse(v, cmf, m);
    return nz;
}
void dl_fdct_line(float *v, const float *cmf) { fdct16_line(v, cmf); }
}
