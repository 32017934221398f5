/*
 * mc_oracle.c -- CPU restatement of the matter-compressor 16^3-block codec.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker + CPU baseline).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it; the
 * product library never links it.  See mc_oracle.h for what pins each part.
 *
 * Build: -O2 -std=c11 -fno-fast-math -ffp-contract=off (no FMA contraction,
 * IEEE single precision in source order), matching the reference's
 * deterministic build (CMakeLists.txt:36-37, src/mc_codec_float.h:34-35).
 *
 * Citations are to SuperOptimizer/matter-compressor.
 */
#define _GNU_SOURCE
#include "mc_oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define NVOX 4096
#define BLK 16

/* ------------------------------------------------------------------------
 * Trained prior endpoint tables (q=1 and q=12), P(bit==0) in 1/4096 units.
 * Same values as src/mc_codec.c:121-140 (format constants: the decoder must
 * start every context from exactly these numbers).  Class order:
 * sig, mag, eob, mask, mask-u, mask-a, flag, dc (src/mc_codec.c:108).
 * Slots not listed are 2048.
 * ------------------------------------------------------------------------ */
static const uint16_t k_lo_head[8][24] = {
    {115,110,112,144,3866,3522,3330,948,3939,3723,3483,2383,4007,3795,3575,3130,4042,3912,3805,3686,4064,4064,4028,4064},
    {1996,1234,948,743,595,539,441,421,333,358,297,305},
    {32,32,32,32,32,32,32,32,32,92,285,2295,4064},
    {4025,3064,2995,827,2107,473,576,50},
    {250,1713,2961,2991},
    {3936,32},
    {1023,4064,4064,4064},
    {3873,366,1820,2056,2040,2055,2041,2045},
};
static const uint16_t k_hi_head[8][24] = {
    {456,335,535,1133,3950,3755,3590,2892,4064,4051,3990,3686},
    {2121,1410,1081,863,746,631,569,508,465,431,399,370},
    {59,32,32,32,32,63,92,315,2273,4064,4064},
    {4025,3064,2995,827,2107,473,576,50},
    {250,1713,2961,2991},
    {3936,32},
    {1023,4064,4064,4064},
    {3873,366,1820,2056,2040,2055,2041,2045},
};
static const int k_lo_n[8] = {24,12,13,8,4,2,4,8};
static const int k_hi_n[8] = {12,12,11,8,4,2,4,8};

static uint16_t g_plo[8][32], g_phi[8][32];
static int g_tables_ready = 0;
static pthread_once_t g_once = PTHREAD_ONCE_INIT;
static pthread_mutex_t g_pri_mu = PTHREAD_MUTEX_INITIALIZER;

static float    g_cmf[16][16];
static uint16_t g_scan[NVOX];
static uint8_t  g_band_of_raster[NVOX];

static int scan_key_cmp(const void *a, const void *b) {
    /* ascending L1 frequency, ties by raster index (src/mc_codec.c:283-287) */
    uint32_t ia = *(const uint32_t *)a, ib = *(const uint32_t *)b;
    uint32_t la = (ia >> 8) + ((ia >> 4) & 15) + (ia & 15);
    uint32_t lb = (ib >> 8) + ((ib >> 4) & 15) + (ib & 15);
    if (la != lb) return la < lb ? -1 : 1;
    return ia < ib ? -1 : (ia > ib);
}

static void default_priors(uint16_t lo[8][32], uint16_t hi[8][32]) {
    for (int c = 0; c < 8; ++c)
        for (int s = 0; s < 32; ++s) {
            lo[c][s] = s < k_lo_n[c] ? k_lo_head[c][s] : 2048;
            hi[c][s] = s < k_hi_n[c] ? k_hi_head[c][s] : 2048;
        }
}

static void init_tables(void) {
    default_priors(g_plo, g_phi);
    /* orthonormal DCT-II matrix in double, rounded to float
     * (src/mc_codec_float.h:70-78) */
    for (int k = 0; k < 16; ++k) {
        double ck = (k == 0) ? sqrt(1.0 / 16) : sqrt(2.0 / 16);
        for (int n = 0; n < 16; ++n)
            g_cmf[k][n] = (float)(ck * cos(3.14159265358979323846 * (2.0 * n + 1.0) * k / 32.0));
    }
    uint32_t ord[NVOX];
    for (uint32_t i = 0; i < NVOX; ++i) ord[i] = i;
    qsort(ord, NVOX, sizeof ord[0], scan_key_cmp);
    for (int i = 0; i < NVOX; ++i) g_scan[i] = (uint16_t)ord[i];
    /* band = min(7, L1*8/48) (src/mc_codec.c:305-308) */
    for (int i = 0; i < NVOX; ++i) {
        int l1 = (i >> 8) + ((i >> 4) & 15) + (i & 15);
        int b = l1 * 8 / 48;
        g_band_of_raster[i] = (uint8_t)(b > 7 ? 7 : b);
    }
    g_tables_ready = 1;
}
static void ensure_tables(void) { pthread_once(&g_once, init_tables); }

void orc_set_priors(const uint16_t *plo, const uint16_t *phi) {
    ensure_tables();
    pthread_mutex_lock(&g_pri_mu);
    if (plo && phi) {
        memcpy(g_plo, plo, sizeof g_plo);
        memcpy(g_phi, phi, sizeof g_phi);
    } else {
        default_priors(g_plo, g_phi);
    }
    pthread_mutex_unlock(&g_pri_mu);
}

float orc_sanitize_q(float q) {
    /* src/mc_codec.c:493-497 */
    if (!(q > 0.0f) || q > 1024.0f) q = 8.0f;
    return q;
}

void orc_priors_for_q(float q, uint16_t pri[8][32]) {
    /* log2(q) interpolation between the q=1 and q=12 tables
     * (src/mc_codec.c:156-167) */
    ensure_tables();
    const float lo = 0.0f, hi = 3.585f;
    float w;
    if (q <= 1.0f) w = 0.0f;
    else w = ((float)(log((double)q) / log(2.0)) - lo) / (hi - lo);
    if (w < 0.0f) w = 0.0f;
    if (w > 1.0f) w = 1.0f;
    for (int c = 0; c < 8; ++c)
        for (int s = 0; s < 32; ++s) {
            int a = g_plo[c][s], b = g_phi[c][s];
            float v = (float)a + (float)(b - a) * w;
            v = v + 0.5f;
            pri[c][s] = (uint16_t)v;
        }
}

void orc_step_table(float q, float step[NVOX], float rstep[NVOX]) {
    /* step = q * (1 + L1)^0.65, rstep = 1/step (src/mc_codec.c:541-552) */
    float wl[46];
    for (int l = 0; l < 46; ++l) wl[l] = powf(1.0f + (float)l, 0.65f);
    for (int i = 0; i < NVOX; ++i) {
        int l1 = (i >> 8) + ((i >> 4) & 15) + (i & 15);
        step[i] = q * wl[l1];
        rstep[i] = 1.0f / step[i];
    }
}

void orc_dct_matrix(float cmf[256]) {
    ensure_tables();
    memcpy(cmf, g_cmf, sizeof g_cmf);
}
void orc_scan_order(uint16_t scan[NVOX]) {
    ensure_tables();
    memcpy(scan, g_scan, sizeof g_scan);
}

/* ------------------------------------------------------------------------
 * Separable 16^3 DCT.  The reference transforms contiguous lines and rotates
 * the cube between passes (src/mc_codec_float.h:131-158); that is an x pass,
 * then a y pass, then a z pass over the original (z,y,x) layout, which is
 * what the strided passes below do.  Each line uses the even/odd butterfly
 * with 8-term sums accumulated in order from 0.0f
 * (src/mc_codec_float.h:83-105); an all-zero line maps to zeros (:109-126).
 * ------------------------------------------------------------------------ */
static void line_fwd(float *v, int stride) {
    float in[16], s[8], d[8];
    int nz = 0;
    for (int i = 0; i < 16; ++i) { in[i] = v[i * stride]; nz |= (in[i] != 0.0f); }
    if (!nz) { for (int i = 0; i < 16; ++i) v[i * stride] = 0.0f; return; }
    for (int n = 0; n < 8; ++n) { s[n] = in[n] + in[15 - n]; d[n] = in[n] - in[15 - n]; }
    for (int k = 0; k < 16; ++k) {
        const float *src = (k & 1) ? d : s;
        float acc = 0.0f;
        for (int n = 0; n < 8; ++n) acc += g_cmf[k][n] * src[n];
        v[k * stride] = acc;
    }
}
static void line_inv(float *v, int stride) {
    float in[16], e[8], o[8];
    int nz = 0;
    for (int i = 0; i < 16; ++i) { in[i] = v[i * stride]; nz |= (in[i] != 0.0f); }
    if (!nz) { for (int i = 0; i < 16; ++i) v[i * stride] = 0.0f; return; }
    for (int n = 0; n < 8; ++n) { e[n] = 0.0f; o[n] = 0.0f; }
    for (int k = 0; k < 16; k += 2) {
        float c = in[k];
        if (c != 0.0f) for (int n = 0; n < 8; ++n) e[n] += g_cmf[k][n] * c;
    }
    for (int k = 1; k < 16; k += 2) {
        float c = in[k];
        if (c != 0.0f) for (int n = 0; n < 8; ++n) o[n] += g_cmf[k][n] * c;
    }
    for (int n = 0; n < 8; ++n) { v[n * stride] = e[n] + o[n]; v[(15 - n) * stride] = e[n] - o[n]; }
}
static void dct3(const float *in, float *out, void (*line)(float *, int)) {
    memcpy(out, in, NVOX * sizeof(float));
    for (int zy = 0; zy < 256; ++zy) line(out + zy * 16, 1);                 /* x */
    for (int z = 0; z < 16; ++z) for (int x = 0; x < 16; ++x) line(out + z * 256 + x, 16); /* y */
    for (int yx = 0; yx < 256; ++yx) line(out + yx, 256);                    /* z */
}
void orc_dct3_fwd(const float in[NVOX], float out[NVOX]) { ensure_tables(); dct3(in, out, line_fwd); }
void orc_dct3_inv(const float in[NVOX], float out[NVOX]) { ensure_tables(); dct3(in, out, line_inv); }

/* ------------------------------------------------------------------------
 * Binary range coder (src/mc_codec.c:95-238).  32-bit range, 12-bit P(0),
 * adaptation shift 4, LZMA-style carry through a cached byte plus a run of
 * pending 0xFF bytes; the decoder reads zeros past the end of the payload.
 * ------------------------------------------------------------------------ */
typedef struct {
    uint8_t *out; size_t cap, n;
    uint64_t low; uint32_t range;
    uint8_t cached; uint64_t pending;
} renc;
typedef struct { const uint8_t *in; size_t len, at; uint32_t code, range; } rdec;

static void re_begin(renc *e, uint8_t *out, size_t cap) {
    e->out = out; e->cap = cap; e->n = 0; e->low = 0; e->range = 0xFFFFFFFFu;
    e->cached = 0; e->pending = 1;
}
static void re_emit(renc *e, uint8_t b) { if (e->n < e->cap) e->out[e->n] = b; e->n++; }
static void re_shift(renc *e) {
    uint32_t top = (uint32_t)(e->low >> 32);
    if (top != 0 || e->low < 0xFF000000ull) {
        uint8_t carry = (uint8_t)top;
        re_emit(e, (uint8_t)(e->cached + carry));
        while (--e->pending) re_emit(e, (uint8_t)(0xFF + carry));
        e->cached = (uint8_t)(e->low >> 24);
    }
    e->pending++;
    e->low = (e->low << 8) & 0xFFFFFFFFull;
}
static void re_norm(renc *e) { while (e->range < (1u << 24)) { re_shift(e); e->range <<= 8; } }
static void re_bit(renc *e, uint16_t *p0, int bit) {
    uint32_t r0 = (e->range >> 12) * (uint32_t)*p0;
    if (!bit) { e->range = r0; *p0 = (uint16_t)(*p0 + ((4096 - *p0) >> 4)); }
    else { e->low += r0; e->range -= r0; *p0 = (uint16_t)(*p0 - (*p0 >> 4)); }
    re_norm(e);
}
static void re_raw(renc *e, int bit) { e->range >>= 1; if (bit) e->low += e->range; re_norm(e); }
static void re_raw_n(renc *e, uint32_t v, int k) {
    if (k > 16) { re_raw_n(e, (v >> (k - 16)) & 0xFFFF, 16); re_raw_n(e, v, k - 16); return; }
    if (k == 0) return;
    e->range >>= k;
    e->low += (uint64_t)(v & ((1u << k) - 1)) * e->range;
    re_norm(e);
}
static void re_finish(renc *e) { for (int i = 0; i < 5; ++i) re_shift(e); }
static int bitlen(uint32_t v) { int k = 0; while (k < 32 && (1u << k) <= v) ++k; return k; }
static void re_expgolomb(renc *e, uint32_t v) {
    /* order-0 Exp-Golomb in bypass bits (src/mc_codec.c:229-233) */
    int nb = bitlen(v + 1) - 1;
    for (int i = 0; i < nb; ++i) re_raw(e, 1);
    re_raw(e, 0);
    if (nb) re_raw_n(e, (v + 1) & ((1u << nb) - 1), nb);
}

static uint8_t rd_byte(rdec *d) { return d->at < d->len ? d->in[d->at++] : 0; }
static void rd_begin(rdec *d, const uint8_t *in, size_t len) {
    d->in = in; d->len = len; d->at = 0; d->code = 0; d->range = 0xFFFFFFFFu;
    for (int i = 0; i < 5; ++i) d->code = (d->code << 8) | rd_byte(d);
}
static void rd_norm(rdec *d) {
    while (d->range < (1u << 24)) { d->code = (d->code << 8) | rd_byte(d); d->range <<= 8; }
}
static int rd_bit(rdec *d, uint16_t *p0) {
    uint32_t r0 = (d->range >> 12) * (uint32_t)*p0;
    int bit;
    if (d->code < r0) { d->range = r0; bit = 0; *p0 = (uint16_t)(*p0 + ((4096 - *p0) >> 4)); }
    else { d->code -= r0; d->range -= r0; bit = 1; *p0 = (uint16_t)(*p0 - (*p0 >> 4)); }
    rd_norm(d);
    return bit;
}
static int rd_raw(rdec *d) {
    d->range >>= 1;
    int bit = d->code >= d->range;
    if (bit) d->code -= d->range;
    rd_norm(d);
    return bit;
}
static uint32_t rd_raw_n(rdec *d, int k) {
    uint32_t v = 0;
    while (k > 16) { v = (v << 16) | rd_raw_n(d, 16); k -= 16; }
    if (k == 0) return v;
    d->range >>= k;
    uint32_t q = d->code / d->range, m = (1u << k) - 1;
    if (q > m) q = m;
    d->code -= q * d->range;
    rd_norm(d);
    return (v << k) | q;
}
static uint32_t rd_expgolomb(rdec *d) {
    uint32_t nb = 0;
    while (rd_raw(d)) nb++;
    if (!nb) return 0;
    return ((1u << (nb & 31)) | rd_raw_n(d, (int)nb)) - 1;
}

size_t orc_rc_encode_ops(int n, const uint8_t *kinds, const uint8_t *ctx_idx,
                         const uint8_t *nbits, const uint32_t *vals, uint8_t *out, size_t cap) {
    renc e; re_begin(&e, out, cap);
    uint16_t ctx[64]; for (int i = 0; i < 64; ++i) ctx[i] = 2048;
    for (int i = 0; i < n; ++i) {
        if (kinds[i] == 0) re_bit(&e, &ctx[ctx_idx[i] & 63], (int)(vals[i] & 1));
        else if (kinds[i] == 1) re_raw(&e, (int)(vals[i] & 1));
        else re_raw_n(&e, vals[i], nbits[i]);
    }
    re_finish(&e);
    return e.n;
}
void orc_rc_decode_ops(int n, const uint8_t *kinds, const uint8_t *ctx_idx, const uint8_t *nbits,
                       const uint8_t *in, size_t len, uint32_t *vals_out) {
    rdec d; rd_begin(&d, in, len);
    uint16_t ctx[64]; for (int i = 0; i < 64; ++i) ctx[i] = 2048;
    for (int i = 0; i < n; ++i) {
        if (kinds[i] == 0) vals_out[i] = (uint32_t)rd_bit(&d, &ctx[ctx_idx[i] & 63]);
        else if (kinds[i] == 1) vals_out[i] = (uint32_t)rd_raw(&d);
        else vals_out[i] = rd_raw_n(&d, nbits[i]);
    }
}

/* ------------------------------------------------------------------------
 * Coefficient context coder (src/mc_codec.c:240-375).
 * ------------------------------------------------------------------------ */
typedef struct { uint16_t sig[32], mag[12], eob[14]; } coef_ctx;
static void coef_ctx_init(coef_ctx *c, const uint16_t pri[8][32]) {
    memcpy(c->sig, pri[0], sizeof c->sig);
    memcpy(c->mag, pri[1], sizeof c->mag);
    memcpy(c->eob, pri[2], sizeof c->eob);
}

static void put_magnitude(renc *e, coef_ctx *c, uint32_t m) {
    /* unary over mag[0..11] then an EG0 tail (src/mc_codec.c:260-267) */
    uint32_t rest = m - 1;
    int k = 0;
    while (k < 11 && rest > 0) {
        re_bit(e, &c->mag[k], 1); rest--; k++;
        if (rest == 0) { re_bit(e, &c->mag[k], 0); return; }
    }
    if (rest == 0) { re_bit(e, &c->mag[k], 0); return; }
    re_bit(e, &c->mag[11], 1);
    re_expgolomb(e, rest);
}
static uint32_t get_magnitude(rdec *d, coef_ctx *c) {
    uint32_t v = 0;
    for (int k = 0; k < 11; ++k) {
        if (!rd_bit(d, &c->mag[k])) return v + 1;
        v++;
    }
    if (!rd_bit(d, &c->mag[11])) return v + 1;
    return v + rd_expgolomb(d) + 1;
}

static void put_eob(renc *e, coef_ctx *c, uint32_t v) {
    /* bit-length prefix over eob[0..k] (no terminator at k=13), then the
     * bits below the MSB in bypass (src/mc_codec.c:314-320) */
    const int kmax = 13;
    int k = bitlen(v);
    for (int i = 0; i < k; ++i) re_bit(e, &c->eob[i], 1);
    if (k < kmax) re_bit(e, &c->eob[k], 0);
    if (k > 1) re_raw_n(e, v & ((1u << (k - 1)) - 1), k - 1);
}
static uint32_t get_eob(rdec *d, coef_ctx *c) {
    const int kmax = 13;
    int k = 0;
    while (k < kmax && rd_bit(d, &c->eob[k])) k++;
    if (k == 0) return 0;
    if (k == 1) return 1;
    return (1u << (k - 1)) | rd_raw_n(d, k - 1);
}

static void put_coefs(renc *e, const int16_t *lv, const uint16_t pri[8][32]) {
    coef_ctx c; coef_ctx_init(&c, pri);
    uint32_t eob = 0;
    for (int p = NVOX - 1; p >= 0; --p) if (lv[g_scan[p]]) { eob = (uint32_t)p + 1; break; }
    put_eob(e, &c, eob);
    uint32_t hist = 0;
    for (uint32_t p = 0; p < eob; ++p) {
        int idx = g_scan[p];
        int v = lv[idx];
        int dens = __builtin_popcount(hist & 0xFFFFu);
        if (dens > 3) dens = 3;
        if (p + 1 != eob) re_bit(e, &c.sig[g_band_of_raster[idx] * 4 + dens], v != 0);
        hist = (hist << 1) | (v != 0);
        if (!v) continue;
        put_magnitude(e, &c, (uint32_t)(v < 0 ? -v : v));
        re_raw(e, v < 0);
    }
}
/* returns number of decoded significant positions */
static int get_coefs(rdec *d, int16_t *lv, const uint16_t pri[8][32]) {
    coef_ctx c; coef_ctx_init(&c, pri);
    memset(lv, 0, NVOX * sizeof *lv);
    uint32_t eob = get_eob(d, &c);
    if (eob > NVOX) eob = NVOX;
    uint32_t hist = 0;
    int nsig = 0;
    for (uint32_t p = 0; p < eob; ++p) {
        int idx = g_scan[p];
        int sig = 1;
        if (p + 1 != eob) {
            int dens = __builtin_popcount(hist & 0xFFFFu);
            if (dens > 3) dens = 3;
            sig = rd_bit(d, &c.sig[g_band_of_raster[idx] * 4 + dens]);
        }
        hist = (hist << 1) | (uint32_t)sig;
        if (!sig) continue;
        uint32_t m = get_magnitude(d, &c);
        int neg = rd_raw(d);
        lv[idx] = (int16_t)(neg ? -(int32_t)m : (int32_t)m);
        nsig++;
    }
    return nsig;
}

/* ------------------------------------------------------------------------
 * Dead-zone quantiser (src/mc_codec.c:686-692, src/mc_codec_float.h:185-190)
 * ------------------------------------------------------------------------ */
static int16_t quant_one(float c, float rstep) {
    float t = fabsf(c) * rstep + (1.0f - 0.80f);
    t = t > 0.0f ? t : 0.0f;
    t = t < 32767.0f ? t : 32767.0f;
    int32_t v = (int32_t)t;
    return (int16_t)(c < 0.0f ? -v : v);
}
static float dequant_one(int32_t lv, float step) {
    if (!lv) return 0.0f;
    float a = (float)(lv < 0 ? -lv : lv);
    float r = (a - 1.0f) * step + 0.80f * step + 0.40f * step;
    return lv < 0 ? -r : r;
}

/* ------------------------------------------------------------------------
 * Block encode / decode (src/mc_codec.c:629-783)
 * ------------------------------------------------------------------------ */
typedef struct { float step[NVOX], rstep[NVOX]; uint16_t pri[8][32]; } qtab;
static void qtab_build(qtab *t, float q) {
    orc_step_table(q, t->step, t->rstep);
    orc_priors_for_q(q, t->pri);
}

uint32_t orc_enc_block_ex(float q, int tau, const uint8_t vox[NVOX], uint8_t *out, size_t cap,
                          int16_t levels[NVOX], int *dc_out, int *ncorr_out) {
    ensure_tables();
    q = orc_sanitize_q(q);
    if (tau < 0) tau = 0;
    long sum = 0, cnt = 0;
    for (int i = 0; i < NVOX; ++i) { sum += vox[i]; cnt += vox[i] != 0; }
    if (!cnt) { if (dc_out) *dc_out = 0; if (ncorr_out) *ncorr_out = 0; return 0; }
    qtab *t = malloc(sizeof *t);
    qtab_build(t, q);
    int dc = (int)((sum + NVOX / 2) / NVOX);
    float *blk = malloc(NVOX * sizeof(float) * 2), *coef = blk + NVOX;
    int16_t lv[NVOX];
    for (int i = 0; i < NVOX; ++i) blk[i] = (float)((int)vox[i] - dc);
    dct3(blk, coef, line_fwd);
    for (int i = 0; i < NVOX; ++i) lv[i] = quant_one(coef[i], t->rstep[i]);

    /* max-error corrections (src/mc_codec.c:694-709) */
    uint16_t *cpos = malloc(NVOX * sizeof *cpos);
    int32_t *cdel = malloc(NVOX * sizeof *cdel);
    int ncorr = 0;
    if (tau > 0) {
        for (int i = 0; i < NVOX; ++i) blk[i] = dequant_one(lv[i], t->step[i]);
        dct3(blk, coef, line_inv);
        for (int i = 0; i < NVOX; ++i) {
            int v = (int)lrintf(coef[i]) + dc;
            if (v < 0) v = 0;
            if (v > 255) v = 255;
            int err = (int)vox[i] - v, ae = err < 0 ? -err : err;
            if (ae > tau) { cpos[ncorr] = (uint16_t)i; cdel[ncorr] = err < 0 ? -(ae - tau) : (ae - tau); ncorr++; }
        }
    }

    renc e; re_begin(&e, out, cap);
    uint16_t flag[2] = { t->pri[6][0], t->pri[6][1] };
    re_bit(&e, &flag[0], 0);
    re_bit(&e, &flag[1], ncorr > 0);
    for (int b = 7; b >= 0; --b) { uint16_t p = t->pri[7][7 - b]; re_bit(&e, &p, (dc >> b) & 1); }
    put_coefs(&e, lv, t->pri);
    if (ncorr > 0) {
        re_expgolomb(&e, (uint32_t)(ncorr - 1));
        uint32_t prev = 0;
        for (int c = 0; c < ncorr; ++c) {
            re_expgolomb(&e, (uint32_t)cpos[c] - prev);
            prev = cpos[c];
            int32_t dd = cdel[c];
            re_raw(&e, dd < 0);
            re_expgolomb(&e, (uint32_t)(dd < 0 ? -dd : dd) - 1);
        }
    }
    re_finish(&e);
    if (levels) memcpy(levels, lv, sizeof lv);
    if (dc_out) *dc_out = dc;
    if (ncorr_out) *ncorr_out = ncorr;
    free(cpos); free(cdel); free(blk); free(t);
    return (uint32_t)e.n;
}
uint32_t orc_enc_block(float q, int tau, const uint8_t vox[NVOX], uint8_t *out, size_t cap) {
    return orc_enc_block_ex(q, tau, vox, out, cap, NULL, NULL, NULL);
}

static void dec_header_and_coefs(rdec *d, const qtab *t, int16_t *lv, int *dc, int *flags, int *nsig) {
    uint16_t flag[2] = { t->pri[6][0], t->pri[6][1] };
    int f = 0, v = 0;
    f |= rd_bit(d, &flag[0]) ? 1 : 0;
    f |= rd_bit(d, &flag[1]) ? 2 : 0;
    for (int b = 0; b < 8; ++b) { uint16_t p = t->pri[7][b]; v = (v << 1) | rd_bit(d, &p); }
    *dc = v; *flags = f;
    *nsig = get_coefs(d, lv, t->pri);
}

void orc_dec_block_levels(float q, const uint8_t *p, uint32_t plen, int16_t levels[NVOX],
                          int *dc, int *flags, uint32_t *ncorr,
                          uint32_t corr_pos[NVOX], int32_t corr_delta[NVOX]) {
    ensure_tables();
    q = orc_sanitize_q(q);
    qtab *t = malloc(sizeof *t);
    qtab_build(t, q);
    rdec d; rd_begin(&d, p, plen);
    int nsig;
    dec_header_and_coefs(&d, t, levels, dc, flags, &nsig);
    uint32_t nc = 0;
    if (*flags & 2) {
        uint32_t n = rd_expgolomb(&d) + 1, pos = 0;
        if (n > NVOX) n = NVOX;
        for (uint32_t c = 0; c < n; ++c) {
            pos += rd_expgolomb(&d);
            int neg = rd_raw(&d);
            uint32_t m = rd_expgolomb(&d) + 1;
            if (pos >= NVOX) break;
            corr_pos[nc] = pos;
            corr_delta[nc] = neg ? (int32_t)(0u - m) : (int32_t)m;
            nc++;
        }
    }
    *ncorr = nc;
    free(t);
}

static void dec_block_with(const qtab *t, const uint8_t *p, uint32_t plen, uint8_t *dst,
                           float *scratch /* 2*NVOX */) {
    rdec d; rd_begin(&d, p, plen);
    int16_t lv[NVOX];
    int dc, flags, nsig;
    dec_header_and_coefs(&d, t, lv, &dc, &flags, &nsig);
    if (nsig == 0 && !(flags & 2)) { memset(dst, dc, NVOX); return; }
    float *coef = scratch, *out = scratch + NVOX;
    for (int i = 0; i < NVOX; ++i) coef[i] = dequant_one(lv[i], t->step[i]);
    dct3(coef, out, line_inv);
    for (int i = 0; i < NVOX; ++i) {
        int v = (int)lrintf(out[i]) + dc;
        if (v < 0) v = 0;
        if (v > 255) v = 255;
        dst[i] = (uint8_t)v;
    }
    if (flags & 2) {
        /* sparse corrections, count and positions clamped (src/mc_codec.c:770-782) */
        uint32_t n = rd_expgolomb(&d) + 1, pos = 0;
        if (n > NVOX) n = NVOX;
        for (uint32_t c = 0; c < n; ++c) {
            pos += rd_expgolomb(&d);
            int neg = rd_raw(&d);
            uint32_t m = rd_expgolomb(&d) + 1;
            if (pos >= NVOX) break;
            /* 32-bit wrapping add, as the reference's int arithmetic compiles */
            uint32_t delta = neg ? 0u - m : m;
            int v = (int)((uint32_t)dst[pos] + delta);
            if (v < 0) v = 0;
            if (v > 255) v = 255;
            dst[pos] = (uint8_t)v;
        }
    }
}

void orc_dec_block(float q, const uint8_t *p, uint32_t plen, uint8_t dst[NVOX]) {
    ensure_tables();
    q = orc_sanitize_q(q);
    qtab *t = malloc(sizeof *t);
    float *scratch = malloc(2 * NVOX * sizeof(float));
    qtab_build(t, q);
    dec_block_with(t, p, plen, dst, scratch);
    free(scratch); free(t);
}

/* ------------------------------------------------------------------------
 * Fraction map (src/mc_codec.c:600-621): 4096 nibbles, 4 adaptive bins each
 * MSB first, context by previous nibble bucket {0, 1..14, 15}, p0 = 2048.
 * ------------------------------------------------------------------------ */
/* ---- decode-side deblocking (src/mc_codec.c:561-595) ----------------------
 * Every 16-aligned block face of the raster, X faces first, then Y, then Z,
 * in place.  A face at coordinate c of a line moves the voxels at c-1 and c
 * toward each other by a clamped delta when both are material and their step
 * is below beta; c-2 and c+1 (air replaced by the side's own value) shape the
 * delta.  The neighbour at c+1 is taken at the flat offset the reference
 * uses; one past the end of the raster reads as 0. */
static void orc_db_face(uint8_t *v, size_t n, size_t at, size_t step, int beta, int tc) {
    const int a = v[at - step], b = v[at];
    if (a == 0 || b == 0) return;
    const int d = b - a;
    if ((d < 0 ? -d : d) >= beta) return;
    const int far_p = v[at - 2 * step], far_q = at + step < n ? v[at + step] : 0;
    const int pp = far_p ? far_p : a, qq = far_q ? far_q : b;
    int delta = (4 * (b - a) + (pp - qq) + 4) >> 3;
    if (delta > tc) delta = tc;
    if (delta < -tc) delta = -tc;
    int na = a + delta, nb = b - delta;
    na = na < 1 ? 1 : na > 255 ? 255 : na;
    nb = nb < 1 ? 1 : nb > 255 ? 255 : nb;
    v[at - step] = (uint8_t)na;
    v[at] = (uint8_t)nb;
}

void orc_deblock(uint8_t *v, int nz, int ny, int nx, float quality) {
    const int beta = (int)(3.0f * quality + 6.0f), tc = (int)(0.5f * quality + 1.0f);
    const size_t sy = (size_t)nx, sz = (size_t)ny * nx, n = sz * (size_t)nz;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 16; x < nx; x += 16) orc_db_face(v, n, z * sz + y * sy + x, 1, beta, tc);
    for (int z = 0; z < nz; ++z)
        for (int y = 16; y < ny; y += 16)
            for (int x = 0; x < nx; ++x) orc_db_face(v, n, z * sz + y * sy + x, sy, beta, tc);
    for (int z = 16; z < nz; z += 16)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) orc_db_face(v, n, z * sz + y * sy + x, sz, beta, tc);
}

uint32_t orc_enc_fracmap(const uint8_t frac[NVOX], uint8_t *out, size_t cap) {
    renc e; re_begin(&e, out, cap);
    uint16_t cx[3][4];
    for (int b = 0; b < 3; ++b) for (int i = 0; i < 4; ++i) cx[b][i] = 2048;
    int prev = 0;
    for (int i = 0; i < NVOX; ++i) {
        int v = frac[i] & 15, pb = prev == 0 ? 0 : (prev == 15 ? 2 : 1);
        for (int b = 3; b >= 0; --b) re_bit(&e, &cx[pb][b], (v >> b) & 1);
        prev = v;
    }
    re_finish(&e);
    return (uint32_t)e.n;
}
void orc_dec_fracmap(const uint8_t *in, uint32_t len, uint8_t frac[NVOX]) {
    rdec d; rd_begin(&d, in, len);
    uint16_t cx[3][4];
    for (int b = 0; b < 3; ++b) for (int i = 0; i < 4; ++i) cx[b][i] = 2048;
    int prev = 0;
    for (int i = 0; i < NVOX; ++i) {
        int v = 0, pb = prev == 0 ? 0 : (prev == 15 ? 2 : 1);
        for (int b = 3; b >= 0; --b) v |= rd_bit(&d, &cx[pb][b]) << b;
        frac[i] = (uint8_t)v;
        prev = v;
    }
}

/* ------------------------------------------------------------------------
 * XXH64 (published xxHash algorithm; the reference restates it at
 * src/mc_codec.c:383-414).
 * ------------------------------------------------------------------------ */
#define P1 0x9E3779B185EBCA87ULL
#define P2 0xC2B2AE3D27D4EB4FULL
#define P3 0x165667B19E3779F9ULL
#define P4 0x85EBCA77C2B2AE63ULL
#define P5 0x27D4EB2F165667C5ULL
static uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
static uint64_t ld64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }
static uint32_t ld32(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
static uint64_t xround(uint64_t acc, uint64_t in) { return rotl64(acc + in * P2, 31) * P1; }
uint64_t orc_xxh64(const void *data, size_t len, uint64_t seed) {
    const uint8_t *p = data, *end = p + len;
    uint64_t h;
    if (len >= 32) {
        uint64_t a = seed + P1 + P2, b = seed + P2, c = seed, d = seed - P1;
        while (p + 32 <= end) {
            a = xround(a, ld64(p)); b = xround(b, ld64(p + 8));
            c = xround(c, ld64(p + 16)); d = xround(d, ld64(p + 24));
            p += 32;
        }
        h = rotl64(a, 1) + rotl64(b, 7) + rotl64(c, 12) + rotl64(d, 18);
        uint64_t lanes[4] = {a, b, c, d};
        for (int i = 0; i < 4; ++i) { h ^= xround(0, lanes[i]); h = h * P1 + P4; }
    } else {
        h = seed + P5;
    }
    h += (uint64_t)len;
    for (; p + 8 <= end; p += 8) { h ^= xround(0, ld64(p)); h = rotl64(h, 27) * P1 + P4; }
    if (p + 4 <= end) { h ^= (uint64_t)ld32(p) * P1; h = rotl64(h, 23) * P2 + P3; p += 4; }
    for (; p < end; ++p) { h ^= (uint64_t)(*p) * P5; h = rotl64(h, 11) * P1; }
    h ^= h >> 33; h *= P2; h ^= h >> 29; h *= P3; h ^= h >> 32;
    return h;
}

/* ------------------------------------------------------------------------
 * Chunk blob v7 (src/mc_codec.c:809-845, src/mc_archive.c:76-88,
 * src/mc_archive.h:147-237):
 *   [f32 q][u64 hash][u16 fml][fmap fml B][512 B bitmap][u16 len x npresent]
 *   [payloads].  hash = xxh64 chain over fmap, bitmap, lens, payloads.
 * ------------------------------------------------------------------------ */
#define BLOB_HDR 14
#define HASH_SEED 0x6D636368756E6BULL

uint64_t orc_blob_hash(const uint8_t *blob, uint64_t blob_len) {
    uint16_t fml; memcpy(&fml, blob + 12, 2);
    const uint8_t *fmap = blob + BLOB_HDR, *bm = fmap + fml;
    int np = 0;
    for (int i = 0; i < 512; ++i) np += __builtin_popcount(bm[i]);
    const uint8_t *lens = bm + 512;
    uint64_t pay = (uint64_t)BLOB_HDR + fml + 512 + (uint64_t)np * 2;
    uint64_t h = orc_xxh64(fmap, fml, HASH_SEED);
    h ^= orc_xxh64(bm, 512, h);
    h ^= orc_xxh64(lens, (size_t)np * 2, h);
    h ^= orc_xxh64(blob + pay, (size_t)(blob_len - pay), h);
    return h;
}

uint64_t orc_blob_len(const uint8_t *blob) {
    uint16_t fml; memcpy(&fml, blob + 12, 2);
    const uint8_t *bm = blob + BLOB_HDR + fml;
    int np = 0;
    for (int i = 0; i < 512; ++i) np += __builtin_popcount(bm[i]);
    uint64_t tot = (uint64_t)BLOB_HDR + fml + 512 + (uint64_t)np * 2;
    for (int s = 0; s < np; ++s) { uint16_t l; memcpy(&l, bm + 512 + 2 * s, 2); tot += l; }
    return tot;
}

static int gather_block(const uint8_t *chunk, int bi, uint8_t *dst) {
    int bz = bi >> 8, by = (bi >> 4) & 15, bx = bi & 15, any = 0;
    for (int z = 0; z < BLK; ++z)
        for (int y = 0; y < BLK; ++y) {
            const uint8_t *row = chunk + ((size_t)(bz * 16 + z) * 256 + (by * 16 + y)) * 256 + bx * 16;
            memcpy(dst + (z * 16 + y) * 16, row, 16);
            for (int x = 0; x < 16; ++x) any |= row[x];
        }
    return any;
}

size_t orc_encode_chunk(float q, int tau, const uint8_t *chunk, uint8_t *out, size_t cap) {
    /* The blob stores q as given (the parallel append path stores its q
     * argument, src/mc_archive.c:865,883); blocks are coded with the
     * sanitised q a codec context would use. */
    ensure_tables();
    float qs = orc_sanitize_q(q);
    uint8_t bm[512] = {0}, frac[NVOX] = {0};
    uint16_t *lens = malloc(NVOX * sizeof *lens);
    size_t pcap = (size_t)NVOX * (NVOX * 4 + 1024);
    uint8_t *pay = malloc(16u << 20), blk[NVOX];
    size_t paycap = 16u << 20, paylen = 0;
    uint8_t *tmp = malloc(NVOX * 4 + 1024);
    (void)pcap;
    int np = 0;
    for (int bi = 0; bi < NVOX; ++bi) {
        if (!gather_block(chunk, bi, blk)) continue;
        int cnt = 0;
        for (int i = 0; i < NVOX; ++i) cnt += blk[i] != 0;
        frac[bi] = (uint8_t)((cnt * 15 + 2048) / 4096);
        if (cnt && !frac[bi]) frac[bi] = 1;
        uint32_t len = orc_enc_block(qs, tau, blk, tmp, NVOX * 4 + 1024);
        if (!len) continue;
        if (paylen + len > paycap) { paycap *= 2; pay = realloc(pay, paycap); }
        memcpy(pay + paylen, tmp, len);
        paylen += len;
        bm[bi >> 3] |= (uint8_t)(1u << (bi & 7));
        lens[np++] = (uint16_t)len;
    }
    size_t total = 0;
    if (np) {
        uint8_t fmap[NVOX / 2 + 64];
        uint32_t fml32 = orc_enc_fracmap(frac, fmap, sizeof fmap);
        if (fml32 > sizeof fmap) fml32 = sizeof fmap;
        uint16_t fml = (uint16_t)fml32;
        total = BLOB_HDR + fml + 512 + (size_t)np * 2 + paylen;
        if (total <= cap) {
            size_t o = 0;
            memcpy(out + o, &q, 4); o += 4;
            memset(out + o, 0, 8); o += 8;
            memcpy(out + o, &fml, 2); o += 2;
            memcpy(out + o, fmap, fml); o += fml;
            memcpy(out + o, bm, 512); o += 512;
            memcpy(out + o, lens, (size_t)np * 2); o += (size_t)np * 2;
            memcpy(out + o, pay, paylen);
            uint64_t h = orc_blob_hash(out, total);
            memcpy(out + 4, &h, 8);
        }
    }
    free(lens); free(pay); free(tmp);
    return total;
}

typedef struct { uint64_t off[NVOX]; uint32_t len[NVOX]; } blob_index;
static void index_blob(const uint8_t *blob, blob_index *ix) {
    uint16_t fml; memcpy(&fml, blob + 12, 2);
    const uint8_t *bm = blob + BLOB_HDR + fml;
    int np = 0;
    for (int i = 0; i < 512; ++i) np += __builtin_popcount(bm[i]);
    uint64_t at = (uint64_t)BLOB_HDR + fml + 512 + (uint64_t)np * 2;
    int s = 0;
    for (int bi = 0; bi < NVOX; ++bi) {
        if ((bm[bi >> 3] >> (bi & 7)) & 1) {
            uint16_t l; memcpy(&l, bm + 512 + 2 * s, 2);
            ix->off[bi] = at; ix->len[bi] = l; at += l; s++;
        } else { ix->off[bi] = 0; ix->len[bi] = 0; }
    }
}

static void scatter_block(uint8_t *out, int bi, const uint8_t *blk) {
    int bz = bi >> 8, by = (bi >> 4) & 15, bx = bi & 15;
    for (int z = 0; z < 16; ++z)
        for (int y = 0; y < 16; ++y)
            memcpy(out + ((size_t)(bz * 16 + z) * 256 + (by * 16 + y)) * 256 + bx * 16, blk + (z * 16 + y) * 16, 16);
}

void orc_decode_blob_block(const uint8_t *blob, uint64_t blob_len, int bz, int by, int bx,
                           uint8_t dst[NVOX]) {
    ensure_tables();
    blob_index *ix = malloc(sizeof *ix);
    index_blob(blob, ix);
    int bi = (bz * 16 + by) * 16 + bx;
    float q; memcpy(&q, blob, 4);
    if (!ix->off[bi] || ix->off[bi] + ix->len[bi] > blob_len) memset(dst, 0, NVOX);
    else orc_dec_block(q, blob + ix->off[bi], ix->len[bi], dst);
    free(ix);
}

void orc_decode_chunk(const uint8_t *blob, uint64_t blob_len, uint8_t *out) {
    ensure_tables();
    blob_index *ix = malloc(sizeof *ix);
    index_blob(blob, ix);
    float q; memcpy(&q, blob, 4);
    qtab *t = malloc(sizeof *t);
    qtab_build(t, orc_sanitize_q(q));
    float *scratch = malloc(2 * NVOX * sizeof(float));
    uint8_t blk[NVOX];
    for (int bi = 0; bi < NVOX; ++bi) {
        if (!ix->off[bi] || ix->off[bi] + ix->len[bi] > blob_len) memset(blk, 0, NVOX);
        else dec_block_with(t, blob + ix->off[bi], ix->len[bi], blk, scratch);
        scatter_block(out, bi, blk);
    }
    free(scratch); free(t); free(ix);
}

/* ------------------------------------------------------------------------
 * Multi-threaded CPU baseline runners.
 * ------------------------------------------------------------------------ */
typedef struct {
    int n; const uint8_t *const *blobs; const uint64_t *lens; uint8_t *out;
    blob_index *ix; qtab *tabs;
    atomic_uint next;
} mt_dec;
static void *mt_dec_worker(void *arg) {
    mt_dec *m = arg;
    float *scratch = malloc(2 * NVOX * sizeof(float));
    uint8_t blk[NVOX];
    for (;;) {
        unsigned w = atomic_fetch_add(&m->next, 1);
        if (w >= (unsigned)m->n * NVOX) break;
        int c = (int)(w / NVOX), bi = (int)(w % NVOX);
        const blob_index *ix = &m->ix[c];
        if (!ix->off[bi] || ix->off[bi] + ix->len[bi] > m->lens[c]) memset(blk, 0, NVOX);
        else dec_block_with(&m->tabs[c], m->blobs[c] + ix->off[bi], ix->len[bi], blk, scratch);
        scatter_block(m->out + (size_t)c * (256u * 256u * 256u), bi, blk);
    }
    free(scratch);
    return NULL;
}
void orc_decode_chunks_mt(int n, const uint8_t *const *blobs, const uint64_t *lens,
                          uint8_t *out, int nthreads) {
    ensure_tables();
    mt_dec m = { .n = n, .blobs = blobs, .lens = lens, .out = out };
    m.ix = malloc((size_t)n * sizeof *m.ix);
    m.tabs = malloc((size_t)n * sizeof *m.tabs);
    for (int c = 0; c < n; ++c) {
        index_blob(blobs[c], &m.ix[c]);
        float q; memcpy(&q, blobs[c], 4);
        qtab_build(&m.tabs[c], orc_sanitize_q(q));
    }
    atomic_init(&m.next, 0);
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = malloc((size_t)nthreads * sizeof *th);
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, mt_dec_worker, &m);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th); free(m.ix); free(m.tabs);
}

typedef struct {
    int n; const uint8_t *chunks; const float *q; const int *tau;
    uint8_t *const *out; const size_t *caps; size_t *lens;
    atomic_uint next;
} mt_enc;
static void *mt_enc_worker(void *arg) {
    mt_enc *m = arg;
    for (;;) {
        unsigned c = atomic_fetch_add(&m->next, 1);
        if (c >= (unsigned)m->n) break;
        m->lens[c] = orc_encode_chunk(m->q[c], m->tau ? m->tau[c] : 0,
                                      m->chunks + (size_t)c * (256u * 256u * 256u),
                                      m->out[c], m->caps[c]);
    }
    return NULL;
}
void orc_encode_chunks_mt(int n, const uint8_t *chunks, const float *q, const int *tau,
                          uint8_t *const *out_blobs, const size_t *caps, size_t *out_lens,
                          int nthreads) {
    ensure_tables();
    mt_enc m = { .n = n, .chunks = chunks, .q = q, .tau = tau, .out = out_blobs, .caps = caps, .lens = out_lens };
    atomic_init(&m.next, 0);
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = malloc((size_t)nthreads * sizeof *th);
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, mt_enc_worker, &m);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
}
