/*
 * mc_host.c -- plain-C host side of the matter-compressor drop-in library:
 * codec contexts, presets, the v7 .mca archive (header, dense node index,
 * crash-safe append and install, metadata, prior blob, verify) and the
 * decode/encode entry points, which hand all block work to the GPU engine
 * through the mchip C ABI.
 *
 * Format constants and install discipline follow SuperOptimizer/
 * matter-compressor src/mc_archive.h:59-237 and src/mc_archive.c:244-1111
 * (node tables 8-byte aligned and published before the chunk offset, blob
 * appended at EOF, offsets released last), so an archive written here is
 * byte-identical to one the reference writes for the same appends.
 */
#include "matter_compressor_amd.h"

#include <errno.h>
#include <fcntl.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "mc_host_internal.h"
#include "mchip.h"

/* Decode dispatch by the archive's block codec id; CABAC is the only codec
 * (src/mc_codec.c:788-792). */
void mc_dec_block_codec(mc_codec_ctx *c, uint32_t codec, const mc_u8 *p, uint32_t plen, mc_u8 *dst) {
    (void)codec;
    mc_dec_block(c, p, plen, dst);
}

/* ------------------------------------------------------------------------
 * Device context per host thread (reentrancy: concurrent callers each get
 * their own HIP stream and workspace).
 * ------------------------------------------------------------------------ */
static pthread_key_t g_dev_key;
static pthread_once_t g_dev_once = PTHREAD_ONCE_INIT;

static void dev_dtor(void *p) { if (p) mchip_ctx_destroy((mchip_ctx *)p); }
static void dev_key_init(void) { pthread_key_create(&g_dev_key, dev_dtor); }

mchip_ctx *mch_dev(void) {
    pthread_once(&g_dev_once, dev_key_init);
    mchip_ctx *c = pthread_getspecific(g_dev_key);
    if (!c) {
        c = mchip_ctx_create();
        if (!c) {
            fprintf(stderr, "matter_compressor_amd: no usable GPU: %s\n", mchip_last_error());
            return NULL;
        }
        pthread_setspecific(g_dev_key, c);
    }
    return c;
}

void mch_dev_error(const char *what) {
    fprintf(stderr, "matter_compressor_amd: %s failed: %s\n", what, mchip_last_error());
}

/* ------------------------------------------------------------------------
 * Codec context and presets (src/mc_codec.c:477-531)
 * ------------------------------------------------------------------------ */
struct mc_codec_ctx {
    float quality;
    int max_err;
};

static float sanitize_q(float q) { return (!(q > 0.0f) || q > 1024.0f) ? 8.0f : q; }

mc_codec_ctx *mc_codec_ctx_new(void) {
    mc_codec_ctx *c = calloc(1, sizeof *c);
    if (!c) return NULL;
    c->quality = 8.0f;
    c->max_err = 0;
    return c;
}
void mc_codec_ctx_free(mc_codec_ctx *c) { free(c); }
void mc_codec_ctx_set_quality(mc_codec_ctx *c, float q) {
    q = sanitize_q(q);
    if (c) c->quality = q;
}
float mc_codec_ctx_get_quality(mc_codec_ctx *c) { return c ? c->quality : 0.0f; }
void mc_codec_ctx_set_max_error(mc_codec_ctx *c, int tau) { if (c) c->max_err = tau < 0 ? 0 : tau; }
int mc_codec_ctx_get_max_error(mc_codec_ctx *c) { return c ? c->max_err : 0; }

static const struct { float q; int tau; const char *name; } k_presets[MC_PRESET_COUNT] = {
    {0.5f, 1, "archival"}, {0.5f, 2, "master"},   {1.0f, 4, "high"},       {2.5f, 8, "balanced"},
    {6.0f, 16, "streaming"}, {16.0f, 32, "fast"}, {32.0f, 64, "ultrafast"}, {64.0f, 128, "preview"},
};
float mc_preset_quality(mc_preset p) {
    if ((unsigned)p >= MC_PRESET_COUNT) p = MC_PRESET_STREAMING;
    return k_presets[p].q;
}
int mc_preset_tau(mc_preset p) {
    if ((unsigned)p >= MC_PRESET_COUNT) p = MC_PRESET_STREAMING;
    return k_presets[p].tau;
}
const char *mc_preset_name(mc_preset p) {
    if ((unsigned)p >= MC_PRESET_COUNT) return "?";
    return k_presets[p].name;
}
float mc_apply_preset(mc_codec_ctx *c, mc_preset p) {
    if ((unsigned)p >= MC_PRESET_COUNT) p = MC_PRESET_STREAMING;
    mc_codec_ctx_set_quality(c, k_presets[p].q);
    mc_codec_ctx_set_max_error(c, k_presets[p].tau);
    return k_presets[p].q;
}

void mc_codec_init(void) {
    if (mchip_init(-1) != 0) mch_dev_error("mchip_init");
}
void mc_codec_set_priors(const uint16_t *plo, const uint16_t *phi) {
    if (mchip_set_priors(plo, phi) != 0) mch_dev_error("mchip_set_priors");
}

void mc_buf_put(mc_buf *b, const void *s, size_t n) {
    if (b->len + n > b->cap) {
        size_t nc = b->cap ? b->cap * 2 : (size_t)1 << 16;
        while (nc < b->len + n) nc *= 2;
        void *t = realloc(b->p, nc);
        if (!t) return;
        b->p = t;
        b->cap = nc;
    }
    memcpy(b->p + b->len, s, n);
    b->len += n;
}

/* ------------------------------------------------------------------------
 * Block codec on the GPU (src/mc_codec.c:629-783)
 * ------------------------------------------------------------------------ */
int mc_enc_block(mc_codec_ctx *c, const mc_u8 *vox, mc_buf *out, uint32_t *len_out) {
    mchip_ctx *d = mch_dev();
    uint8_t pay[MCH_SLOT_BYTES];
    uint8_t *pp = pay;
    uint32_t len = 0;
    int st = 0;
    float q = c ? c->quality : 8.0f;
    int tau = c ? c->max_err : 0;
    if (!d || mchip_encode_blocks_host(d, vox, &q, &tau, 1, &pp, &len, &st) != 0) {
        mch_dev_error("mc_enc_block");
        *len_out = 0;
        return 0;
    }
    if (st != 0) {                     /* the reference aborts on a payload over its scratch */
        fprintf(stderr, "mc_enc_block: scratch overflow (%u)\n", len);
        abort();
    }
    *len_out = len;
    if (!len) return 0;
    mc_buf_put(out, pay, len);
    return 1;
}

void mc_dec_block(mc_codec_ctx *c, const mc_u8 *p, uint32_t plen, mc_u8 *dst) {
    mchip_ctx *d = mch_dev();
    float q = c ? c->quality : 8.0f;
    const uint8_t *pp = p;
    static const uint8_t empty[1] = {0};
    if (!pp) pp = empty;
    if (!d || mchip_decode_blocks_host(d, &pp, &plen, &q, 1, dst) != 0) {
        mch_dev_error("mc_dec_block");
        memset(dst, 0, 4096);
    }
}

/* Decode-side deblocking on the GPU (src/mc_codec.c:561-595).  On a device
 * failure the volume is left unfiltered and the failure is reported. */
void mc_deblock(mc_u8 *v, int nz, int ny, int nx, float quality) {
    if (!v || nz <= 0 || ny <= 0 || nx <= 0) return;
    mchip_ctx *d = mch_dev();
    if (!d || mchip_deblock_host(d, v, nz, ny, nx, quality) != 0) mch_dev_error("mc_deblock");
}

/* ------------------------------------------------------------------------
 * Host range coder for the per-chunk fraction map only (4096 nibbles,
 * src/mc_codec.c:600-621).  Block payloads never pass through here.
 * ------------------------------------------------------------------------ */
typedef struct { uint8_t *out; size_t cap, n; uint64_t low; uint32_t range; uint8_t cache; uint64_t pend; } hrc_enc;
static void he_put(hrc_enc *e, uint8_t b) { if (e->n < e->cap) e->out[e->n] = b; e->n++; }
static void he_shift(hrc_enc *e) {
    uint32_t top = (uint32_t)(e->low >> 32);
    if (top || e->low < 0xFF000000ull) {
        he_put(e, (uint8_t)(e->cache + top));
        while (--e->pend) he_put(e, (uint8_t)(0xFF + top));
        e->cache = (uint8_t)(e->low >> 24);
    }
    e->pend++;
    e->low = (e->low << 8) & 0xFFFFFFFFull;
}
static void he_bit(hrc_enc *e, uint16_t *p0, int bit) {
    uint32_t r0 = (e->range >> 12) * (uint32_t)*p0;
    if (!bit) { e->range = r0; *p0 = (uint16_t)(*p0 + ((4096 - *p0) >> 4)); }
    else { e->low += r0; e->range -= r0; *p0 = (uint16_t)(*p0 - (*p0 >> 4)); }
    while (e->range < (1u << 24)) { he_shift(e); e->range <<= 8; }
}
typedef struct { const uint8_t *in; size_t len, at; uint32_t code, range; } hrc_dec;
static uint8_t hd_byte(hrc_dec *d) { return d->at < d->len ? d->in[d->at++] : 0; }
static int hd_bit(hrc_dec *d, uint16_t *p0) {
    uint32_t r0 = (d->range >> 12) * (uint32_t)*p0;
    int bit;
    if (d->code < r0) { d->range = r0; bit = 0; *p0 = (uint16_t)(*p0 + ((4096 - *p0) >> 4)); }
    else { d->code -= r0; d->range -= r0; bit = 1; *p0 = (uint16_t)(*p0 - (*p0 >> 4)); }
    while (d->range < (1u << 24)) { d->code = (d->code << 8) | hd_byte(d); d->range <<= 8; }
    return bit;
}

uint32_t mc_enc_fracmap(const mc_u8 *frac, mc_u8 *out, size_t cap) {
    hrc_enc e = {out, cap, 0, 0, 0xFFFFFFFFu, 0, 1};
    uint16_t cx[3][4];
    for (int b = 0; b < 3; ++b) for (int i = 0; i < 4; ++i) cx[b][i] = 2048;
    int prev = 0;
    for (int i = 0; i < 4096; ++i) {
        int v = frac[i] & 15, pb = prev == 0 ? 0 : (prev == 15 ? 2 : 1);
        for (int b = 3; b >= 0; --b) he_bit(&e, &cx[pb][b], (v >> b) & 1);
        prev = v;
    }
    for (int i = 0; i < 5; ++i) he_shift(&e);
    return (uint32_t)e.n;
}
void mc_dec_fracmap(const mc_u8 *in, uint32_t len, mc_u8 *frac) {
    hrc_dec d = {in, len, 0, 0, 0xFFFFFFFFu};
    for (int i = 0; i < 5; ++i) d.code = (d.code << 8) | hd_byte(&d);
    uint16_t cx[3][4];
    for (int b = 0; b < 3; ++b) for (int i = 0; i < 4; ++i) cx[b][i] = 2048;
    int prev = 0;
    for (int i = 0; i < 4096; ++i) {
        int v = 0, pb = prev == 0 ? 0 : (prev == 15 ? 2 : 1);
        for (int b = 3; b >= 0; --b) v |= hd_bit(&d, &cx[pb][b]) << b;
        frac[i] = (uint8_t)v;
        prev = v;
    }
}

/* ------------------------------------------------------------------------
 * XXH64 (published xxHash algorithm) for blob integrity.
 * ------------------------------------------------------------------------ */
#define XP1 0x9E3779B185EBCA87ULL
#define XP2 0xC2B2AE3D27D4EB4FULL
#define XP3 0x165667B19E3779F9ULL
#define XP4 0x85EBCA77C2B2AE63ULL
#define XP5 0x27D4EB2F165667C5ULL
static uint64_t rl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
static uint64_t rd64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }
static uint32_t rd32(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
static uint64_t xr(uint64_t a, uint64_t in) { return rl(a + in * XP2, 31) * XP1; }
uint64_t mch_xxh64(const void *data, size_t len, uint64_t seed) {
    const uint8_t *p = data, *end = p + len;
    uint64_t h;
    if (len >= 32) {
        uint64_t v[4] = {seed + XP1 + XP2, seed + XP2, seed, seed - XP1};
        while (p + 32 <= end) {
            for (int i = 0; i < 4; ++i) v[i] = xr(v[i], rd64(p + 8 * i));
            p += 32;
        }
        h = rl(v[0], 1) + rl(v[1], 7) + rl(v[2], 12) + rl(v[3], 18);
        for (int i = 0; i < 4; ++i) { h ^= xr(0, v[i]); h = h * XP1 + XP4; }
    } else {
        h = seed + XP5;
    }
    h += (uint64_t)len;
    for (; p + 8 <= end; p += 8) { h ^= xr(0, rd64(p)); h = rl(h, 27) * XP1 + XP4; }
    if (p + 4 <= end) { h ^= (uint64_t)rd32(p) * XP1; h = rl(h, 23) * XP2 + XP3; p += 4; }
    for (; p < end; ++p) { h ^= (uint64_t)(*p) * XP5; h = rl(h, 11) * XP1; }
    h ^= h >> 33; h *= XP2; h ^= h >> 29; h *= XP3; h ^= h >> 32;
    return h;
}

/* blob v7: [f32 q][u64 hash][u16 fml][fmap][512 B bitmap][u16 lens][payloads] */
uint64_t mc_chunk_compute_hash(const uint8_t *blob, uint64_t blob_len) {
    uint16_t fml; memcpy(&fml, blob + 12, 2);
    const uint8_t *fmap = blob + MCH_BLOB_HDR, *bm = fmap + fml;
    int np = 0;
    for (int i = 0; i < 512; ++i) np += __builtin_popcount(bm[i]);
    uint64_t pay = (uint64_t)MCH_BLOB_HDR + fml + 512 + (uint64_t)np * 2;
    uint64_t h = mch_xxh64(fmap, fml, MCH_HASH_SEED);
    h ^= mch_xxh64(bm, 512, h);
    h ^= mch_xxh64(bm + 512, (size_t)np * 2, h);
    h ^= mch_xxh64(blob + pay, (size_t)(blob_len - pay), h);
    return h;
}

/* Structural length of the blob at off, every read bounded by end; 0 if it
 * does not fit (src/mc_archive.h:212-226). */
uint64_t mch_blob_len_bounded(const uint8_t *arc, uint64_t off, uint64_t end) {
    if (off > end || end - off < (uint64_t)MCH_BLOB_HDR) return 0;
    uint16_t fml; memcpy(&fml, arc + off + 12, 2);
    uint64_t bm = off + MCH_BLOB_HDR + fml;
    if (bm > end || end - bm < 512) return 0;
    int np = 0;
    for (int i = 0; i < 512; ++i) np += __builtin_popcount(arc[bm + i]);
    uint64_t lens = bm + 512;
    if (end - lens < (uint64_t)np * 2) return 0;
    uint64_t tot = 0;
    for (int s = 0; s < np; ++s) { uint16_t l; memcpy(&l, arc + lens + 2 * s, 2); tot += l; }
    uint64_t total = lens + (uint64_t)np * 2 + tot - off;
    if (off + total > end) return 0;
    return total;
}

/* ------------------------------------------------------------------------
 * Archive: header, node tree, mmap writer (src/mc_archive.c:244-564)
 * ------------------------------------------------------------------------ */
#define MC_RESERVE_MAX (10ull * 1024 * 1024 * 1024 * 1024)
#define MC_GROW (1ull << 30)

static uint64_t reserve_for(int nx, int ny, int nz) {
    uint64_t need = 0, dz = (uint64_t)nz, dy = (uint64_t)ny, dx = (uint64_t)nx;
    for (int l = 0; l < 8; ++l) {
        need += ((dz + 255) / 256 * 256) * ((dy + 255) / 256 * 256) * ((dx + 255) / 256 * 256);
        dz = (dz + 1) / 2; dy = (dy + 1) / 2; dx = (dx + 1) / 2;
    }
    uint64_t r = need + need / 2 + (1ull << 30);
    return r > MC_RESERVE_MAX ? MC_RESERVE_MAX : r;
}

static int grow_to(mc_archive *a, uint64_t need) {
    uint64_t fl = atomic_load_explicit(&a->file_len, memory_order_acquire);
    if (need <= fl) return 0;
    pthread_mutex_lock(&a->grow_mu);
    fl = atomic_load_explicit(&a->file_len, memory_order_relaxed);
    int rc = 0;
    if (need > fl) {
        uint64_t nf = fl;
        while (nf < need) nf += MC_GROW;
        if (nf > a->reserve) {
            fprintf(stderr, "mc_archive: grow beyond reservation\n");
            rc = -1;
        } else {
            int g = fallocate(a->fd, 0, (off_t)fl, (off_t)(nf - fl));
            if (g != 0 && (errno == EOPNOTSUPP || errno == ENOSYS)) g = ftruncate(a->fd, (off_t)nf);
            if (g != 0) rc = -1;
            else atomic_store_explicit(&a->file_len, nf, memory_order_release);
        }
    }
    pthread_mutex_unlock(&a->grow_mu);
    return rc;
}
static uint64_t alloc_at_eof(mc_archive *a, uint64_t n) {
    uint64_t off = atomic_fetch_add_explicit(&a->cursor, n, memory_order_relaxed);
    return grow_to(a, off + n) ? ~0ull : off;
}
static uint64_t alloc_aligned(mc_archive *a, uint64_t n, uint64_t align) {
    uint64_t cur = atomic_load_explicit(&a->cursor, memory_order_relaxed);
    for (;;) {
        uint64_t off = (cur + align - 1) & ~(align - 1);
        if (atomic_compare_exchange_weak_explicit(&a->cursor, &cur, off + n, memory_order_relaxed,
                                                  memory_order_relaxed))
            return grow_to(a, off + n) ? ~0ull : off;
    }
}
static _Atomic uint64_t *slot_at(mc_archive *a, uint64_t at) { return (_Atomic uint64_t *)(void *)(a->base + at); }

static inline int nib(int c, int level) { return (c >> (4 * level)) & 15; }

/* Offset of the shard slot for (lod,cz,cy,cx), creating zeroed node tables
 * on the way (root, inner, shard; each published by CAS into its parent). */
static uint64_t shard_slot(mc_archive *a, int lod, int cz, int cy, int cx) {
    uint64_t root = atomic_load_explicit(slot_at(a, MCH_ROOTOFF + (uint64_t)lod * 8), memory_order_acquire);
    if (!root) {
        uint64_t no = alloc_aligned(a, MCH_NODE_BYTES, 8);
        if (no == ~0ull) return ~0ull;
        memset(a->base + no, 0, MCH_NODE_BYTES);
        uint64_t exp = 0;
        root = atomic_compare_exchange_strong_explicit(slot_at(a, MCH_ROOTOFF + (uint64_t)lod * 8), &exp, no,
                                                       memory_order_release, memory_order_acquire) ? no : exp;
    }
    uint64_t node = root;
    for (int lv = 2; lv >= 1; --lv) {
        uint64_t at = node + (uint64_t)((nib(cz, lv) * 16 + nib(cy, lv)) * 16 + nib(cx, lv)) * 8;
        uint64_t child = atomic_load_explicit(slot_at(a, at), memory_order_acquire);
        if (!child) {
            uint64_t no = alloc_aligned(a, MCH_NODE_BYTES, 8);
            if (no == ~0ull) return ~0ull;
            memset(a->base + no, 0, MCH_NODE_BYTES);
            uint64_t exp = 0;
            child = atomic_compare_exchange_strong_explicit(slot_at(a, at), &exp, no, memory_order_release,
                                                            memory_order_acquire) ? no : exp;
        }
        node = child;
    }
    return node + (uint64_t)((nib(cz, 0) * 16 + nib(cy, 0)) * 16 + nib(cx, 0)) * 8;
}

static int install_blob(mc_archive *a, int lod, int cz, int cy, int cx, const uint8_t *blob, size_t len) {
    pthread_mutex_lock(&a->write_mu);
    uint64_t slot = shard_slot(a, lod, cz, cy, cx);
    uint64_t off = slot == ~0ull ? ~0ull : alloc_at_eof(a, len);
    if (off == ~0ull) { pthread_mutex_unlock(&a->write_mu); return -1; }
    memcpy(a->base + off, blob, len);
    atomic_store_explicit(slot_at(a, slot), off, memory_order_release);
    atomic_fetch_add_explicit(&a->gen, 1, memory_order_release);
    atomic_store_explicit(slot_at(a, MCH_TOTLEN), atomic_load_explicit(&a->cursor, memory_order_relaxed),
                          memory_order_relaxed);
    pthread_mutex_unlock(&a->write_mu);
    return 0;
}
static int mark_zero(mc_archive *a, int lod, int cz, int cy, int cx) {
    pthread_mutex_lock(&a->write_mu);
    uint64_t slot = shard_slot(a, lod, cz, cy, cx);
    if (slot == ~0ull) { pthread_mutex_unlock(&a->write_mu); return -1; }
    atomic_store_explicit(slot_at(a, slot), MCH_SLOT_ZERO, memory_order_release);
    atomic_fetch_add_explicit(&a->gen, 1, memory_order_release);
    pthread_mutex_unlock(&a->write_mu);
    return 0;
}

/* The archive's prior override (header @128 -> "PRIS" blob) becomes the
 * process-wide priors; absent or invalid restores the built-in tables
 * (src/mc_archive.c:666-672). */
void priors_load(const uint8_t *base) {
    uint64_t off; memcpy(&off, base + MCH_PRIOROFF, 8);
    uint32_t magic = 0;
    if (off) memcpy(&magic, base + off, 4);
    if (!off || magic != MCH_PRIORS_MAGIC) { mc_codec_set_priors(NULL, NULL); return; }
    uint16_t lo[256], hi[256];
    memcpy(lo, base + off + 8, sizeof lo);
    memcpy(hi, base + off + 8 + sizeof lo, sizeof hi);
    mc_codec_set_priors(lo, hi);
}

mc_archive *mc_archive_open(const char *path, int dim, float quality) {
    return mc_archive_open_dims(path, dim, dim, dim, quality);
}

mc_archive *mc_archive_open_dims(const char *path, int nx, int ny, int nz, float quality) {
    if (nx <= 0 || ny <= 0 || nz <= 0) { fprintf(stderr, "mc_archive_open: bad dims\n"); return NULL; }
    mc_codec_init();
    int fd = open(path, O_RDWR | O_CREAT, 0644);
    if (fd < 0) { perror("mc_archive_open: open"); return NULL; }
    struct stat sb;
    if (fstat(fd, &sb) != 0) { perror("fstat"); close(fd); return NULL; }
    int fresh = sb.st_size == 0;
    uint64_t reserve = reserve_for(nx, ny, nz);
    uint64_t len0 = fresh ? MCH_META_END : (uint64_t)sb.st_size;
    if (fresh && ftruncate(fd, (off_t)len0) != 0) { perror("ftruncate"); close(fd); return NULL; }
    uint8_t *base = mmap(NULL, reserve, PROT_READ | PROT_WRITE, MAP_SHARED | MAP_NORESERVE, fd, 0);
    if (base == MAP_FAILED) { perror("mmap"); close(fd); return NULL; }
    mc_archive *a = calloc(1, sizeof *a);
    if (!a) {
        fprintf(stderr, "mc_archive_open: out of memory\n");
        munmap(base, reserve);
        close(fd);
        return NULL;
    }
    a->fd = fd; a->base = base; a->reserve = reserve;
    a->nx = nx; a->ny = ny; a->nz = nz; a->dim = nx; a->quality = quality; a->cov = NULL;
    pthread_mutex_init(&a->grow_mu, NULL);
    pthread_mutex_init(&a->write_mu, NULL);
    pthread_mutex_init(&a->aux_mu, NULL);
    atomic_store(&a->file_len, len0);
    if (fresh) {
        memset(base, 0, MCH_HDR);
        uint32_t magic = MCH_MAGIC, ver = MCH_VERSION, ux = (uint32_t)nx, uy = (uint32_t)ny, uz = (uint32_t)nz;
        memcpy(base + 0, &magic, 4); memcpy(base + 4, &ver, 4);
        memcpy(base + 12, &ux, 4); memcpy(base + 16, &uy, 4); memcpy(base + 20, &uz, 4);
        uint64_t metaoff = MCH_HDR, metacap = MCH_META_END - MCH_HDR, totlen = MCH_META_END;
        memcpy(base + 96, &metaoff, 8); memcpy(base + 104, &metacap, 8); memcpy(base + MCH_TOTLEN, &totlen, 8);
        memcpy(base + 120, &quality, 4);
        a->block_codec = MC_BLOCKCODEC_CABAC;
        atomic_store(&a->cursor, MCH_META_END);
    } else {
        uint32_t magic, ver, ux, uy, uz;
        memcpy(&magic, base, 4); memcpy(&ver, base + 4, 4);
        memcpy(&ux, base + 12, 4); memcpy(&uy, base + 16, 4); memcpy(&uz, base + 20, 4);
        if (magic != MCH_MAGIC || ver != MCH_VERSION || (int)ux != nx || (int)uy != ny || (int)uz != nz) {
            fprintf(stderr, "mc_archive_open: %s is not a matching mc archive (magic/ver/dims)\n", path);
            munmap(base, reserve); close(fd);
            pthread_mutex_destroy(&a->grow_mu); pthread_mutex_destroy(&a->write_mu);
            pthread_mutex_destroy(&a->aux_mu);
            free(a);
            return NULL;
        }
        uint64_t tot; memcpy(&tot, base + MCH_TOTLEN, 8);
        if (tot < MCH_META_END) tot = MCH_META_END;
        atomic_store(&a->cursor, tot);
        a->block_codec = MC_BLOCKCODEC_CABAC;
        priors_load(base);
    }
    return a;
}

void mc_archive_close(mc_archive *a) {
    if (!a) return;
    uint64_t cur = atomic_load(&a->cursor);
    memcpy(a->base + MCH_TOTLEN, &cur, 8);
    msync(a->base, cur, MS_SYNC);
    munmap(a->base, a->reserve);
    if (ftruncate(a->fd, (off_t)cur) != 0) perror("mc_archive_close: ftruncate");
    close(a->fd);
    pthread_mutex_destroy(&a->grow_mu);
    pthread_mutex_destroy(&a->write_mu);
    pthread_mutex_destroy(&a->aux_mu);
    free(a->fr_cache);
    free(a);
}

int mc_archive_set_block_codec(mc_archive *a, uint32_t codec) {
    if (!a || codec != MC_BLOCKCODEC_CABAC) return -1;
    a->block_codec = codec;
    memcpy(a->base + 124, &codec, 4);
    return 0;
}
uint32_t mc_archive_block_codec(const mc_archive *a) { return a ? a->block_codec : MC_BLOCKCODEC_CABAC; }

int mc_archive_reserve_index(mc_archive *a, int lod, int cz, int cy, int cx) {
    if (!a || lod < 0 || lod > 7) return -1;
    return shard_slot(a, lod, cz, cy, cx) == ~0ull ? -1 : 0;
}

uint64_t mc_archive_data_len(mc_archive *a) {
    return a ? atomic_load_explicit(&a->cursor, memory_order_relaxed) : 0;
}

int mc_archive_set_metadata(mc_archive *a, const void *data, size_t len) {
    if (!a || (len && !data)) return -1;
    if (len > MCH_META_END - MCH_HDR) return -1;
    if (len) memcpy(a->base + MCH_HDR, data, len);
    atomic_thread_fence(memory_order_release);
    uint64_t l = len;
    memcpy(a->base + 112, &l, 8);
    return 0;
}
const char *mc_metadata(const uint8_t *arc, size_t *out_len) {
    if (!arc) { if (out_len) *out_len = 0; return NULL; }
    uint64_t len; memcpy(&len, arc + 112, 8);
    if (len > MCH_META_END - MCH_HDR) len = MCH_META_END - MCH_HDR;
    if (out_len) *out_len = (size_t)len;
    return (const char *)(arc + MCH_HDR);
}
const char *mc_archive_metadata(mc_archive *a, size_t *out_len) {
    if (!a) { if (out_len) *out_len = 0; return NULL; }
    return mc_metadata(a->base, out_len);
}

int mc_archive_set_priors(struct mc_archive *a, const uint16_t plo[8][32], const uint16_t phi[8][32]) {
    if (!a || !plo || !phi) return -1;
    uint64_t off = alloc_at_eof(a, MCH_PRIORS_BYTES);
    if (off == ~0ull) return -1;
    uint32_t magic = MCH_PRIORS_MAGIC, ver = 1;
    memcpy(a->base + off, &magic, 4);
    memcpy(a->base + off + 4, &ver, 4);
    memcpy(a->base + off + 8, plo, 512);
    memcpy(a->base + off + 8 + 512, phi, 512);
    atomic_thread_fence(memory_order_release);
    memcpy(a->base + MCH_PRIOROFF, &off, 8);
    mc_codec_set_priors((const uint16_t *)plo, (const uint16_t *)phi);
    return 0;
}

/* Walk the dense node tree; every slot read is bounded by end
 * (src/mc_archive.h:133-145). */
uint64_t mch_resolve(const uint8_t *arc, uint64_t root, int cz, int cy, int cx, uint64_t end) {
    uint64_t node = root;
    for (int lv = 2; lv >= 0; --lv) {
        if (!node) return 0;
        int idx = (nib(cz, lv) * 16 + nib(cy, lv)) * 16 + nib(cx, lv);
        if (node > end || (uint64_t)idx * 8 + 8 > end - node) return 0;
        uint64_t child; memcpy(&child, arc + node + (uint64_t)idx * 8, 8);
        node = child;
    }
    return node;
}

uint64_t mc_archive_chunk_offset(mc_archive *a, int lod, int cz, int cy, int cx) {
    if (!a || lod < 0 || lod > 7) return 0;
    uint64_t root; memcpy(&root, a->base + MCH_ROOTOFF + (uint64_t)lod * 8, 8);
    return mch_resolve(a->base, root, cz, cy, cx, atomic_load_explicit(&a->file_len, memory_order_relaxed));
}

mc_cover mc_archive_chunk_coverage(mc_archive *a, int lod, int cz, int cy, int cx) {
    if (!a || lod < 0 || lod > 7) return MC_ABSENT;
    uint64_t off = mc_archive_chunk_offset(a, lod, cz, cy, cx);
    if (!off) return MC_ABSENT;
    return off == MCH_SLOT_ZERO ? MC_ZERO : MC_PRESENT;
}

/* Payload range of block bi inside the chunk blob at co (0 = absent). */
int mch_block_range(mc_archive *a, uint64_t co, int bi, uint64_t *off, uint32_t *len) {
    uint64_t end = atomic_load_explicit(&a->cursor, memory_order_acquire);
    const uint8_t *arc = a->base;
    if (co + MCH_BLOB_HDR > end) return 0;
    uint16_t fml; memcpy(&fml, arc + co + 12, 2);
    uint64_t bm = co + MCH_BLOB_HDR + fml;
    if (bm + 512 > end) return 0;
    if (!((arc[bm + (bi >> 3)] >> (bi & 7)) & 1)) return 0;
    int np = 0, rank = 0;
    for (int i = 0; i < 512; ++i) {
        int c = __builtin_popcount(arc[bm + i]);
        np += c;
        if (i < (bi >> 3)) rank += c;
    }
    rank += __builtin_popcount(arc[bm + (bi >> 3)] & ((1u << (bi & 7)) - 1u));
    uint64_t lens = bm + 512;
    if (lens + 2ull * np > end) return 0;
    uint64_t pay = lens + 2ull * np;
    for (int s = 0; s < rank; ++s) { uint16_t l; memcpy(&l, arc + lens + 2 * s, 2); pay += l; }
    uint16_t l; memcpy(&l, arc + lens + 2 * rank, 2);
    if (pay + l > end) return 0;
    *off = pay;
    *len = l;
    return 1;
}

int mc_archive_block_blob(mc_archive *a, uint64_t chunk_off, int bz, int by, int bx, const uint8_t **ptr,
                          uint32_t *len) {
    if (ptr) *ptr = NULL;
    if (len) *len = 0;
    if (!a || chunk_off <= MCH_SLOT_ZERO) return 0;
    uint64_t o; uint32_t l;
    if (!mch_block_range(a, chunk_off, (bz * 16 + by) * 16 + bx, &o, &l)) return 0;
    if (ptr) *ptr = a->base + o;
    if (len) *len = l;
    return 1;
}

void mc_archive_decode_block(mc_archive *a, uint64_t chunk_off, int bz, int by, int bx, mc_u8 *dst) {
    if (!a || chunk_off <= MCH_SLOT_ZERO) { memset(dst, 0, 4096); return; }
    uint64_t o; uint32_t l;
    if (!mch_block_range(a, chunk_off, (bz * 16 + by) * 16 + bx, &o, &l)) { memset(dst, 0, 4096); return; }
    float q; memcpy(&q, a->base + chunk_off, 4);
    mchip_ctx *d = mch_dev();
    const uint8_t *p = a->base + o;
    if (!d || mchip_decode_blocks_host(d, &p, &l, &q, 1, dst) != 0) {
        mch_dev_error("mc_archive_decode_block");
        memset(dst, 0, 4096);
    }
}

void mc_archive_decode_chunk(mc_archive *a, uint64_t chunk_off, mc_u8 *out, int nthreads) {
    (void)nthreads;
    if (!a || !out) return;
    if (chunk_off <= MCH_SLOT_ZERO) { memset(out, 0, (size_t)MCH_CHUNK_BYTES); return; }
    uint64_t end = atomic_load_explicit(&a->cursor, memory_order_acquire);
    uint64_t len = mch_blob_len_bounded(a->base, chunk_off, (uint64_t)-1);
    if (chunk_off >= end) len = 0;
    else if (!len || len > end - chunk_off) len = end - chunk_off;   /* the payload bound is the live cursor */
    mchip_ctx *d = mch_dev();
    const uint8_t *b = a->base + chunk_off;
    /* concurrent callers join one batched decode on the device */
    if (!d || mchip_decode_chunk_merged(d, b, len, out) != 0) {
        mch_dev_error("mc_archive_decode_chunk");
        memset(out, 0, (size_t)MCH_CHUNK_BYTES);
    }
}

/* ------------------------------------------------------------------------
 * Append paths (src/mc_archive.c:566-614, :822-901)
 * ------------------------------------------------------------------------ */
/* Encode one chunk on the GPU; returns blob length (0 = all air), -1 error. */
static long long encode_one(const mc_u8 *vox, float q, int tau, float q_store, uint8_t **blob) {
    mchip_ctx *d = mch_dev();
    if (!d) return -1;
    uint64_t cap = 24ull << 20, blen = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        uint8_t *buf = malloc(cap);
        if (!buf) return -1;
        int st = 0;
        if (mchip_encode_chunks_host(d, vox, &q, &tau, &q_store, 1, &buf, &cap, &blen, &st) != 0) {
            mch_dev_error("chunk encode");
            free(buf);
            return -1;
        }
        if (st == 0) { *blob = buf; return (long long)blen; }
        free(buf);
        if (st == 1) { fprintf(stderr, "mc_enc_block: scratch overflow\n"); abort(); }
        cap = blen + 4096;        /* st == 2: capacity too small */
    }
    return -1;
}

int mc_archive_append_chunk_par(mc_archive *a, int lod, int cz, int cy, int cx, const mc_u8 vox[256 * 256 * 256],
                                float q, int nthreads) {
    (void)nthreads;
    if (!a || lod < 0 || lod > 7 || !vox) return -1;
    float qq = q > 0 ? q : a->quality;
    uint8_t *blob = NULL;
    long long n = encode_one(vox, qq, 0, qq, &blob);
    if (n < 0) return -1;
    int rc = n ? install_blob(a, lod, cz, cy, cx, blob, (size_t)n) : 0;   /* all air: slot stays absent */
    free(blob);
    return rc;
}

int mc_archive_append_chunk_ctx(mc_archive *a, mc_codec_ctx *C, int lod, int cz, int cy, int cx,
                                const mc_u8 vox[256 * 256 * 256]) {
    if (!a || !C || lod < 0 || lod > 7 || !vox) return -1;
    uint8_t *blob = NULL;
    long long n = encode_one(vox, C->quality, C->max_err, C->quality, &blob);
    if (n < 0) return -1;
    int rc = n ? install_blob(a, lod, cz, cy, cx, blob, (size_t)n) : mark_zero(a, lod, cz, cy, cx);
    free(blob);
    return rc;
}

int mc_archive_append_chunk_raw_q(mc_archive *a, int lod, int cz, int cy, int cx, const mc_u8 vox[256 * 256 * 256],
                                  float q) {
    if (!a || lod < 0 || lod > 7 || !vox) return -1;
    mc_codec_ctx C = {sanitize_q(q), 0};
    return mc_archive_append_chunk_ctx(a, &C, lod, cz, cy, cx, vox);
}

int mc_archive_append_chunk_raw(mc_archive *a, int lod, int cz, int cy, int cx, const mc_u8 vox[256 * 256 * 256]) {
    return mc_archive_append_chunk_raw_q(a, lod, cz, cy, cx, vox, a ? a->quality : 8.0f);
}

/* Rate control (src/mc_archive.c:618-648): the 256 blocks (bz, by, (bz+by)&15)
 * are sample-encoded at the archive's quality on the GPU (one launch); their
 * payload bytes, scaled to 4096 blocks, set q once by bytes ~ q^-0.75; the
 * chunk is then encoded at that q. */
int mc_archive_append_chunk_target(mc_archive *a, int lod, int cz, int cy, int cx,
                                   const mc_u8 vox[256 * 256 * 256], float target_ratio, float *q_out) {
    if (!a || lod < 0 || lod > 7 || !vox || !(target_ratio > 1.0f)) return -1;
    mc_codec_ctx C = {8.0f, 0};
    mc_codec_ctx_set_quality(&C, a->quality);
    uint8_t *blk = malloc(256 * 4096);
    uint8_t *pay = malloc((size_t)256 * MCH_SLOT_BYTES);
    uint8_t *pp[256];
    uint32_t lens[256];
    float qv[256];
    int tv[256], st[256], nb = 0;
    if (!blk || !pay) { free(blk); free(pay); return -1; }
    for (int bz = 0; bz < 16; ++bz)
        for (int by = 0; by < 16; ++by) {
            const int bx = (bz + by) & 15;
            uint8_t *d = blk + (size_t)nb * 4096;
            int any = 0;
            for (int z = 0; z < 16; ++z)
                for (int y = 0; y < 16; ++y) {
                    const mc_u8 *src = vox + ((size_t)(16 * bz + z) * 256 + (16 * by + y)) * 256 + 16 * bx;
                    memcpy(d + (z * 16 + y) * 16, src, 16);
                    for (int x = 0; x < 16; ++x) any |= src[x];
                }
            if (!any) continue;                       /* air: counted as sampled, no bytes */
            pp[nb] = pay + (size_t)nb * MCH_SLOT_BYTES;
            qv[nb] = C.quality;
            tv[nb] = 0;
            nb++;
        }
    size_t sample_bytes = 0;
    if (nb) {
        mchip_ctx *d = mch_dev();
        if (!d || mchip_encode_blocks_host(d, blk, qv, tv, nb, pp, lens, st) != 0) {
            mch_dev_error("mc_archive_append_chunk_target");
            free(blk); free(pay);
            return -1;
        }
        for (int i = 0; i < nb; ++i) sample_bytes += lens[i];
    }
    free(blk); free(pay);
    const float q0 = a->quality;                      /* as given; the sample encode used it sanitised */
    float q = q0;
    if (sample_bytes) {
        const double est_total = (double)sample_bytes * (4096.0 / 256);
        const double want_total = (double)(256.0 * 256.0 * 256.0) / target_ratio;
        q = (float)(q0 * pow(est_total / want_total, 1.0 / 0.75f));
        if (q < 0.5f) q = 0.5f;
        if (q > 24.0f) q = 24.0f;
    }
    if (q_out) *q_out = q;
    return mc_archive_append_chunk_raw_q(a, lod, cz, cy, cx, vox, q);
}

int mc_archive_append_chunk_compressed(mc_archive *a, int lod, int cz, int cy, int cx, const uint8_t *blob,
                                       size_t len) {
    if (!a || lod < 0 || lod > 7 || !blob || !len) return -1;
    return install_blob(a, lod, cz, cy, cx, blob, len);
}

/* ------------------------------------------------------------------------
 * Fraction map, occupancy, sampler (src/mc_archive.c:903-948).  The map of
 * a chunk is decoded once and cached (keyed by offset and stored hash);
 * the sampler's draws and acceptance test are unchanged.
 * ------------------------------------------------------------------------ */
int mc_archive_block_present(mc_archive *a, int lod, int bz, int by, int bx) {
    if (!a || lod < 0 || lod > 7 || bz < 0 || by < 0 || bx < 0) return 0;
    uint64_t co = mc_archive_chunk_offset(a, lod, bz >> 4, by >> 4, bx >> 4);
    if (co <= MCH_SLOT_ZERO) return 0;
    uint16_t fml; memcpy(&fml, a->base + co + 12, 2);
    const uint8_t *bm = a->base + co + MCH_BLOB_HDR + fml;
    int bi = ((bz & 15) * 16 + (by & 15)) * 16 + (bx & 15);
    return (bm[bi >> 3] >> (bi & 7)) & 1;
}

static const uint8_t *fracmap_of(mc_archive *a, uint64_t co) {
    uint64_t tag; memcpy(&tag, a->base + co + 4, 8);
    pthread_mutex_lock(&a->aux_mu);
    if (!a->fr_cache) {
        a->fr_cache = calloc(1024, sizeof *a->fr_cache);
        if (!a->fr_cache) { pthread_mutex_unlock(&a->aux_mu); return NULL; }
        a->fr_cap = 1024;
    }
    uint64_t h = (co * 0x9E3779B97F4A7C15ull) >> 20;
    for (uint32_t p = 0; p < a->fr_cap; ++p) {
        mch_frac_entry *e = &a->fr_cache[(h + p) & (a->fr_cap - 1)];
        if (e->co == co && e->tag == tag) { pthread_mutex_unlock(&a->aux_mu); return e->frac; }
        if (e->co == 0 || p == a->fr_cap - 1) {
            if (e->co != 0) e = &a->fr_cache[h & (a->fr_cap - 1)];     /* full: overwrite the home slot */
            uint16_t fml; memcpy(&fml, a->base + co + 12, 2);
            mc_dec_fracmap(a->base + co + MCH_BLOB_HDR, fml, e->frac);
            if (!fml) memset(e->frac, 0, sizeof e->frac);
            e->co = co;
            e->tag = tag;
            pthread_mutex_unlock(&a->aux_mu);
            return e->frac;
        }
    }
    pthread_mutex_unlock(&a->aux_mu);
    return NULL;
}

float mc_archive_block_fraction(mc_archive *a, int lod, int bz, int by, int bx) {
    if (!a || lod < 0 || lod > 7 || bz < 0 || by < 0 || bx < 0) return 0.0f;
    uint64_t co = mc_archive_chunk_offset(a, lod, bz >> 4, by >> 4, bx >> 4);
    if (co <= MCH_SLOT_ZERO) return 0.0f;
    uint16_t fml; memcpy(&fml, a->base + co + 12, 2);
    if (!fml) return 0.0f;
    const uint8_t *fr = fracmap_of(a, co);
    if (!fr) return 0.0f;
    return (float)fr[((bz & 15) * 16 + (by & 15)) * 16 + (bx & 15)] / 15.0f;
}

static uint64_t xorshift64(uint64_t *s) {
    uint64_t x = *s;
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    *s = x;
    return x;
}
int mc_archive_sample_boxes(mc_archive *a, int lod, uint64_t seed, int count, long dz, long dy, long dx,
                            float min_frac, mc_box *out) {
    if (!a || !out || count <= 0 || dz <= 0 || dy <= 0 || dx <= 0) return 0;
    uint32_t unx, uny, unz;
    memcpy(&unx, a->base + 12, 4); memcpy(&uny, a->base + 16, 4); memcpy(&unz, a->base + 20, 4);
    long NX = (long)unx >> lod, NY = (long)uny >> lod, NZ = (long)unz >> lod;
    if (NZ < dz || NY < dy || NX < dx) return 0;
    uint64_t s = seed ? seed : 0x9E3779B97F4A7C15ull;
    int got = 0;
    long attempts = 0, max_attempts = (long)count * 256;
    while (got < count && attempts++ < max_attempts) {
        long z0 = (long)(xorshift64(&s) % (uint64_t)(NZ - dz + 1));
        long y0 = (long)(xorshift64(&s) % (uint64_t)(NY - dy + 1));
        long x0 = (long)(xorshift64(&s) % (uint64_t)(NX - dx + 1));
        long bz0 = z0 >> 4, bz1 = (z0 + dz - 1) >> 4, by0 = y0 >> 4, by1 = (y0 + dy - 1) >> 4;
        long bx0 = x0 >> 4, bx1 = (x0 + dx - 1) >> 4;
        double fsum = 0;
        long nb = 0;
        for (long bz = bz0; bz <= bz1; ++bz)
            for (long by = by0; by <= by1; ++by)
                for (long bx = bx0; bx <= bx1; ++bx) {
                    fsum += mc_archive_block_fraction(a, lod, (int)bz, (int)by, (int)bx);
                    nb++;
                }
        if (nb && fsum / nb >= min_frac) { out[got].z0 = z0; out[got].y0 = y0; out[got].x0 = x0; got++; }
    }
    return got;
}

/* ------------------------------------------------------------------------
 * Verify (src/mc_archive.c:1072-1111)
 *
 * The index walk, its bounds checks and the per-blob structure check run
 * here as in the reference; the blobs whose structure fits are hashed on the
 * device (mchip_verify_blobs_host), and the CORRUPT lines are printed
 * afterwards in walk order.
 * ------------------------------------------------------------------------ */
typedef struct {
    int lod, n2, n1, n0;
    uint64_t co;
} verify_ent;

long mc_verify_archive(const uint8_t *arc, size_t len, int verbose) {
    if (!arc || len < MCH_HDR) return -1;
    const uint64_t NB = MCH_NODE_BYTES;
#define NODE_OK(o) ((o) >= MCH_HDR && (o) <= len && len - (o) >= NB)
    long bad = 0, total = 0;
    size_t n = 0, cap = 0;
    verify_ent *ent = NULL;
    const uint8_t **ptr = NULL;
    uint64_t *blen = NULL;
    for (int lod = 0; lod < 8; ++lod) {
        uint64_t root; memcpy(&root, arc + MCH_ROOTOFF + (uint64_t)lod * 8, 8);
        if (!root) continue;
        if (!NODE_OK(root)) { bad++; continue; }
        for (int n2 = 0; n2 < 4096; ++n2) {
            uint64_t inner; memcpy(&inner, arc + root + (size_t)n2 * 8, 8);
            if (!inner) continue;
            if (!NODE_OK(inner)) { bad++; continue; }
            for (int n1 = 0; n1 < 4096; ++n1) {
                uint64_t shard; memcpy(&shard, arc + inner + (size_t)n1 * 8, 8);
                if (!shard) continue;
                if (!NODE_OK(shard)) { bad++; continue; }
                for (int n0 = 0; n0 < 4096; ++n0) {
                    uint64_t co; memcpy(&co, arc + shard + (size_t)n0 * 8, 8);
                    if (co <= MCH_SLOT_ZERO) continue;
                    total++;
                    const uint64_t bl = mch_blob_len_bounded(arc, co, len);
                    if (!bl) { bad++; continue; }
                    if (n == cap) {
                        cap = cap ? 2 * cap : 1024;
                        verify_ent *e2 = realloc(ent, sizeof *ent * cap);
                        const uint8_t **p2 = realloc(ptr, sizeof *ptr * cap);
                        uint64_t *l2 = realloc(blen, sizeof *blen * cap);
                        if (e2) ent = e2;
                        if (p2) ptr = p2;
                        if (l2) blen = l2;
                        if (!e2 || !p2 || !l2) { bad = -1; goto out; }
                    }
                    ent[n] = (verify_ent){lod, n2, n1, n0, co};
                    ptr[n] = arc + co;
                    blen[n] = bl;
                    n++;
                }
            }
        }
    }
#undef NODE_OK
    if (n) {
        if (n > 0x7fffffff) { bad = -1; goto out; }
        mchip_ctx *d = mch_dev();
        uint8_t *flag = malloc(n);
        if (!d || !flag) { free(flag); bad = -1; goto out; }
        if (mchip_verify_blobs_host(d, ptr, blen, (int)n, NULL, flag) < 0) {
            mch_dev_error("mc_verify_archive");
            free(flag);
            bad = -1;
            goto out;
        }
        for (size_t i = 0; i < n; ++i) {
            if (!flag[i]) continue;
            bad++;
            if (verbose)
                fprintf(stderr, "mc_verify: lod %d node %d/%d/%d chunk@%llu CORRUPT\n", ent[i].lod, ent[i].n2,
                        ent[i].n1, ent[i].n0, (unsigned long long)ent[i].co);
        }
        free(flag);
    }
    if (verbose) fprintf(stderr, "mc_verify: %ld chunks checked, %ld corrupt\n", total, bad);
out:
    free(ent); free(ptr); free(blen);
    return bad;
}
