/*
 * mc_host_stream.c -- the read-only archive reader of the drop-in API
 * (src/matter_compressor.h:313-361, behaviour of src/mc_stream.c): flat
 * readers over an archive in memory and streaming readers over a byte-range
 * callback, with whole-blob or partial fetch, chunk resolution through a
 * node-table cache and the blob accessors.  Block decodes run on the GPU.
 *
 * mc_reader_decode_blocks (additive) feeds the device in bulk: it gathers
 * the payloads of many blocks through the reader (flat pointers, fetched
 * chunk blobs, or per-block range reads in partial mode) and decodes them
 * in one launch.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mc_host_internal.h"

#define RD_NODE_CACHE 512           /* cached 32 KiB node tables per streaming reader */
#define RD_HDR_CAP (MCH_BLOB_HDR + 512 + 4096 * 2)

struct mc_reader {
    const uint8_t *arc;             /* flat: the archive bytes; streaming: NULL */
    uint64_t len;
    uint64_t roots[8];
    int nx, ny, nz;
    float quality;
    uint8_t priors[MCH_PRIORS_BYTES];
    int has_priors;
    mc_read_fn read;
    void *read_ud;
    /* streaming, whole-blob mode: the last fetched chunk blob */
    uint8_t *cbuf;
    uint64_t cbuf_off, cbuf_len;
    /* streaming, partial mode: the last chunk's header, bitmap and lengths */
    int partial;
    uint8_t hdr[RD_HDR_CAP];
    uint64_t hdr_off;
    int hdr_np;
    uint16_t hdr_fml;
    /* FIFO of fetched node tables */
    uint64_t ntab_off[RD_NODE_CACHE];
    uint8_t *ntab[RD_NODE_CACHE];
    int ntab_next;
};

static int rd(mc_reader *r, uint64_t off, uint32_t n, uint8_t *dst) { return r->read(r->read_ud, off, n, dst); }

/* the per-chunk quality of an untrusted blob (src/mc_stream.c:265-266) */
static float reader_q(float q) { return (!(q > 0.0f) || q > 1024.0f) ? 1.0f : q; }

static void load_header(mc_reader *r, const uint8_t *h) {
    for (int l = 0; l < 8; ++l) memcpy(&r->roots[l], h + MCH_ROOTOFF + l * 8, 8);
    uint32_t ux, uy, uz;
    memcpy(&ux, h + 12, 4); memcpy(&uy, h + 16, 4); memcpy(&uz, h + 20, 4);
    r->nx = (int)ux; r->ny = (int)uy; r->nz = (int)uz;
    memcpy(&r->quality, h + 120, 4);
}

mc_reader *mc_open(const uint8_t *arc, size_t len) {
    if (!arc || len < MCH_HDR) return NULL;
    uint32_t magic;
    memcpy(&magic, arc, 4);
    if (magic != MCH_MAGIC) return NULL;
    mc_codec_init();
    mc_reader *r = calloc(1, sizeof *r);
    if (!r) return NULL;
    r->arc = arc;
    r->len = len;
    r->hdr_off = ~0ull;
    load_header(r, arc);
    uint64_t poff;
    memcpy(&poff, arc + MCH_PRIOROFF, 8);
    if (poff && poff <= len && len - poff >= (uint64_t)MCH_PRIORS_BYTES) {
        uint32_t pm;
        memcpy(&pm, arc + poff, 4);
        if (pm == MCH_PRIORS_MAGIC) {
            memcpy(r->priors, arc + poff, MCH_PRIORS_BYTES);
            r->has_priors = 1;
            priors_load(arc);
        }
    }
    return r;
}

mc_reader *mc_open_streaming(mc_read_fn read, void *ud, uint64_t total_len) {
    if (!read) return NULL;
    mc_codec_init();
    mc_reader *r = calloc(1, sizeof *r);
    if (!r) return NULL;
    r->read = read;
    r->read_ud = ud;
    r->len = total_len;
    r->hdr_off = ~0ull;
    uint8_t h[MCH_HDR];
    uint32_t magic = 0;
    if (read(ud, 0, MCH_HDR, h) != 0 || (memcpy(&magic, h, 4), magic != MCH_MAGIC)) {
        free(r);
        return NULL;
    }
    load_header(r, h);
    uint64_t poff;
    memcpy(&poff, h + MCH_PRIOROFF, 8);
    uint32_t pm = 0;
    if (poff && read(ud, poff, MCH_PRIORS_BYTES, r->priors) == 0) memcpy(&pm, r->priors, 4);
    if (pm == MCH_PRIORS_MAGIC) {
        r->has_priors = 1;
        mc_codec_set_priors((const uint16_t *)(r->priors + 8), (const uint16_t *)(r->priors + 8 + 512));
    } else {
        mc_codec_set_priors(NULL, NULL);
    }
    return r;
}

void mc_close(mc_reader *r) {
    if (!r) return;
    for (int i = 0; i < RD_NODE_CACHE; ++i) free(r->ntab[i]);
    free(r->cbuf);
    free(r);
}

void mc_reader_set_quality(mc_reader *r, float q) { (void)r; (void)q; /* each chunk's own q decodes it */ }
void mc_reader_set_partial_fetch(mc_reader *r, int on) {
    if (r) { r->partial = on; r->hdr_off = ~0ull; }
}
void mc_reader_dims(mc_reader *r, int *nx, int *ny, int *nz) {
    if (nx) *nx = r ? r->nx : 0;
    if (ny) *ny = r ? r->ny : 0;
    if (nz) *nz = r ? r->nz : 0;
}
float mc_reader_quality(mc_reader *r) { return r ? r->quality : 0.0f; }
int mc_reader_nlods(mc_reader *r) {
    if (!r) return 0;
    int n = 0;
    for (int l = 0; l < 8; ++l) if (r->roots[l]) n = l + 1;
    return n;
}
int mc_reader_priors(mc_reader *r, const uint16_t **plo, const uint16_t **phi) {
    if (!r || !r->has_priors) return 0;
    if (plo) *plo = (const uint16_t *)(r->priors + 8);
    if (phi) *phi = (const uint16_t *)(r->priors + 8 + 512);
    return 1;
}

/* ---- chunk resolution ---------------------------------------------------- */
static const uint8_t *node_table(mc_reader *r, uint64_t off) {
    for (int i = 0; i < RD_NODE_CACHE; ++i)
        if (r->ntab[i] && r->ntab_off[i] == off) return r->ntab[i];
    const int s = r->ntab_next;
    r->ntab_next = (s + 1) % RD_NODE_CACHE;
    if (!r->ntab[s]) r->ntab[s] = malloc(MCH_NODE_BYTES);
    if (!r->ntab[s] || rd(r, off, MCH_NODE_BYTES, r->ntab[s]) != 0) {
        free(r->ntab[s]);
        r->ntab[s] = NULL;
        return NULL;
    }
    r->ntab_off[s] = off;
    return r->ntab[s];
}

uint64_t mc_chunk_offset_chk(mc_reader *r, int lod, int cz, int cy, int cx, int *err) {
    if (err) *err = 0;
    if (!r || lod < 0 || lod > 7) return 0;
    if (r->arc) return mch_resolve(r->arc, r->roots[lod], cz, cy, cx, r->len);
    uint64_t node = r->roots[lod];
    for (int lv = 2; lv >= 0; --lv) {
        if (!node) return 0;                         /* absent */
        const uint8_t *t = node_table(r, node);
        if (!t) { if (err) *err = 1; return 0; }     /* the read failed: not known to be absent */
        const int idx = ((((cz >> (4 * lv)) & 15) * 16 + ((cy >> (4 * lv)) & 15)) * 16) + ((cx >> (4 * lv)) & 15);
        memcpy(&node, t + (size_t)idx * 8, 8);
    }
    return node;
}

uint64_t mc_chunk_offset(mc_reader *r, int lod, int cz, int cy, int cx) {
    return mc_chunk_offset_chk(r, lod, cz, cy, cx, NULL);
}

/* ---- blob accessors --------------------------------------------------------- */
uint64_t mc_reader_chunk_blob_len(mc_reader *r, uint64_t chunk_off) {
    if (!r || !chunk_off) return 0;
    if (r->arc) return mch_blob_len_bounded(r->arc, chunk_off, (uint64_t)-1);
    uint8_t h[MCH_BLOB_HDR], bm[512];
    if (rd(r, chunk_off, MCH_BLOB_HDR, h) != 0) return 0;
    uint16_t fml;
    memcpy(&fml, h + 12, 2);
    const uint64_t bmo = chunk_off + MCH_BLOB_HDR + fml;
    if (rd(r, bmo, 512, bm) != 0) return 0;
    int np = 0;
    for (int i = 0; i < 512; ++i) np += __builtin_popcount(bm[i]);
    if (!np) return bmo + 512 - chunk_off;
    uint8_t *lens = malloc((size_t)np * 2);
    if (!lens || rd(r, bmo + 512, (uint32_t)(np * 2), lens) != 0) { free(lens); return 0; }
    uint64_t pay = 0;
    for (int i = 0; i < np; ++i) { uint16_t l; memcpy(&l, lens + 2 * i, 2); pay += l; }
    free(lens);
    return bmo + 512 + 2ull * np + pay - chunk_off;
}

int mc_reader_read_blob(mc_reader *r, uint64_t chunk_off, size_t len, uint8_t *dst) {
    if (!r || !chunk_off || !len || !dst) return -1;
    if (r->arc) {                /* a range past the archive's end is refused, not read */
        if (chunk_off > r->len || len > r->len - chunk_off) return -1;
        memcpy(dst, r->arc + chunk_off, len);
        return 0;
    }
    for (size_t done = 0; done < len;) {
        const uint32_t n = len - done > 0x40000000u ? 0x40000000u : (uint32_t)(len - done);
        if (rd(r, chunk_off + done, n, dst + done) != 0) return -1;
        done += n;
    }
    return 0;
}

/* ---- locating one block's payload ------------------------------------------ */
/* Payload of block bi of the blob whose first byte is b (bytes [0, end) of it
 * readable): 1 and (offset in b, length), or 0 for an absent block. */
static int blob_block(const uint8_t *b, uint64_t end, int bi, uint64_t *off, uint32_t *len) {
    uint16_t fml;
    memcpy(&fml, b + 12, 2);
    const uint64_t bm = MCH_BLOB_HDR + (uint64_t)fml;
    if (!((b[bm + (bi >> 3)] >> (bi & 7)) & 1)) return 0;
    int np = 0, rank = 0;
    for (int i = 0; i < 512; ++i) {
        const int c = __builtin_popcount(b[bm + i]);
        np += c;
        if (i < (bi >> 3)) rank += c;
    }
    rank += __builtin_popcount(b[bm + (bi >> 3)] & ((1u << (bi & 7)) - 1u));
    uint64_t pay = bm + 512 + 2ull * np;
    for (int s = 0; s < rank; ++s) { uint16_t l; memcpy(&l, b + bm + 512 + 2 * s, 2); pay += l; }
    uint16_t l;
    memcpy(&l, b + bm + 512 + 2 * rank, 2);
    if (pay + l > end) return 0;
    *off = pay;
    *len = l;
    return 1;
}

/* Streaming whole-blob fetch: header, bitmap and lengths give the blob's
 * length, then the blob is read in one request (src/mc_stream.c:234-255). */
static const uint8_t *fetch_blob(mc_reader *r, uint64_t co, uint64_t *blen) {
    if (r->cbuf && r->cbuf_off == co) { *blen = r->cbuf_len; return r->cbuf; }
    const uint64_t total = mc_reader_chunk_blob_len(r, co);
    if (!total || total > 0xFFFFFFFFull) return NULL;
    uint8_t *t = realloc(r->cbuf, total);
    if (!t) return NULL;
    r->cbuf = t;
    r->cbuf_off = ~0ull;
    if (rd(r, co, (uint32_t)total, r->cbuf) != 0) return NULL;
    r->cbuf_off = co;
    r->cbuf_len = total;
    *blen = total;
    return r->cbuf;
}

/* Streaming partial fetch: the chunk's header, bitmap and lengths (cached
 * per chunk), then only the block's payload (src/mc_stream.c:258-286).
 * Returns 1 with the payload in *buf (malloc'd, grown as needed), 0 for an
 * absent block, -1 on a failed read. */
static int fetch_partial(mc_reader *r, uint64_t co, int bi, float *q, uint8_t **buf, size_t *cap, uint32_t *len) {
    if (r->hdr_off != co) {
        r->hdr_off = ~0ull;
        if (rd(r, co, MCH_BLOB_HDR, r->hdr) != 0) return -1;
        uint16_t fml;
        memcpy(&fml, r->hdr + 12, 2);
        r->hdr_fml = fml;
        if (rd(r, co + MCH_BLOB_HDR + fml, 512, r->hdr + MCH_BLOB_HDR) != 0) return -1;
        int np = 0;
        for (int i = 0; i < 512; ++i) np += __builtin_popcount(r->hdr[MCH_BLOB_HDR + i]);
        if (np && rd(r, co + MCH_BLOB_HDR + fml + 512, (uint32_t)(np * 2), r->hdr + MCH_BLOB_HDR + 512) != 0)
            return -1;
        r->hdr_np = np;
        r->hdr_off = co;
    }
    float cq;
    memcpy(&cq, r->hdr, 4);
    *q = reader_q(cq);
    const uint8_t *bm = r->hdr + MCH_BLOB_HDR;
    if (!((bm[bi >> 3] >> (bi & 7)) & 1)) return 0;
    int rank = 0;
    for (int i = 0; i < (bi >> 3); ++i) rank += __builtin_popcount(bm[i]);
    rank += __builtin_popcount(bm[bi >> 3] & ((1u << (bi & 7)) - 1u));
    const uint8_t *lens = bm + 512;
    uint64_t cum = 0;
    for (int s = 0; s < rank; ++s) { uint16_t l; memcpy(&l, lens + 2 * s, 2); cum += l; }
    uint16_t l;
    memcpy(&l, lens + 2 * rank, 2);
    if (*cap < (size_t)l + 1) {
        uint8_t *t = realloc(*buf, (size_t)l + 1);
        if (!t) return -1;
        *buf = t;
        *cap = (size_t)l + 1;
    }
    const uint64_t pay = co + MCH_BLOB_HDR + r->hdr_fml + 512 + 2ull * r->hdr_np + cum;
    if (l && rd(r, pay, l, *buf) != 0) return -1;
    *len = l;
    return 1;
}

/* ---- decode ------------------------------------------------------------------ */
/* Payload, length and quality of one block for a batch; NULL payload: the
 * block reads as zeros.  Streaming payloads are copied into store[i]. */
typedef struct {
    const uint8_t *p;
    uint32_t len;
    float q;
} blk_src;

static void locate(mc_reader *r, uint64_t co, int bz, int by, int bx, blk_src *out, uint8_t **own, size_t *own_cap) {
    out->p = NULL;
    out->len = 0;
    out->q = 1.0f;
    if (co <= MCH_SLOT_ZERO || bz < 0 || by < 0 || bx < 0 || bz > 15 || by > 15 || bx > 15) return;
    const int bi = (bz * 16 + by) * 16 + bx;
    if (r->arc) {
        const uint64_t total = mch_blob_len_bounded(r->arc, co, r->len);
        if (!total) return;                          /* the blob's structure does not fit the archive */
        uint64_t o;
        uint32_t l;
        if (!blob_block(r->arc + co, total, bi, &o, &l)) return;
        float cq;
        memcpy(&cq, r->arc + co, 4);
        out->p = r->arc + co + o;
        out->len = l;
        out->q = reader_q(cq);
        return;
    }
    if (r->partial) {
        float q;
        uint32_t l = 0;
        if (fetch_partial(r, co, bi, &q, own, own_cap, &l) != 1) return;
        out->p = *own;
        out->len = l;
        out->q = q;
        return;
    }
    uint64_t blen;
    const uint8_t *b = fetch_blob(r, co, &blen);
    if (!b) return;
    uint64_t o;
    uint32_t l;
    if (!blob_block(b, blen, bi, &o, &l)) return;
    if (*own_cap < (size_t)l + 1) {
        uint8_t *t = realloc(*own, (size_t)l + 1);
        if (!t) return;
        *own = t;
        *own_cap = (size_t)l + 1;
    }
    memcpy(*own, b + o, l);
    float cq;
    memcpy(&cq, b, 4);
    out->p = *own;
    out->len = l;
    out->q = reader_q(cq);
}

int mc_reader_decode_blocks(mc_reader *r, int n, const uint64_t *chunk_off, const int32_t *bzyx, mc_u8 *dst) {
    if (!r || n < 0 || (n && (!chunk_off || !bzyx || !dst))) return -1;
    if (!n) return 0;
    blk_src *src = calloc((size_t)n, sizeof *src);
    uint8_t **own = calloc((size_t)n, sizeof *own);
    size_t *cap = calloc((size_t)n, sizeof *cap);
    const uint8_t **pp = calloc((size_t)n, sizeof *pp);
    uint32_t *ll = calloc((size_t)n, sizeof *ll);
    float *qq = calloc((size_t)n, sizeof *qq);
    int *idx = calloc((size_t)n, sizeof *idx);
    int rc = -1, m = 0;
    if (!src || !own || !cap || !pp || !ll || !qq || !idx) goto out;
    for (int i = 0; i < n; ++i) {
        locate(r, chunk_off[i], bzyx[3 * i], bzyx[3 * i + 1], bzyx[3 * i + 2], &src[i], &own[i], &cap[i]);
        if (!src[i].p) { memset(dst + (size_t)i * 4096, 0, 4096); continue; }
        pp[m] = src[i].p;
        ll[m] = src[i].len;
        qq[m] = src[i].q;
        idx[m++] = i;
    }
    if (m) {
        mchip_ctx *d = mch_dev();
        uint8_t *tmp = malloc((size_t)m * 4096);
        if (!d || !tmp || mchip_decode_blocks_host(d, pp, ll, qq, m, tmp) != 0) {
            mch_dev_error("mc_reader_decode_blocks");
            free(tmp);
            for (int k = 0; k < m; ++k) memset(dst + (size_t)idx[k] * 4096, 0, 4096);
            goto out;
        }
        for (int k = 0; k < m; ++k) memcpy(dst + (size_t)idx[k] * 4096, tmp + (size_t)k * 4096, 4096);
        free(tmp);
    }
    rc = 0;
out:
    if (own) for (int i = 0; i < n; ++i) free(own[i]);
    free(src); free(own); free(cap); free(pp); free(ll); free(qq); free(idx);
    return rc;
}

void mc_decode_block(mc_reader *r, uint64_t chunk_off, int bz, int by, int bx, mc_u8 *dst) {
    if (!dst) return;
    if (!r) { memset(dst, 0, 4096); return; }
    const int32_t c[3] = {bz, by, bx};
    if (mc_reader_decode_blocks(r, 1, &chunk_off, c, dst) != 0) memset(dst, 0, 4096);
}
