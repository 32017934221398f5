/*
 * mc_host_region.c -- region reads of the drop-in library
 * (mc_archive_read_region / mc_archive_read_regions, SuperOptimizer/
 * matter-compressor src/mc_archive.c:950-1042).
 *
 * The host plans: it walks the touched 16^3 blocks of every box, resolves
 * their chunks through the node index once per chunk, checks the bitmap and
 * builds a block map per box with each distinct present block decoded once
 * (the reference decodes a block again for every box that touches it).  The
 * GPU decodes the distinct blocks into an arena and assembles every box from
 * it; the boxes come back in one copy.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mc_host_internal.h"

typedef struct {                 /* chunk resolved during one call */
    long cz, cy, cx;
    uint64_t co;
    float q;
    uint64_t *off;               /* per block payload offset (0 = absent), lazily built */
    uint32_t *len;
} plan_chunk;

typedef struct {
    mc_archive *a;
    int lod;
    uint64_t end;
    plan_chunk *ch;
    int nch, capch;
    uint64_t *hkeys;             /* distinct-block hash: key = (chunk index << 12 | bi) + 1 */
    int32_t *hval;
    uint32_t hcap;
    const uint8_t **pay;
    uint32_t *plen;
    float *pq;
    int nblk, capblk;
} plan_t;

/* Grow *p (elements of size sz, capacity *cap) to hold at least need;
 * 0 on success, -1 when the allocation fails (*p is then unchanged). */
static int grow(void **p, int *cap, int need, size_t sz) {
    if (need <= *cap) return 0;
    int nc = *cap ? *cap : 64;
    while (nc < need) nc *= 2;
    void *t = realloc(*p, sz * (size_t)nc);
    if (!t) return -1;
    *p = t;
    *cap = nc;
    return 0;
}

/* The plan's entry of chunk (cz, cy, cx), resolved on first use; NULL when
 * memory runs out. */
static plan_chunk *plan_chunk_get(plan_t *P, long cz, long cy, long cx, int *idx) {
    for (int i = P->nch - 1; i >= 0 && i >= P->nch - 8; --i)
        if (P->ch[i].cz == cz && P->ch[i].cy == cy && P->ch[i].cx == cx) { *idx = i; return &P->ch[i]; }
    for (int i = 0; i < P->nch; ++i)
        if (P->ch[i].cz == cz && P->ch[i].cy == cy && P->ch[i].cx == cx) { *idx = i; return &P->ch[i]; }
    if (grow((void **)&P->ch, &P->capch, P->nch + 1, sizeof *P->ch)) return NULL;
    plan_chunk *c = &P->ch[P->nch];
    memset(c, 0, sizeof *c);
    c->cz = cz; c->cy = cy; c->cx = cx;
    c->co = mc_archive_chunk_offset(P->a, P->lod, (int)cz, (int)cy, (int)cx);
    if (c->co > MCH_SLOT_ZERO && c->co + 4 <= P->end) memcpy(&c->q, P->a->base + c->co, 4);
    *idx = P->nch++;
    return c;
}

/* Per-chunk (offset, length) of every block, bounded by the live cursor
 * like the reference's decode path (src/mc_archive.c:728-740).  -1 when
 * memory runs out. */
static int plan_chunk_index(plan_t *P, plan_chunk *c) {
    c->off = calloc(4096, sizeof *c->off);
    c->len = calloc(4096, sizeof *c->len);
    if (!c->off || !c->len) {
        free(c->off); free(c->len);
        c->off = NULL; c->len = NULL;
        return -1;
    }
    const uint8_t *arc = P->a->base;
    uint64_t co = c->co, end = P->end;
    if (co + MCH_BLOB_HDR > end) return 0;
    uint16_t fml; memcpy(&fml, arc + co + 12, 2);
    uint64_t bm = co + MCH_BLOB_HDR + fml;
    if (bm + 512 > end) return 0;
    int np = 0;
    for (int i = 0; i < 512; ++i) np += __builtin_popcount(arc[bm + i]);
    uint64_t lens = bm + 512, pay = lens + 2ull * np;
    if (pay > end) return 0;
    int s = 0;
    for (int bi = 0; bi < 4096; ++bi)
        if ((arc[bm + (bi >> 3)] >> (bi & 7)) & 1) {
            uint16_t l; memcpy(&l, arc + lens + 2 * s, 2);
            if (pay + l <= end) { c->off[bi] = pay; c->len[bi] = l; }
            pay += l;
            s++;
        }
    return 0;
}

/* Arena slot of block bi of plan chunk ci: the block's index among the
 * distinct present blocks, -1 for an absent block, -2 when memory runs out. */
static int plan_block(plan_t *P, int ci, int bi) {
    uint64_t key = ((uint64_t)ci << 12 | (uint64_t)bi) + 1;
    if ((uint32_t)P->nblk * 2 + 2 > P->hcap) {            /* grow + rehash */
        uint32_t nc = P->hcap ? P->hcap * 2 : 1u << 14;
        uint64_t *nk = calloc(nc, sizeof *nk);
        int32_t *nv = malloc(nc * sizeof *nv);
        if (!nk || !nv) { free(nk); free(nv); return -2; }
        for (uint32_t i = 0; i < P->hcap; ++i)
            if (P->hkeys[i]) {
                uint32_t h = (uint32_t)((P->hkeys[i] * 0x9E3779B97F4A7C15ull) >> 40) & (nc - 1);
                while (nk[h]) h = (h + 1) & (nc - 1);
                nk[h] = P->hkeys[i];
                nv[h] = P->hval[i];
            }
        free(P->hkeys); free(P->hval);
        P->hkeys = nk; P->hval = nv; P->hcap = nc;
    }
    uint32_t h = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 40) & (P->hcap - 1);
    while (P->hkeys[h]) {
        if (P->hkeys[h] == key) return P->hval[h];
        h = (h + 1) & (P->hcap - 1);
    }
    plan_chunk *c = &P->ch[ci];
    if (!c->off && plan_chunk_index(P, c)) return -2;
    if (!c->off[bi]) { P->hkeys[h] = key; P->hval[h] = -1; return -1; }
    if (P->nblk == P->capblk) {
        int cp = P->capblk, cl = P->capblk, cq = P->capblk;
        if (grow((void **)&P->pay, &cp, P->nblk + 1, sizeof *P->pay) ||
            grow((void **)&P->plen, &cl, P->nblk + 1, sizeof *P->plen) ||
            grow((void **)&P->pq, &cq, P->nblk + 1, sizeof *P->pq))
            return -2;
        P->capblk = cp < cl ? (cp < cq ? cp : cq) : (cl < cq ? cl : cq);
    }
    int slot = P->nblk++;
    P->pay[slot] = P->a->base + c->off[bi];
    P->plen[slot] = c->len[bi];
    P->pq[slot] = c->q;
    P->hkeys[h] = key;
    P->hval[h] = slot;
    return slot;
}

static void plan_free(plan_t *P) {
    for (int i = 0; i < P->nch; ++i) { free(P->ch[i].off); free(P->ch[i].len); }
    free(P->ch); free(P->hkeys); free(P->hval); free(P->pay); free(P->plen); free(P->pq);
}

static void zero_boxes(mc_u8 *out, int n, size_t batch_stride, size_t box) {
    for (int b = 0; b < n; ++b) memset(out + (size_t)b * batch_stride, 0, box);
}

/* Read n boxes (origins in boxes) into out + i*batch_stride (dense).  When
 * host memory or the device fails, the boxes read as zeros and the failure
 * is reported on stderr. */
static void read_boxes(mc_archive *a, int lod, const long *org, int n, long dz, long dy, long dx, mc_u8 *out,
                       size_t batch_stride) {
    const int gz = (int)((dz + 15) >> 4) + 1, gy = (int)((dy + 15) >> 4) + 1, gx = (int)((dx + 15) >> 4) + 1;
    const size_t cells = (size_t)gz * gy * gx;
    const size_t box = (size_t)(dz * dy * dx);
    int32_t *bmap = malloc(sizeof(int32_t) * cells * n);
    int64_t *origin = malloc(sizeof(int64_t) * 3 * n);
    plan_t P;
    memset(&P, 0, sizeof P);
    P.a = a; P.lod = lod;
    P.end = atomic_load_explicit(&a->cursor, memory_order_acquire);
    int oom = !bmap || !origin;
    for (int b = 0; b < n && !oom; ++b) {
        const long z0 = org[3 * b], y0 = org[3 * b + 1], x0 = org[3 * b + 2];
        origin[3 * b] = z0; origin[3 * b + 1] = y0; origin[3 * b + 2] = x0;
        int32_t *bm = bmap + cells * b;
        for (size_t k = 0; k < cells; ++k) bm[k] = -1;
        const long bz0 = z0 >> 4, by0 = y0 >> 4, bx0 = x0 >> 4;
        const long nbz = ((z0 + dz + 15) >> 4) - bz0, nby = ((y0 + dy + 15) >> 4) - by0,
                   nbx = ((x0 + dx + 15) >> 4) - bx0;
        for (long iz = 0; iz < nbz && !oom; ++iz)
            for (long iy = 0; iy < nby && !oom; ++iy)
                for (long ix = 0; ix < nbx; ++ix) {
                    const long bz = bz0 + iz, by = by0 + iy, bx = bx0 + ix;
                    int ci;
                    plan_chunk *c = plan_chunk_get(&P, bz >> 4, by >> 4, bx >> 4, &ci);
                    if (!c) { oom = 1; break; }
                    if (c->co <= MCH_SLOT_ZERO) continue;
                    const int bi = (int)(((bz & 15) * 16 + (by & 15)) * 16 + (bx & 15));
                    const int slot = plan_block(&P, ci, bi);
                    if (slot == -2) { oom = 1; break; }
                    bm[(iz * gy + iy) * gx + ix] = slot;
                }
    }
    if (oom) {
        fprintf(stderr, "matter_compressor_amd: region read: out of host memory\n");
        zero_boxes(out, n, batch_stride, box);
    } else {
        mchip_ctx *d = mch_dev();
        if (!d || mchip_read_regions_host(d, P.nblk, P.pay, P.plen, P.pq, bmap, origin, n, gz, gy, gx, dz, dy, dx,
                                          out, batch_stride) != 0) {
            mch_dev_error("region read");
            zero_boxes(out, n, batch_stride, box);
        }
    }
    plan_free(&P);
    free(bmap);
    free(origin);
}

void mc_archive_read_regions(mc_archive *a, int lod, const mc_box *boxes, int n, long dz, long dy, long dx,
                             mc_u8 *out, size_t batch_stride, int nthreads) {
    (void)nthreads;
    if (!a || !boxes || n <= 0 || !out) return;
    if (lod < 0 || lod > 7 || dz <= 0 || dy <= 0 || dx <= 0) return;   /* read_region's own guard */
    long *org = malloc(sizeof(long) * 3 * n);
    if (!org) {
        fprintf(stderr, "matter_compressor_amd: mc_archive_read_regions: out of host memory\n");
        zero_boxes(out, n, batch_stride, (size_t)(dz * dy * dx));
        return;
    }
    for (int i = 0; i < n; ++i) { org[3 * i] = boxes[i].z0; org[3 * i + 1] = boxes[i].y0; org[3 * i + 2] = boxes[i].x0; }
    read_boxes(a, lod, org, n, dz, dy, dx, out, batch_stride);
    free(org);
}

void mc_archive_read_region(mc_archive *a, int lod, long z0, long y0, long x0, long dz, long dy, long dx, mc_u8 *out,
                            size_t sz, size_t sy, int nthreads) {
    (void)nthreads;
    if (!a || lod < 0 || lod > 7 || !out || dz <= 0 || dy <= 0 || dx <= 0) return;
    long org[3] = {z0, y0, x0};
    if (sy == (size_t)dx && sz == (size_t)(dy * dx)) {
        read_boxes(a, lod, org, 1, dz, dy, dx, out, (size_t)(dz * dy * dx));
        return;
    }
    mc_u8 *tmp = malloc((size_t)(dz * dy * dx));
    if (!tmp) {
        fprintf(stderr, "matter_compressor_amd: mc_archive_read_region: out of host memory\n");
        for (long z = 0; z < dz; ++z)
            for (long y = 0; y < dy; ++y) memset(out + (size_t)z * sz + (size_t)y * sy, 0, (size_t)dx);
        return;
    }
    read_boxes(a, lod, org, 1, dz, dy, dx, tmp, (size_t)(dz * dy * dx));
    for (long z = 0; z < dz; ++z)
        for (long y = 0; y < dy; ++y)
            memcpy(out + (size_t)z * sz + (size_t)y * sy, tmp + ((size_t)z * dy + y) * dx, (size_t)dx);
    free(tmp);
}
