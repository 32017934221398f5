/*
 * mc_host_serve.c -- HBM-resident serving of one LOD of an archive.
 *
 * mc_serve_open walks the chunk index of one LOD, mc_serve_open_levels of
 * several consecutive LODs (src/mc_archive.h:133-145); every present chunk
 * blob is uploaded once and the device is given the payload range and
 * quality of every block of each level's chunk bounding box.  Requests
 * (boxes of the first level, or quad renders over the levels) then run
 * through the device serving loop (mchip_server_*), which decodes each block
 * at most once while it stays in the HBM arena.  The voxels of every box
 * equal mc_archive_read_regions' for the same box (absent chunks and blocks,
 * and everything outside the LOD's chunks, read as 0).
 *
 * Coherence with appends (the reference's mc_cache re-resolves when the
 * archive's generation changes and drops re-appended chunks,
 * src/mc_cache.c:690-697, :722-736): every submit compares the archive's
 * generation with the one last seen; after a change it walks the levels'
 * indexes again and replaces the device payloads of every chunk whose offset
 * changed (mc_serve_refresh), and mc_serve_invalidate_chunk does that for
 * one chunk on request.  Chunks outside the grid a level had at open are not
 * seen (mc_serve_refresh counts them).
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mc_host_internal.h"

typedef struct {                 /* one level of the store */
    int lo[3];                   /* first chunk coordinate of the level's grid */
    int nc[3];                   /* grid size in chunks */
    int gz, gy, gx;              /* grid size in blocks */
    uint64_t base;               /* store index of the level's first block */
    uint64_t *co;                /* chunk offset last uploaded, per grid chunk */
} serve_grid;

struct mc_serve {
    mchip_store *st;
    mchip_server *sv;
    long oz, oy, ox;             /* voxel origin of the store grid */
    mc_archive *a;
    int lod0, nlods;
    serve_grid g[8];
    _Atomic uint64_t gen;        /* archive generation of the last refresh */
    pthread_mutex_t mu;          /* refreshes */
};

typedef struct {
    int cz, cy, cx;
    uint64_t co;
} serve_chunk;

typedef struct {
    serve_chunk *v;
    int n, cap;
    int oom;                     /* an allocation failed: the list is incomplete */
} chunk_list;

static void walk(const uint8_t *arc, uint64_t end, uint64_t node, int lv, int cz, int cy, int cx, chunk_list *out) {
    if (!node || node > end || MCH_NODE_BYTES > end - node) return;
    for (int idx = 0; idx < 4096 && !out->oom; ++idx) {
        uint64_t child;
        memcpy(&child, arc + node + (uint64_t)idx * 8, 8);
        if (!child) continue;
        const int z = (cz << 4) | (idx >> 8), y = (cy << 4) | ((idx >> 4) & 15), x = (cx << 4) | (idx & 15);
        if (lv > 0) {
            walk(arc, end, child, lv - 1, z, y, x, out);
        } else if (child > MCH_SLOT_ZERO) {
            if (out->n == out->cap) {
                const int nc = out->cap ? out->cap * 2 : 256;
                serve_chunk *t = realloc(out->v, sizeof *out->v * nc);
                if (!t) { out->oom = 1; return; }
                out->v = t;
                out->cap = nc;
            }
            out->v[out->n++] = (serve_chunk){z, y, x, child};
        }
    }
}

static void serve_free_grids(mc_serve *s) {
    for (int l = 0; l < 8; ++l) {
        free(s->g[l].co);
        s->g[l].co = NULL;
    }
}

/* Upload the blocks of chunk (cz, cy, cx) of level l as the blob at co (0:
 * absent) holds them, replacing the store's.  1 when the chunk lies outside
 * the level's grid, -1 on failure, else 0. */
static int serve_update_chunk(mc_serve *s, int l, int cz, int cy, int cx, uint64_t co) {
    serve_grid *G = &s->g[l];
    const int rz = cz - G->lo[0], ry = cy - G->lo[1], rx = cx - G->lo[2];
    if (rz < 0 || ry < 0 || rx < 0 || rz >= G->nc[0] || ry >= G->nc[1] || rx >= G->nc[2]) return 1;
    const uint8_t *arc = s->a->base;
    const uint64_t end = atomic_load_explicit(&s->a->cursor, memory_order_acquire);
    uint64_t *ids = malloc(4096 * sizeof *ids), *off = calloc(4096, sizeof *off);
    uint32_t *len = calloc(4096, sizeof *len);
    float *q = calloc(4096, sizeof *q);
    uint8_t *pay = NULL;
    uint64_t npay = 0;
    int rc = -1;
    if (!ids || !off || !len || !q) goto out;
    for (int bi = 0; bi < 4096; ++bi) {
        const int bz = 16 * rz + (bi >> 8), by = 16 * ry + ((bi >> 4) & 15), bx = 16 * rx + (bi & 15);
        ids[bi] = G->base + ((uint64_t)bz * G->gy + by) * G->gx + bx;
    }
    if (co > MCH_SLOT_ZERO && co + MCH_BLOB_HDR <= end) {
        float cq;
        memcpy(&cq, arc + co, 4);
        uint16_t fml;
        memcpy(&fml, arc + co + 12, 2);
        const uint64_t bm = co + MCH_BLOB_HDR + (uint64_t)fml;
        if (bm + 512 <= end) {
            int np = 0;
            for (int k = 0; k < 512; ++k) np += __builtin_popcount(arc[bm + k]);
            const uint64_t lens = bm + 512, pay0 = lens + 2ull * np;
            if (pay0 <= end) {
                /* the payloads are consecutive: upload them as one range */
                uint64_t p = pay0, ext = pay0;
                int rank = 0;
                for (int bi = 0; bi < 4096; ++bi) {
                    if (!((arc[bm + (bi >> 3)] >> (bi & 7)) & 1)) continue;
                    uint16_t bl;
                    memcpy(&bl, arc + lens + 2ull * rank, 2);
                    rank++;
                    if (bl && p + bl <= end) {
                        off[bi] = p - pay0;
                        len[bi] = bl;
                        q[bi] = cq;
                        if (p + bl > ext) ext = p + bl;
                    }
                    p += bl;
                }
                pay = (uint8_t *)(arc + pay0);
                npay = ext - pay0;
            }
        }
    }
    if (mchip_server_update_blocks(s->sv, pay, npay, 4096, ids, off, len, q) != 0) {
        mch_dev_error("mc_serve chunk update");
        goto out;
    }
    G->co[((size_t)rz * G->nc[1] + ry) * G->nc[2] + rx] = co;
    rc = 0;
out:
    free(ids); free(off); free(len); free(q);
    return rc;
}

/* Per-level chunk list, bounding box and block grid. */
typedef struct {
    chunk_list cl;
    int lo[3], hi[3];
    int gz, gy, gx;
    uint64_t base;                   /* store index of the level's first block */
} serve_level;

mc_serve *mc_serve_open_levels(mc_archive *a, int lod0, int nlods, const int32_t *dims, uint64_t arena_bytes,
                               int lanes) {
    if (!a || lod0 < 0 || nlods < 1 || lod0 + nlods > 8 || lanes <= 0) return NULL;
    const uint8_t *arc = a->base;
    const uint64_t gen0 = atomic_load_explicit(&a->gen, memory_order_acquire);
    const uint64_t end = atomic_load_explicit(&a->cursor, memory_order_acquire);
    serve_level lv[8];
    memset(lv, 0, sizeof lv);
    uint64_t nb = 0;
    int ok = 1;
    for (int l = 0; l < nlods; ++l) {
        serve_level *L = &lv[l];
        uint64_t root;
        memcpy(&root, arc + MCH_ROOTOFF + (uint64_t)(lod0 + l) * 8, 8);
        walk(arc, end, root, 2, 0, 0, 0, &L->cl);
        if (L->cl.oom) {
            fprintf(stderr, "mc_serve_open: out of host memory\n");
            ok = 0;
        }
        for (int k = 0; k < 3; ++k) { L->lo[k] = 4096; L->hi[k] = -1; }
        for (int i = 0; i < L->cl.n; ++i) {
            const int c[3] = {L->cl.v[i].cz, L->cl.v[i].cy, L->cl.v[i].cx};
            for (int k = 0; k < 3; ++k) {
                if (c[k] < L->lo[k]) L->lo[k] = c[k];
                if (c[k] > L->hi[k]) L->hi[k] = c[k];
            }
        }
        if (L->cl.n == 0) for (int k = 0; k < 3; ++k) { L->lo[k] = 0; L->hi[k] = 0; }
        L->gz = 16 * (L->hi[0] - L->lo[0] + 1);
        L->gy = 16 * (L->hi[1] - L->lo[1] + 1);
        L->gx = 16 * (L->hi[2] - L->lo[2] + 1);
        L->base = nb;
        nb += (uint64_t)L->gz * L->gy * L->gx;
    }
    if (nb > (1ull << 28)) {                          /* each level's grid is its chunks' bounding box */
        fprintf(stderr, "mc_serve_open: chunk bounding boxes of %llu blocks are too large\n", (unsigned long long)nb);
        ok = 0;
    }
    int nchunks = 0;
    for (int l = 0; l < nlods; ++l) nchunks += lv[l].cl.n;
    uint64_t *off = ok ? calloc(nb, sizeof *off) : NULL;
    uint32_t *len = ok ? calloc(nb, sizeof *len) : NULL;
    float *q = ok ? calloc(nb, sizeof *q) : NULL;
    uint64_t total = 0;
    uint64_t *blob_at = malloc(sizeof *blob_at * (nchunks ? nchunks : 1));
    uint64_t *blob_len = malloc(sizeof *blob_len * (nchunks ? nchunks : 1));
    uint8_t *stage = NULL;
    mc_serve *s = NULL;
    if (!ok || !off || !len || !q || !blob_at || !blob_len) goto done;
    /* Pass 1: per-block payload ranges, each bounded by the archive's live
     * end like mc_archive_read_regions (src/mc_archive.c:728-740); a chunk's
     * upload extent is the end of its last valid payload.  Pass 2 fills the
     * per-block records and the upload. */
    for (int pass = 0; pass < 2; ++pass) {
        int ci = 0;
        for (int l = 0; l < nlods; ++l) {
            const serve_level *L = &lv[l];
            for (int i = 0; i < L->cl.n; ++i, ++ci) {
                const uint64_t co = L->cl.v[i].co;
                uint64_t ext = 0;
                if (co + MCH_BLOB_HDR <= end) {
                    float cq;
                    memcpy(&cq, arc + co, 4);
                    uint16_t fml;
                    memcpy(&fml, arc + co + 12, 2);
                    const uint64_t bm = co + MCH_BLOB_HDR + (uint64_t)fml;
                    if (bm + 512 <= end) {
                        int np = 0;
                        for (int k = 0; k < 512; ++k) np += __builtin_popcount(arc[bm + k]);
                        const uint64_t lens = bm + 512;
                        uint64_t pay = lens + 2ull * np;
                        if (pay <= end) {
                            int rank = 0;
                            for (int bi = 0; bi < 4096; ++bi) {
                                if (!((arc[bm + (bi >> 3)] >> (bi & 7)) & 1)) continue;
                                uint16_t bl;
                                memcpy(&bl, arc + lens + 2ull * rank, 2);
                                rank++;
                                if (pay + bl <= end && bl) {
                                    if (pay + bl - co > ext) ext = pay + bl - co;
                                    if (pass) {
                                        const int bz = 16 * (L->cl.v[i].cz - L->lo[0]) + (bi >> 8);
                                        const int by = 16 * (L->cl.v[i].cy - L->lo[1]) + ((bi >> 4) & 15);
                                        const int bx = 16 * (L->cl.v[i].cx - L->lo[2]) + (bi & 15);
                                        const uint64_t id = L->base + ((uint64_t)bz * L->gy + by) * L->gx + bx;
                                        off[id] = blob_at[ci] + (pay - co);
                                        len[id] = bl;
                                        q[id] = cq;
                                    }
                                }
                                pay += bl;
                            }
                        }
                    }
                }
                if (!pass) {
                    blob_len[ci] = ext;
                    blob_at[ci] = total;
                    total += (ext + 15) & ~15ull;
                } else if (blob_len[ci]) {
                    memcpy(stage + blob_at[ci], arc + co, blob_len[ci]);
                }
            }
        }
        if (!pass) {
            stage = malloc(total ? total : 1);
            if (!stage) goto done;
        }
    }
    int32_t geom[9 * 8];
    for (int l = 0; l < nlods; ++l) {
        const serve_level *L = &lv[l];
        int32_t *g = geom + 9 * l;
        g[0] = L->gz; g[1] = L->gy; g[2] = L->gx;
        g[3] = 16 * L->lo[0]; g[4] = 16 * L->lo[1]; g[5] = 16 * L->lo[2];
        const int sh = lod0 + l;
        g[6] = dims ? dims[3 * l] : (int32_t)(((int64_t)a->nz + (1 << sh) - 1) >> sh);
        g[7] = dims ? dims[3 * l + 1] : (int32_t)(((int64_t)a->ny + (1 << sh) - 1) >> sh);
        g[8] = dims ? dims[3 * l + 2] : (int32_t)(((int64_t)a->nx + (1 << sh) - 1) >> sh);
    }
    s = calloc(1, sizeof *s);
    if (!s) goto done;
    s->oz = 256L * lv[0].lo[0]; s->oy = 256L * lv[0].lo[1]; s->ox = 256L * lv[0].lo[2];
    s->a = a;
    s->lod0 = lod0;
    s->nlods = nlods;
    atomic_store(&s->gen, gen0);
    pthread_mutex_init(&s->mu, NULL);
    for (int l = 0; l < nlods; ++l) {
        serve_grid *G = &s->g[l];
        for (int k = 0; k < 3; ++k) {
            G->lo[k] = lv[l].lo[k];
            G->nc[k] = lv[l].hi[k] - lv[l].lo[k] + 1;
        }
        G->gz = lv[l].gz; G->gy = lv[l].gy; G->gx = lv[l].gx;
        G->base = lv[l].base;
        G->co = calloc((size_t)G->nc[0] * G->nc[1] * G->nc[2], sizeof *G->co);
        if (!G->co) { serve_free_grids(s); free(s); s = NULL; goto done; }
        for (int i = 0; i < lv[l].cl.n; ++i) {
            const serve_chunk *c = &lv[l].cl.v[i];
            G->co[((size_t)(c->cz - G->lo[0]) * G->nc[1] + (c->cy - G->lo[1])) * G->nc[2] + (c->cx - G->lo[2])] = c->co;
        }
    }
    if (!mch_dev()) { serve_free_grids(s); free(s); s = NULL; goto done; }
    s->st = mchip_store_create_levels(stage, total, nlods, geom, off, len, q);
    if (!s->st) { mch_dev_error("mc_serve_open"); serve_free_grids(s); free(s); s = NULL; goto done; }
    s->sv = mchip_server_create(s->st, arena_bytes, lanes);
    if (!s->sv) {
        mch_dev_error("mc_serve_open");
        mchip_store_destroy(s->st);
        serve_free_grids(s);
        free(s);
        s = NULL;
        goto done;
    }
done:
    free(stage); free(off); free(len); free(q); free(blob_at); free(blob_len);
    for (int l = 0; l < nlods; ++l) free(lv[l].cl.v);
    return s;
}

int mc_serve_refresh(mc_serve *s) {
    if (!s) return -1;
    const uint64_t gen = atomic_load_explicit(&s->a->gen, memory_order_acquire);
    if (gen == atomic_load_explicit(&s->gen, memory_order_acquire)) return 0;
    pthread_mutex_lock(&s->mu);
    int outside = 0, rc = 0;
    if (gen != atomic_load_explicit(&s->gen, memory_order_relaxed)) {
        const uint8_t *arc = s->a->base;
        const uint64_t end = atomic_load_explicit(&s->a->cursor, memory_order_acquire);
        for (int l = 0; l < s->nlods && rc >= 0; ++l) {
            serve_grid *G = &s->g[l];
            const size_t nch = (size_t)G->nc[0] * G->nc[1] * G->nc[2];
            uint64_t *now = calloc(nch ? nch : 1, sizeof *now);
            chunk_list cl;
            memset(&cl, 0, sizeof cl);
            uint64_t root;
            memcpy(&root, arc + MCH_ROOTOFF + (uint64_t)(s->lod0 + l) * 8, 8);
            walk(arc, end, root, 2, 0, 0, 0, &cl);
            if (!now || cl.oom) {
                fprintf(stderr, "mc_serve_refresh: out of host memory\n");
                rc = -1;
            }
            for (int i = 0; i < cl.n && rc >= 0; ++i) {
                const serve_chunk *c = &cl.v[i];
                const int rz = c->cz - G->lo[0], ry = c->cy - G->lo[1], rx = c->cx - G->lo[2];
                if (rz < 0 || ry < 0 || rx < 0 || rz >= G->nc[0] || ry >= G->nc[1] || rx >= G->nc[2]) {
                    outside++;
                    continue;
                }
                now[((size_t)rz * G->nc[1] + ry) * G->nc[2] + rx] = c->co;
            }
            for (size_t k = 0; k < nch && rc >= 0; ++k) {
                if (now[k] == G->co[k]) continue;
                const int rz = (int)(k / ((size_t)G->nc[1] * G->nc[2])), ry = (int)((k / G->nc[2]) % G->nc[1]),
                          rx = (int)(k % G->nc[2]);
                if (serve_update_chunk(s, l, G->lo[0] + rz, G->lo[1] + ry, G->lo[2] + rx, now[k]) < 0) rc = -1;
            }
            free(now);
            free(cl.v);
        }
        if (rc >= 0) atomic_store_explicit(&s->gen, gen, memory_order_release);
    }
    pthread_mutex_unlock(&s->mu);
    return rc < 0 ? -1 : outside;
}

int mc_serve_invalidate_chunk(mc_serve *s, int lod, int cz, int cy, int cx) {
    if (!s || lod < s->lod0 || lod >= s->lod0 + s->nlods) return -1;
    pthread_mutex_lock(&s->mu);
    const uint64_t co = mc_archive_chunk_offset(s->a, lod, cz, cy, cx);
    const int rc = serve_update_chunk(s, lod - s->lod0, cz, cy, cx, co);
    pthread_mutex_unlock(&s->mu);
    return rc;
}

int mc_serve_resize(mc_serve *s, uint64_t arena_bytes) {
    if (!s) return -1;
    if (mchip_server_resize(s->sv, arena_bytes) != 0) {
        mch_dev_error("mc_serve_resize");
        return -1;
    }
    return 0;
}

mc_serve *mc_serve_open(mc_archive *a, int lod, uint64_t arena_bytes, int lanes) {
    if (!a || lod < 0 || lod > 7) return NULL;
    return mc_serve_open_levels(a, lod, 1, NULL, arena_bytes, lanes);
}

int64_t mc_serve_render_quad(mc_serve *s, const float *grid, int gw, int gh, float x0, float y0, float step, int w,
                             int h, int filter, int comp, float t0, float t1, float dt, void *d_out) {
    if (!s || mc_serve_refresh(s) < 0) return -1;
    return mchip_server_submit_render(s->sv, grid, gw, gh, x0, y0, step, w, h, filter, comp, t0, t1, dt,
                                      (uint8_t *)d_out);
}

int64_t mc_serve_submit(mc_serve *s, const mc_box *boxes, int n, long dz, long dy, long dx, void *d_out,
                        size_t batch_stride) {
    if (!s || !boxes || n <= 0 || dz <= 0 || dy <= 0 || dx <= 0 || !d_out) return -1;
    if (mc_serve_refresh(s) < 0) return -1;
    int64_t *org = malloc(sizeof *org * 3 * (size_t)n);
    if (!org) return -1;
    for (int i = 0; i < n; ++i) {
        org[3 * i] = boxes[i].z0 - s->oz;
        org[3 * i + 1] = boxes[i].y0 - s->oy;
        org[3 * i + 2] = boxes[i].x0 - s->ox;
    }
    const int64_t t = mchip_server_submit(s->sv, org, n, dz, dy, dx, (uint8_t *)d_out, batch_stride);
    free(org);
    return t;
}

int mc_serve_poll(mc_serve *s, int64_t ticket) { return s ? mchip_server_poll(s->sv, ticket) : -1; }

int mc_serve_wait(mc_serve *s, int64_t ticket, double *latency_ms) {
    return s ? mchip_server_wait(s->sv, ticket, latency_ms) : -1;
}

int mc_serve_preload(mc_serve *s, uint64_t *loaded) {
    if (!s) return -1;
    return mchip_server_preload(s->sv, loaded);
}

mchip_server *mc_serve_device(mc_serve *s) { return s ? s->sv : NULL; }

void mc_serve_close(mc_serve *s) {
    if (!s) return;
    mchip_server_destroy(s->sv);
    mchip_store_destroy(s->st);
    serve_free_grids(s);
    pthread_mutex_destroy(&s->mu);
    free(s);
}
