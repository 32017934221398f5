/*
 * mc_host_serve.c -- HBM-resident serving of one LOD of an archive.
 *
 * mc_serve_open walks the LOD's chunk index (src/mc_archive.h:133-145),
 * uploads every present chunk blob once and gives the device the payload
 * range and quality of every block of the chunks' bounding box; requests
 * of boxes then run through the device serving loop (mchip_server_*), which
 * decodes each block at most once while it stays in the HBM arena.  The
 * voxels of every box equal mc_archive_read_regions' for the same box
 * (absent chunks and blocks, and everything outside the LOD's chunks, read
 * as 0).  Chunks appended after mc_serve_open are not seen.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mc_host_internal.h"

struct mc_serve {
    mchip_store *st;
    mchip_server *sv;
    long oz, oy, ox;             /* voxel origin of the store grid */
};

typedef struct {
    int cz, cy, cx;
    uint64_t co;
} serve_chunk;

typedef struct {
    serve_chunk *v;
    int n, cap;
} chunk_list;

static void walk(const uint8_t *arc, uint64_t end, uint64_t node, int lv, int cz, int cy, int cx, chunk_list *out) {
    if (!node || node > end || MCH_NODE_BYTES > end - node) return;
    for (int idx = 0; idx < 4096; ++idx) {
        uint64_t child;
        memcpy(&child, arc + node + (uint64_t)idx * 8, 8);
        if (!child) continue;
        const int z = (cz << 4) | (idx >> 8), y = (cy << 4) | ((idx >> 4) & 15), x = (cx << 4) | (idx & 15);
        if (lv > 0) {
            walk(arc, end, child, lv - 1, z, y, x, out);
        } else if (child > MCH_SLOT_ZERO) {
            if (out->n == out->cap) {
                out->cap = out->cap ? out->cap * 2 : 256;
                out->v = realloc(out->v, sizeof *out->v * out->cap);
            }
            out->v[out->n++] = (serve_chunk){z, y, x, child};
        }
    }
}

mc_serve *mc_serve_open(mc_archive *a, int lod, uint64_t arena_bytes, int lanes) {
    if (!a || lod < 0 || lod > 7 || lanes <= 0) return NULL;
    const uint8_t *arc = a->base;
    const uint64_t end = atomic_load_explicit(&a->cursor, memory_order_acquire);
    uint64_t root;
    memcpy(&root, arc + MCH_ROOTOFF + (uint64_t)lod * 8, 8);
    chunk_list cl = {0};
    walk(arc, end, root, 2, 0, 0, 0, &cl);
    int lo[3] = {4096, 4096, 4096}, hi[3] = {-1, -1, -1};
    for (int i = 0; i < cl.n; ++i) {
        const int c[3] = {cl.v[i].cz, cl.v[i].cy, cl.v[i].cx};
        for (int k = 0; k < 3; ++k) {
            if (c[k] < lo[k]) lo[k] = c[k];
            if (c[k] > hi[k]) hi[k] = c[k];
        }
    }
    if (cl.n == 0) { lo[0] = lo[1] = lo[2] = 0; hi[0] = hi[1] = hi[2] = 0; }
    const int gz = 16 * (hi[0] - lo[0] + 1), gy = 16 * (hi[1] - lo[1] + 1), gx = 16 * (hi[2] - lo[2] + 1);
    const uint64_t nb = (uint64_t)gz * gy * gx;
    if (nb > (1ull << 28)) {                          /* the block grid is the chunks' bounding box */
        fprintf(stderr, "mc_serve_open: chunk bounding box of %llu blocks is too large\n", (unsigned long long)nb);
        free(cl.v);
        return NULL;
    }
    uint64_t *off = calloc(nb, sizeof *off);
    uint32_t *len = calloc(nb, sizeof *len);
    float *q = calloc(nb, sizeof *q);
    uint64_t total = 0;
    uint64_t *blob_at = malloc(sizeof *blob_at * (cl.n ? cl.n : 1));
    uint64_t *blob_len = malloc(sizeof *blob_len * (cl.n ? cl.n : 1));
    uint8_t *stage = NULL;
    mc_serve *s = NULL;
    if (!off || !len || !q || !blob_at || !blob_len) goto done;
    /* Pass 1: per-block payload ranges, each bounded by the archive's live
     * end like mc_archive_read_regions (src/mc_archive.c:728-740); a chunk's
     * upload extent is the end of its last valid payload. */
    for (int pass = 0; pass < 2; ++pass) {
        for (int i = 0; i < cl.n; ++i) {
            const uint64_t co = cl.v[i].co;
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
                            uint16_t l;
                            memcpy(&l, arc + lens + 2ull * rank, 2);
                            rank++;
                            if (pay + l <= end && l) {
                                if (pay + l - co > ext) ext = pay + l - co;
                                if (pass) {
                                    const int bz = 16 * (cl.v[i].cz - lo[0]) + (bi >> 8);
                                    const int by = 16 * (cl.v[i].cy - lo[1]) + ((bi >> 4) & 15);
                                    const int bx = 16 * (cl.v[i].cx - lo[2]) + (bi & 15);
                                    const uint64_t id = ((uint64_t)bz * gy + by) * gx + bx;
                                    off[id] = blob_at[i] + (pay - co);
                                    len[id] = l;
                                    q[id] = cq;
                                }
                            }
                            pay += l;
                        }
                    }
                }
            }
            if (!pass) {
                blob_len[i] = ext;
                blob_at[i] = total;
                total += (ext + 15) & ~15ull;
            } else if (blob_len[i]) {
                memcpy(stage + blob_at[i], arc + co, blob_len[i]);
            }
        }
        if (!pass) {
            stage = malloc(total ? total : 1);
            if (!stage) goto done;
        }
    }
    s = calloc(1, sizeof *s);
    if (!s) goto done;
    s->oz = 256L * lo[0]; s->oy = 256L * lo[1]; s->ox = 256L * lo[2];
    if (!mch_dev()) { free(s); s = NULL; goto done; }
    s->st = mchip_store_create(stage, total, off, len, q, gz, gy, gx);
    if (!s->st) { mch_dev_error("mc_serve_open"); free(s); s = NULL; goto done; }
    s->sv = mchip_server_create(s->st, arena_bytes, lanes);
    if (!s->sv) { mch_dev_error("mc_serve_open"); mchip_store_destroy(s->st); free(s); s = NULL; goto done; }
done:
    free(stage); free(off); free(len); free(q); free(blob_at); free(blob_len); free(cl.v);
    return s;
}

int64_t mc_serve_submit(mc_serve *s, const mc_box *boxes, int n, long dz, long dy, long dx, void *d_out,
                        size_t batch_stride) {
    if (!s || !boxes || n <= 0 || dz <= 0 || dy <= 0 || dx <= 0 || !d_out) return -1;
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
    free(s);
}
