/*
 * mchip.h -- C ABI of the MI355X (gfx950) codec engine behind the
 * matter-compressor drop-in library.  Plain pointers and sizes only.
 *
 * These entry points are the device half of the boundary: the plain-C host
 * library (include/matter_compressor_amd.h) implements the reference's own
 * mc_codec / mc_archive API on top of them, and a caller that keeps chunk
 * blobs resident in HBM can call them directly.  Each one replaces the work
 * loop of a reference function (SuperOptimizer/matter-compressor):
 *
 *   mchip_decode_chunks      mc_archive_decode_chunk's worker pool
 *                            (src/mc_archive.c:783-820) over n chunks at once
 *   mchip_decode_blocks      mc_archive_decode_block / mc_dec_block
 *                            (src/mc_archive.c:728-762, src/mc_codec.c:739-783)
 *   mchip_encode_chunks      mc_archive_append_chunk_par's encode + stitch
 *                            (src/mc_archive.c:822-901) and, with tau > 0,
 *                            encode_chunk_blob_codec (src/mc_codec.c:809-845)
 *   mchip_encode_blocks      mc_enc_block (src/mc_codec.c:629-737)
 *
 * All functions return 0 on success and a negative code on failure; the
 * message of the last failure on the calling thread is mchip_last_error().
 * Functions taking an mchip_ctx enqueue work on that context's HIP stream;
 * calls on one context are serialised, distinct contexts may run
 * concurrently.
 */
#ifndef MCHIP_H
#define MCHIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mchip_ctx mchip_ctx;

/* Select the HIP device and upload the constant tables.  Idempotent. */
int         mchip_init(int device);
const char *mchip_last_error(void);

/* Per-volume prior override (u16[8][32] q=1 and q=12 endpoint tables);
 * NULL,NULL restores the trained defaults.  Mirrors mc_codec_set_priors
 * (src/mc_codec.c:141-154, :533). */
int         mchip_set_priors(const uint16_t *plo, const uint16_t *phi);

/* Launch-shape tuning (process-wide, read as one snapshot per call):
 * "decoder" selects the entropy decoder variant (0 = one loop step per scan
 * position; 2 = one bin per loop step, branch-free, payload in an LDS
 * window; 3 = the same bins with the loop in assembly and lanes that meet a
 * rare event paused and served together, the default); "persistent",
 * "persist_min", "persist_wg" and "persist_win" (payload window words per
 * lane: 16, 32 = default, 64) shape the persistent entropy launch of batched
 * decodes; "len_shift" and
 * "group_shift" the payload-length buckets of the lane order and of the
 * persistent group order; "recon_bpw" the reconstruction blocks per
 * workgroup; "slices" the chunk slices of a batched persistent decode, whose
 * reconstruction overlaps the next slice's entropy pass on a second stream
 * (0 = two from 16 chunks up, else one); "tuned_clear" drops the autotuned shapes; "enc_batch" sets the
 * chunks per encode launch sequence (0 = sized from free device memory),
 * "enc_slices" the chunk slices of one encode sequence, run alternately on
 * two streams (0 = one slice; more slices measured slower on the mixed
 * workload, as each slice's entropy pass then ends in its own tail),
 * "enc_order" the entropy pass's chunk order (1 = heaviest chunk first, the
 * default; 0 = sequence order), "enc_lanes" its block order within a chunk
 * (0 = raster, the default; 1 = costliest blocks first, so that each wave's
 * 64 lanes code blocks of like cost: faster on mixed-quality chunks, slower
 * where every wave is resident at once), and
 * "enc_keep_mib" the encode workspace a context keeps after a call;
 * "merge" (default 1), "merge_inflight" and "merge_min" shape the merging of
 * concurrent one-chunk host decodes.  Results
 * are identical; only speed differs. */
int         mchip_set_tuning(const char *key, int value);

/* A context owns one HIP stream and a growable device workspace. */
mchip_ctx  *mchip_ctx_create(void);
void        mchip_ctx_destroy(mchip_ctx *ctx);
void       *mchip_ctx_stream(mchip_ctx *ctx);        /* the hipStream_t */
int         mchip_ctx_sync(mchip_ctx *ctx);

/* Device memory helpers (so callers need no HIP headers). */
void       *mchip_malloc(size_t bytes);
void        mchip_free(void *dptr);
/* Free and total device memory in bytes. */
int         mchip_mem_info(size_t *free_bytes, size_t *total_bytes);
int         mchip_h2d(mchip_ctx *ctx, void *dst, const void *src, size_t bytes);
int         mchip_d2h(mchip_ctx *ctx, void *dst, const void *src, size_t bytes);
int         mchip_memset(mchip_ctx *ctx, void *dst, int value, size_t bytes);
/* device-to-device copy on the context's stream (asynchronous) */
int         mchip_d2d(mchip_ctx *ctx, void *dst, const void *src, size_t bytes);

/* Decode n chunk blobs (format v7) that are resident in device memory.
 * Blob i lives at d_blobs + blob_off[i] with blob_len[i] readable bytes; q[i]
 * is the quality stored in its header.  Chunk i is written as a dense 256^3
 * z,y,x raster at d_out + i*out_stride (absent blocks -> 0).  Asynchronous on
 * the context's stream.  Host arrays are read before the call returns. */
int mchip_decode_chunks(mchip_ctx *ctx, const uint8_t *d_blobs, const uint64_t *blob_off,
                        const uint64_t *blob_len, const float *q, int n,
                        uint8_t *d_out, uint64_t out_stride);

/* The decode of mchip_decode_chunks with no allocation of any kind: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  Its workspace is a fixed device array of 128 MiB that
 * the library holds on each device it runs on until the process ends; the
 * quality tables and chunk descriptors travel in the kernel arguments, so
 * the call issues kernel launches only.  At most 4 chunks per call, and
 * their workspace (about 0.6 MiB per chunk plus 32 bytes per blob byte, so
 * three q=6 chunks of the workload) must fit in the array: larger calls are
 * refused with an error.  mchip_init must have run.  Work is enqueued on
 * stream (a hipStream_t; NULL = the null stream) and is asynchronous.
 * Every call on a device shares that device's array, which keeps the last
 * call's contents: calls that overlap on one device race and give wrong
 * voxels, so serialise them (one stream, or one call at a time). */
int mchip_decode_chunks_static(const uint8_t *d_blobs, const uint64_t *blob_off, const uint64_t *blob_len,
                               const float *q, int n, uint8_t *d_out, uint64_t out_stride, void *stream);

/* Host-memory convenience: upload n blobs, decode, download the rasters to
 * out + i*out_stride, synchronise. */
int mchip_decode_chunks_host(mchip_ctx *ctx, const uint8_t *const *blobs, const uint64_t *blob_len,
                             int n, uint8_t *out, uint64_t out_stride);

/* One chunk blob in host memory decoded into the 256^3 raster at out (host),
 * synchronously, as mchip_decode_chunks_host with n = 1, but merged with the
 * calls other threads make at the same time: each call uploads its blob on
 * ctx's stream into a slot of a shared device arena, one worker issues the
 * queued calls as one batched decode when no batch is in flight, or when
 * "merge_min" calls (default 8) wait and fewer than "merge_inflight" batches
 * (default 2) are in flight, and each call copies its chunk back on ctx's
 * stream.  Blobs over 4 MiB, or tuning "merge" = 0, take the caller's own
 * decode.  Replaces the per-call worker pool of mc_archive_decode_chunk
 * (src/mc_archive.c:783-820) for concurrent callers. */
int mchip_decode_chunk_merged(mchip_ctx *ctx, const uint8_t *blob, uint64_t blob_len, uint8_t *out);
/* Batches issued and chunks decoded by the merged path so far. */
int mchip_merge_stats(uint64_t *batches, uint64_t *chunks);

/* Decode n independent block payloads (host memory) at quality q[i] into
 * out + 4096*i (16^3 z,y,x).  A NULL payload or blob-absent block writes
 * zeros.  Synchronous. */
int mchip_decode_blocks_host(mchip_ctx *ctx, const uint8_t *const *payloads, const uint32_t *lens,
                             const float *q, int n, uint8_t *out);

/* Encode n dense 256^3 chunks resident in device memory (chunk i at
 * d_in + i*in_stride, 16-byte aligned) into v7 blobs at d_out + out_off[i]
 * (capacity out_cap[i]; the buffer must stay readable 16 bytes past each
 * region).  q[i] selects the quantiser and priors (sanitised like a codec
 * context); tau[i] > 0 adds max-error corrections (NULL = all 0); the blob
 * header stores q_store[i] (NULL = q[i] as given).  On return blob_len[i]
 * is the blob length (0 = all-air chunk, no blob) and status[i] (may be
 * NULL) is 0, 1 (a block payload exceeded 17408 bytes) or 2 (out_cap too
 * small; blob_len then holds the needed length).  Synchronous. */
int mchip_encode_chunks(mchip_ctx *ctx, const uint8_t *d_in, uint64_t in_stride, const float *q,
                        const int *tau, const float *q_store, int n, uint8_t *d_out,
                        const uint64_t *out_off, const uint64_t *out_cap, uint64_t *blob_len,
                        int *status);

/* Host-memory convenience for mchip_encode_chunks: chunk i at
 * chunks + i*16777216, blob i written to blobs[i] (capacity caps[i]). */
int mchip_encode_chunks_host(mchip_ctx *ctx, const uint8_t *chunks, const float *q, const int *tau,
                             const float *q_store, int n, uint8_t *const *blobs, const uint64_t *caps,
                             uint64_t *blob_len, int *status);

/* Encode n independent 16^3 blocks (host memory, block i at vox + 4096*i,
 * z,y,x) at (q[i], tau[i]) (tau may be NULL).  payloads[i] receives up to
 * 17408 bytes; lens[i] = payload length (0 = all-zero block); status[i]
 * (may be NULL) = 1 if the payload exceeded 17408 bytes.  Synchronous. */
int mchip_encode_blocks_host(mchip_ctx *ctx, const uint8_t *vox, const float *q, const int *tau, int n,
                             uint8_t *const *payloads, uint32_t *lens, int *status);

/* Decode n block payloads resident in device memory into 4096-byte slots of
 * a device arena: payload i at d_pay + pay_off[i] (pay_len[i] bytes,
 * quality q[i]) becomes the 16^3 block at d_arena + 4096*slot[i].
 * Asynchronous; host arrays are consumed before the call returns. */
int mchip_decode_block_jobs(mchip_ctx *ctx, const uint8_t *d_pay, const uint64_t *pay_off,
                            const uint32_t *pay_len, const float *q, const uint32_t *slot, int n,
                            uint8_t *d_arena);

/* Assemble n boxes of dz x dy x dx voxels from arena blocks into
 * d_out + i*batch_stride (dense z,y,x).  Box i has origin[3i..3i+2] =
 * (z0,y0,x0); its touched block grid starts at block origin>>4 and holds
 * gz x gy x gx cells in bmap[i*gz*gy*gx ...] (z,y,x order), each an arena
 * slot or -1 for zeros.  Asynchronous. */
int mchip_crop_boxes(mchip_ctx *ctx, const uint8_t *d_arena, const int32_t *bmap, const int64_t *origin,
                     int n, int gz, int gy, int gx, int64_t dz, int64_t dy, int64_t dx, uint8_t *d_out,
                     uint64_t batch_stride);

/* Host-memory region read: the nblk payloads (host pointers) are decoded
 * into slots 0..nblk-1 of a device arena, the n boxes are assembled on the
 * device (as mchip_crop_boxes) and copied to out + i*batch_stride.
 * Synchronous.  Replaces the worker pool of mc_archive_read_regions
 * (src/mc_archive.c:950-1042). */
int mchip_read_regions_host(mchip_ctx *ctx, int nblk, const uint8_t *const *payloads, const uint32_t *lens,
                            const float *q, const int32_t *bmap, const int64_t *origin, int n, int gz,
                            int gy, int gx, int64_t dz, int64_t dy, int64_t dx, uint8_t *out,
                            uint64_t batch_stride);

/* XXH64 of n byte ranges of device memory: range i is d_base + off[i],
 * len[i] bytes, hashed with seed[i] (the published XXH64 algorithm, the
 * mc_xxh64 of src/mc_archive.c); out[i] (host) receives the hash.  Any
 * alignment.  Synchronous. */
int mchip_xxh64_ranges(mchip_ctx *ctx, const uint8_t *d_base, const uint64_t *off, const uint64_t *len,
                       const uint64_t *seed, int n, uint64_t *out);

/* Chunk-blob verification of n device-resident blobs: blob i is d_base +
 * off[i], len[i] bytes.  hash_out[i] receives the recomputed blob hash
 * (mc_chunk_compute_hash, src/mc_archive.c:76-88) and bad[i] 1 when the
 * blob's structure does not fit in len[i] or the recomputed hash differs from
 * the one in its header (src/mc_archive.c:1093-1100), else 0.  Either output
 * may be NULL.  Synchronous; returns the number of bad blobs, or -1. */
long mchip_verify_blobs(mchip_ctx *ctx, const uint8_t *d_base, const uint64_t *off, const uint64_t *len, int n,
                        uint64_t *hash_out, uint8_t *bad);

/* The same for blobs in host memory (blobs[i], len[i] bytes): gathered by
 * host threads into pinned slabs, uploaded and hashed while the next slab
 * is gathered.  Synchronous; returns the number of bad blobs, or -1. */
long mchip_verify_blobs_host(mchip_ctx *ctx, const uint8_t *const *blobs, const uint64_t *len, int n,
                             uint64_t *hash_out, uint8_t *bad);

/* Quad render over an LOD pyramid of dense rasters in device memory
 * (d_levels[l], dims[3l..3l+2] = nz, ny, nx): as mchip_server_submit_render,
 * into the w x h image at d_out, on the context's stream (asynchronous).
 * *lod_out (optional) receives the level rendered. */
int mchip_render_quad(mchip_ctx *ctx, int nlods, const uint8_t *const *d_levels, const int32_t *dims,
                      const float *grid, int gw, int gh, float x0, float y0, float step, int w, int h,
                      int filter, int comp, float t0, float t1, float dt, uint8_t *d_out, int *lod_out);

/* 2x LOD decimation (src/mc_archive.c:61-72): the mean of the nonzero voxels
 * of each 2x2x2 cell, rounded half up, all-air cells 0.  mchip_decimate maps
 * a dense dz*dy*dx raster to ((dz+1)/2) x ((dy+1)/2) x ((dx+1)/2);
 * mchip_decimate_chunks maps a 2x2x2 group of 256^3 chunks (d_src[(z*2+y)*2+x],
 * NULL for an absent all-air chunk) to one 256^3 chunk.  Asynchronous. */
int mchip_decimate(mchip_ctx *ctx, const uint8_t *d_src, int dz, int dy, int dx, uint8_t *d_out);
int mchip_decimate_chunks(mchip_ctx *ctx, const uint8_t *const *d_src, uint8_t *d_out);

/* Decode-side deblocking (mc_deblock, src/mc_codec.c:561-595) of a dense
 * nz x ny x nx z,y,x raster in device memory, in place, on the context's
 * stream (asynchronous).  A face whose far neighbour lies past the end of the
 * raster reads that neighbour as 0. */
int mchip_deblock(mchip_ctx *ctx, uint8_t *d_vol, int nz, int ny, int nx, float quality);
/* The same for a raster in host memory (uploaded, filtered, copied back;
 * synchronous). */
int mchip_deblock_host(mchip_ctx *ctx, uint8_t *vol, int nz, int ny, int nx, float quality);

/* Kernel timing: when enabled, every kernel launched on the context is
 * bracketed by HIP events on the context's stream; stats accumulate per
 * kernel name ("index", "entropy", "recon", ...). */
int mchip_ctx_enable_timing(mchip_ctx *ctx, int on);
int mchip_ctx_kernel_time(mchip_ctx *ctx, const char *kernel, double *total_ms, long *launches);
int mchip_ctx_reset_timing(mchip_ctx *ctx);

/* ---- Batched decode as a HIP graph, and the launch-shape autotuner ---------
 * mchip_decode_graph_create captures the launches of one mchip_decode_chunks
 * call (the index, entropy and reconstruction kernels of n resident blobs,
 * with the tuning in force) into a HIP graph over a private workspace;
 * mchip_graph_launch replays it on ctx's stream, decoding the blobs found at
 * the captured device addresses into the captured output.  The graph holds
 * no copies or memsets. */
typedef struct mchip_graph mchip_graph;
mchip_graph *mchip_decode_graph_create(mchip_ctx *ctx, const uint8_t *d_blobs, const uint64_t *blob_off,
                                       const uint64_t *blob_len, const float *q, int n, uint8_t *d_out,
                                       uint64_t out_stride);
int          mchip_graph_launch(mchip_graph *g);
void         mchip_graph_destroy(mchip_graph *g);
/* Time the decode launch shapes on this batch by coordinate descent (entropy
 * decoder, persistent launch, its workgroup count and length buckets, the
 * k_index lane-order buckets, k_recon blocks per workgroup), keep the
 * fastest for batches of the same signature (chunk count and mean blob
 * length as bit lengths, mixed or single quality) -- later decodes of such
 * batches use it -- and write the timings as JSON to report (cap bytes). */
int          mchip_autotune_decode(mchip_ctx *ctx, const uint8_t *d_blobs, const uint64_t *blob_off,
                                   const uint64_t *blob_len, const float *q, int n, uint8_t *d_out,
                                   uint64_t out_stride, char *report, size_t cap);

/* Time the encode shapes on these chunks by coordinate descent (the
 * entropy pass's chunk order, chunks per launch sequence, chunk slices),
 * each candidate a whole mchip_encode_chunks of the caller's chunks into the
 * caller's output (the blobs are rewritten, byte for byte the same); keep
 * the fastest for encodes of the same signature (chunk count as a bit length,
 * mixed or single quality, corrections or none) -- later mchip_encode_chunks
 * calls with that signature use it -- and write the timings as JSON to report
 * (cap bytes).  No counterpart in the reference, whose only encode knob is
 * the thread count of mc_archive_append_chunk_par (src/mc_archive.c:805-809). */
int          mchip_autotune_encode(mchip_ctx *ctx, const uint8_t *d_in, uint64_t in_stride, const float *q,
                                   const int *tau, const float *q_store, int n, uint8_t *d_out,
                                   const uint64_t *out_off, const uint64_t *out_cap, char *report, size_t cap);

/* The launches of one mchip_encode_chunks sequence (n chunks, at most
 * one encode sequence) captured as a HIP graph over a private workspace;
 * mchip_graph_launch replays it on ctx's stream, encoding the rasters at
 * the captured addresses into the captured output, and
 * mchip_encode_graph_results (after the replay) gives blob_len and status
 * as mchip_encode_chunks does. */
mchip_graph *mchip_encode_graph_create(mchip_ctx *ctx, const uint8_t *d_in, uint64_t in_stride, const float *q,
                                       const int *tau, const float *q_store, int n, uint8_t *d_out,
                                       const uint64_t *out_off, const uint64_t *out_cap);
int          mchip_encode_graph_results(mchip_graph *g, uint64_t *blob_len, int *status);

/* ---- Device-resident block store and serving loop ------------------------
 * Replaces the reference's decoded-block cache (mc_cache, S3-FIFO over host
 * memory, src/mc_cache.c:287-384) for box-batch serving: compressed blocks
 * stay resident in HBM, decoded blocks live in an HBM arena, and requests
 * are admitted, decoded and assembled on the device. */
typedef struct mchip_store mchip_store;
typedef struct mchip_server mchip_server;

/* A store over a dense grid of gz*gy*gx 16^3 blocks; block (z,y,x) has
 * index (z*gy + y)*gx + x.  The nbytes host bytes at payload are uploaded
 * once; block i's payload is payload[off[i] .. off[i]+len[i]) coded at
 * quality q[i]; len[i] == 0 marks an absent block (reads as zeros). */
mchip_store *mchip_store_create(const uint8_t *payload, uint64_t nbytes, const uint64_t *off,
                                const uint32_t *len, const float *q, int gz, int gy, int gx);
/* A store over the levels of an LOD pyramid.  Level l is described by
 * geom[9l .. 9l+8] = gz, gy, gx (its block grid), obz, oby, obx (the level
 * block coordinates of grid block (0,0,0)) and nz, ny, nx (the level's voxel
 * dims, the bounds its samplers use).  Blocks are indexed level after level,
 * each level's grid in (z*gy + y)*gx + x order; off/len/q cover all levels.
 * Box requests read level 0 (grid coordinates); render requests choose a
 * level. */
mchip_store *mchip_store_create_levels(const uint8_t *payload, uint64_t nbytes, int nlevels,
                                       const int32_t *geom, const uint64_t *off, const uint32_t *len,
                                       const float *q);
void         mchip_store_destroy(mchip_store *st);
uint64_t     mchip_store_device_bytes(const mchip_store *st);

typedef struct {
    uint64_t requests;        /* submitted */
    uint64_t completed;
    uint64_t boxes;
    uint64_t block_refs;      /* distinct present blocks per request, summed */
    uint64_t hits;            /* of those, already resident (or being decoded) */
    uint64_t misses;          /* decoded into the arena */
    uint64_t evictions;
    uint64_t deferred;        /* admissions postponed for lanes or arena slots */
    uint64_t arena_slots;     /* usable decoded-block slots now */
    uint64_t renders;         /* render requests submitted */
    uint64_t arena_grows;     /* arena segments mapped on demand */
    uint64_t arena_mapped_bytes;  /* device memory backing the arena */
    uint64_t arena_budget_bytes;  /* the most the arena may grow to */
    uint64_t updated_blocks;  /* block payloads replaced by mchip_server_update_blocks */
} mchip_server_stats;

/* A serving loop over st with an arena of at most arena_bytes / 4096
 * decoded-block slots (0: half of the device memory free at creation) and up
 * to lanes requests in flight, each on its own stream.  The arena is a
 * reserved device address range backed on demand (HIP virtual memory): it
 * starts with one 256 MiB segment and grows by segments when a request needs
 * slots, as long as its budget allows and the device keeps 2 GiB free;
 * otherwise resident blocks are replaced (CLOCK).  A request that finds no
 * slots or memory waits for the requests in flight; it fails only when none
 * is in flight.  A worker thread admits the
 * submitted requests in arrival order as lanes and arena slots free up
 * (CLOCK replacement; blocks of requests in flight are pinned), so requests
 * that arrive while others run overlap with them on the device.  A request
 * that needs a block another request in flight is decoding waits for that
 * decode on the device instead of decoding it again. */
mchip_server *mchip_server_create(mchip_store *st, uint64_t arena_bytes, int lanes);
/* Submit n boxes of dz x dy x dx voxels; box i has origin (z, y, x) =
 * origins[3i..3i+2] in voxels and is written densely to d_out + i*batch_stride
 * (device memory that stays valid until the request completes).  Voxels
 * outside the grid and in absent blocks are 0.  Returns a ticket >= 0, or -1. */
int64_t      mchip_server_submit(mchip_server *sv, const int64_t *origins, int n, int64_t dz, int64_t dy,
                                 int64_t dx, uint8_t *d_out, uint64_t batch_stride);
/* Submit a quad render (mc_render_quad_lod, src/mc_render.c:2302-2314): the
 * gw x gh control grid (host, (z,y,x) per point in level-0 voxels, copied) is
 * rendered over the store's levels into the w x h image at d_out (device).
 * filter: 0 nearest, 1 trilinear; comp: 0 none, 1 min, 2 mean, 3 max,
 * 4 alpha, 5 stddev, 7 percentile, 8 depth (the reference's mc_filter /
 * mc_comp values, with the reference's default alpha_min, alpha_opacity and
 * percentile); the slab [t0, t1] along the
 * normal is walked at dt.  The level is chosen and the points remapped as the
 * reference does; the render reads only decoded blocks of that level, which
 * are admitted and decoded like a box request's.  Returns a ticket, or -1. */
int64_t      mchip_server_submit_render(mchip_server *sv, const float *grid, int gw, int gh, float x0,
                                        float y0, float step, int w, int h, int filter, int comp, float t0,
                                        float t1, float dt, uint8_t *d_out);
/* Level a pending render request renders (valid until it is waited for). */
int          mchip_server_render_level(mchip_server *sv, int64_t ticket);
/* 1 once the request has completed, 0 while it is pending, -1 on error. */
int          mchip_server_poll(mchip_server *sv, int64_t ticket);
/* Block until the request completes and release its ticket; latency_ms
 * (optional) receives the time from submission to completion. */
int          mchip_server_wait(mchip_server *sv, int64_t ticket, double *latency_ms);
/* Decode the store's present blocks into free or evictable arena slots, in
 * grid order, until every block is resident or the arena is full (waits
 * until no request is pending or in flight).  With an arena that holds the
 * whole store, every later request is served from decoded blocks only. */
int          mchip_server_preload(mchip_server *sv, uint64_t *loaded);
int          mchip_server_get_stats(mchip_server *sv, mchip_server_stats *out);
/* Change the arena budget while the server runs.  Growing takes effect as
 * requests need slots; shrinking drops the blocks held in slots beyond the
 * new size, waits until no request in flight holds one of them and returns
 * their memory to the device.  (The reference's mc_cache_resize,
 * src/mc_cache.c:200.) */
int          mchip_server_resize(mchip_server *sv, uint64_t arena_bytes);
/* Replace the payloads of n blocks of the store behind sv (after a chunk is
 * appended again, src/mc_cache.c:690-697): block ids[i] (store index) gets
 * bytes payload[off[i] .. off[i]+len[i]) coded at quality q[i] (len 0: the
 * block is now absent).  The bytes are uploaded; decoded copies of these
 * blocks leave the arena, so requests admitted afterwards decode the new
 * payloads, while requests already admitted finish with the old ones. */
int          mchip_server_update_blocks(mchip_server *sv, const uint8_t *payload, uint64_t nbytes, int n,
                                        const uint64_t *ids, const uint64_t *off, const uint32_t *len,
                                        const float *q);
/* Waits for the requests in flight, then frees the lanes and the arena. */
void         mchip_server_destroy(mchip_server *sv);

#ifdef __cplusplus
}
#endif
#endif
