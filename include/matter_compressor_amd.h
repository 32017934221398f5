/*
 * matter_compressor_amd.h -- drop-in C API of the matter-compressor codec and
 * archive, implemented for one MI355X (gfx950).
 *
 * Every declaration below keeps the name, signature, argument meaning and
 * error behaviour of the reference declaration it replaces
 * (SuperOptimizer/matter-compressor src/matter_compressor.h; line ranges are
 * given per entry).  Block and chunk coding run on the GPU through the mchip
 * engine (include/mchip.h); the archive file format (v7 .mca) and its
 * index, install and metadata logic are plain C on the host.  A program
 * built against the reference's codec/archive section links against
 * libmc_amd.so instead, unchanged, unless it calls mc_build or
 * mc_build_to_file (src/matter_compressor.h:151-154, the whole-volume
 * builders over a voxel callback): those are not exported, since the builder is outside
 * this library; such a program appends its chunks with
 * mc_archive_append_chunk_par instead.
 */
#ifndef MATTER_COMPRESSOR_AMD_H
#define MATTER_COMPRESSOR_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint8_t mc_u8;
typedef int32_t mc_i32;

#define MC_BLK 16           /* block edge */
#define MC_CHUNK 256        /* chunk edge (16 blocks) */
#define MC_DZ_FRAC 0.80f    /* dead-zone width fraction */
#define MC_HF_EXP 0.65f     /* step = quality * (1 + L1)^MC_HF_EXP */
#define MC_BLOCKCODEC_CABAC 0u

/* ---- codec context (src/matter_compressor.h:36-59) ----------------------- */
typedef struct mc_codec_ctx mc_codec_ctx;
mc_codec_ctx *mc_codec_ctx_new(void);                                /* :42 */
void  mc_codec_ctx_free(mc_codec_ctx *ctx);                          /* :43 */
void  mc_codec_ctx_set_quality(mc_codec_ctx *ctx, float quality);    /* :44 */
float mc_codec_ctx_get_quality(mc_codec_ctx *ctx);                   /* :45 */
void  mc_codec_ctx_set_max_error(mc_codec_ctx *ctx, int tau);        /* :58 */
int   mc_codec_ctx_get_max_error(mc_codec_ctx *ctx);                 /* :59 */

/* ---- preset ladder (src/matter_compressor.h:74-92) ------------------------ */
typedef enum {
    MC_PRESET_ARCHIVAL = 0, MC_PRESET_MASTER = 1, MC_PRESET_HIGH = 2, MC_PRESET_BALANCED = 3,
    MC_PRESET_STREAMING = 4, MC_PRESET_FAST = 5, MC_PRESET_ULTRAFAST = 6, MC_PRESET_PREVIEW = 7,
    MC_PRESET_COUNT = 8,
} mc_preset;
float       mc_apply_preset(mc_codec_ctx *ctx, mc_preset level);    /* :88 */
float       mc_preset_quality(mc_preset level);                     /* :89 */
int         mc_preset_tau(mc_preset level);                         /* :90 */
const char *mc_preset_name(mc_preset level);                        /* :91 */
void  mc_codec_init(void);                                          /* :92 */
void  mc_codec_set_priors(const uint16_t *plo, const uint16_t *phi); /* :96 */

/* ---- buffers and block codec (src/matter_compressor.h:100-118) ----------- */
typedef struct { mc_u8 *p; size_t len, cap; } mc_buf;
void  mc_buf_put(mc_buf *b, const void *s, size_t n);               /* :100 */
/* encode one 16^3 block: appends the payload to out; returns 1 if coded, 0
 * for an all-zero block (no payload).  Runs on the GPU. (:111) */
int   mc_enc_block(mc_codec_ctx *ctx, const mc_u8 *vox, mc_buf *out, uint32_t *len_out);
/* decode one block payload into dst (16^3).  Runs on the GPU. (:114) */
void  mc_dec_block(mc_codec_ctx *ctx, const mc_u8 *payload, uint32_t plen, mc_u8 *dst);
/* decode by archive block codec id (CABAC only; src/mc_codec.h:20) */
void  mc_dec_block_codec(mc_codec_ctx *ctx, uint32_t codec, const mc_u8 *payload, uint32_t plen, mc_u8 *dst);
/* decode-side deblocking of a decoded nz x ny x nx volume in place; runs on
 * the GPU (:106) */
void  mc_deblock(mc_u8 *vol, int nz, int ny, int nx, float quality);
/* per-chunk material-fraction map (:117-118) */
uint32_t mc_enc_fracmap(const mc_u8 *frac, mc_u8 *out, size_t cap);
void     mc_dec_fracmap(const mc_u8 *in, uint32_t len, mc_u8 *frac);

/* ---- archive (src/matter_compressor.h:166-312, :364-379) ------------------ */
typedef struct mc_archive mc_archive;
typedef enum { MC_ABSENT = 0, MC_PRESENT = 1, MC_ZERO = 2 } mc_cover;
typedef struct { long z0, y0, x0; } mc_box;

mc_archive *mc_archive_open(const char *path, int dim, float quality);                  /* :177 */
mc_archive *mc_archive_open_dims(const char *path, int nx, int ny, int nz, float quality); /* :180 */
int      mc_archive_set_block_codec(mc_archive *a, uint32_t codec);                     /* :188 */
uint32_t mc_archive_block_codec(const mc_archive *a);                                   /* :189 */
int mc_archive_block_blob(mc_archive *a, uint64_t chunk_off, int bz, int by, int bx,
                          const uint8_t **ptr, uint32_t *len);                           /* :194 */
int mc_archive_append_chunk_raw(mc_archive *a, int lod, int cz, int cy, int cx,
                                const mc_u8 vox[256 * 256 * 256]);                       /* :201 */
int mc_archive_append_chunk_raw_q(mc_archive *a, int lod, int cz, int cy, int cx,
                                  const mc_u8 vox[256 * 256 * 256], float q);            /* :206 */
int mc_archive_append_chunk_ctx(mc_archive *a, mc_codec_ctx *C, int lod, int cz, int cy, int cx,
                                const mc_u8 vox[256 * 256 * 256]);                       /* :212 */
/* rate control: sample-encode 256 blocks at the archive quality, scale q by
 * bytes ~ q^-0.75 toward 256^3 / target_ratio bytes, encode at that q (:221) */
int mc_archive_append_chunk_target(mc_archive *a, int lod, int cz, int cy, int cx,
                                   const mc_u8 vox[256 * 256 * 256], float target_ratio, float *q_out);
int mc_archive_append_chunk_compressed(mc_archive *a, int lod, int cz, int cy, int cx,
                                       const uint8_t *blob, size_t len);                 /* :227 */
int mc_archive_reserve_index(mc_archive *a, int lod, int cz, int cy, int cx);           /* :235 */
mc_cover mc_archive_chunk_coverage(mc_archive *a, int lod, int cz, int cy, int cx);     /* :238 */
uint64_t mc_archive_data_len(mc_archive *a);                                            /* :243 */
int mc_archive_set_metadata(mc_archive *a, const void *data, size_t len);               /* :250 */
const char *mc_archive_metadata(mc_archive *a, size_t *out_len);                        /* :251 */
uint64_t mc_archive_chunk_offset(mc_archive *a, int lod, int cz, int cy, int cx);       /* :255 */
void mc_archive_decode_block(mc_archive *a, uint64_t chunk_off, int bz, int by, int bx,
                             mc_u8 *dst);                                                /* :260 */
/* decode a whole 256^3 chunk (z,y,x raster); nthreads is accepted and
 * ignored: the work runs on the GPU.  chunk_off <= 1 zero-fills. (:265) */
void mc_archive_decode_chunk(mc_archive *a, uint64_t chunk_off, mc_u8 *out, int nthreads);
/* encode a chunk at q (q <= 0 -> the archive's quality) with tau = 0 and
 * install the blob; an all-air chunk is left absent.  0 on success. (:269) */
int mc_archive_append_chunk_par(mc_archive *a, int lod, int cz, int cy, int cx,
                                const mc_u8 vox[256 * 256 * 256], float q, int nthreads);
void mc_archive_read_region(mc_archive *a, int lod, long z0, long y0, long x0,
                            long dz, long dy, long dx, mc_u8 *out, size_t sz, size_t sy,
                            int nthreads);                                               /* :279 */
int   mc_archive_block_present(mc_archive *a, int lod, int bz, int by, int bx);         /* :287 */
float mc_archive_block_fraction(mc_archive *a, int lod, int bz, int by, int bx);        /* :292 */
int mc_archive_sample_boxes(mc_archive *a, int lod, uint64_t seed, int count,
                            long dz, long dy, long dx, float min_frac, mc_box *out);      /* :300 */
void mc_archive_read_regions(mc_archive *a, int lod, const mc_box *boxes, int n,
                             long dz, long dy, long dx, mc_u8 *out, size_t batch_stride,
                             int nthreads);                                              /* :307 */
void mc_archive_close(mc_archive *a);                                                   /* :312 */
const char *mc_metadata(const uint8_t *arc, size_t *out_len);                           /* :364 */
uint64_t mc_chunk_compute_hash(const uint8_t *blob, uint64_t blob_len);                 /* :369 */
long mc_verify_archive(const uint8_t *arc, size_t len, int verbose);                    /* :370 */
int mc_archive_set_priors(struct mc_archive *a, const uint16_t plo[8][32],
                          const uint16_t phi[8][32]);                                   /* :377 */
/* install an archive's prior override as the process priors (used by the
 * reference's reader, src/mc_stream.c; declared in src/mc_archive.h:27) */
void priors_load(const uint8_t *base);

/* ---- read-only readers over archive bytes (src/matter_compressor.h:313-361)
 * A flat reader reads an archive held in memory; a streaming reader pulls
 * byte ranges through a callback, fetching whole chunk blobs or (partial
 * fetch) only a chunk's bitmap and lengths plus each block's own payload.
 * Block decodes run on the GPU; voxels equal the reference reader's. */
typedef struct mc_reader mc_reader;
typedef int (*mc_read_fn)(void *ud, uint64_t off, uint32_t len, uint8_t *dst);
mc_reader *mc_open(const uint8_t *arc, size_t len);                                      /* :316 */
void       mc_close(mc_reader *r);                                                       /* :317 */
void       mc_reader_set_quality(mc_reader *r, float q);                                 /* :318 */
uint64_t   mc_chunk_offset(mc_reader *r, int lod, int cz, int cy, int cx);               /* :319 */
uint64_t   mc_chunk_offset_chk(mc_reader *r, int lod, int cz, int cy, int cx, int *err); /* :323 */
void       mc_decode_block(mc_reader *r, uint64_t chunk_off, int bz, int by, int bx,
                           mc_u8 *dst);                                                  /* :324 */
mc_reader *mc_open_streaming(mc_read_fn read, void *ud, uint64_t total_len);             /* :334 */
void       mc_reader_set_partial_fetch(mc_reader *r, int on);                            /* :341 */
uint64_t   mc_reader_chunk_blob_len(mc_reader *r, uint64_t chunk_off);                   /* :346 */
int        mc_reader_read_blob(mc_reader *r, uint64_t chunk_off, size_t len, uint8_t *dst); /* :352 */
int        mc_reader_priors(mc_reader *r, const uint16_t **plo, const uint16_t **phi);   /* :356 */
void       mc_reader_dims(mc_reader *r, int *nx, int *ny, int *nz);                      /* :359 */
float      mc_reader_quality(mc_reader *r);                                              /* :360 */
int        mc_reader_nlods(mc_reader *r);                                                /* :361 */
/* (additive) decode n blocks through the reader in one GPU launch: block i is
 * (bzyx[3i], bzyx[3i+1], bzyx[3i+2]) of the chunk blob at chunk_off[i], written
 * to dst + 4096*i exactly as mc_decode_block would.  0, or -1 on a device
 * failure (dst zeroed). */
int        mc_reader_decode_blocks(mc_reader *r, int n, const uint64_t *chunk_off, const int32_t *bzyx,
                                   mc_u8 *dst);

/* ---- HBM-resident serving (additive; replaces the reference's decoded-block
 * cache for box batches, mc_cache_new_archive + mc_cache_get_copy,
 * src/matter_compressor.h:425-444, src/mc_cache.c:737-741) -------------------
 * mc_serve_open uploads the present chunks of one LOD to the GPU and starts a
 * serving loop with an HBM arena of arena_bytes (0: half of the free device
 * memory) of decoded blocks and up to `lanes` requests in flight.  A request
 * is a batch of n boxes of dz x dy x dx voxels written to device memory at
 * d_out + i*batch_stride; its voxels equal mc_archive_read_regions' for the
 * same boxes.  Requests submitted while others run are admitted as lanes and
 * arena slots free up and overlap with them on the device.  The arena grows
 * on demand within arena_bytes while the device keeps memory free, and can
 * be resized live.  Appends are seen: each submit compares the archive's
 * generation with the last one seen and, after a change, uploads the chunks
 * whose offsets changed (the reference's mc_cache re-resolution and
 * invalidation, src/mc_cache.c:690-697, :722-736); chunks outside the grid
 * a level had at open are not seen. */
typedef struct mc_serve mc_serve;
mc_serve *mc_serve_open(mc_archive *a, int lod, uint64_t arena_bytes, int lanes);
/* LODs lod0 .. lod0+nlods-1 in one store sharing one arena: box requests read
 * lod0, render requests choose among the levels.  dims (optional) gives each
 * level's voxel dims nz, ny, nx (the bounds its samplers use); NULL takes
 * the archive's dims halved per level, rounded up. */
mc_serve *mc_serve_open_levels(mc_archive *a, int lod0, int nlods, const int32_t *dims, uint64_t arena_bytes,
                               int lanes);
/* Quad render over the levels (mc_render_quad_lod, src/mc_render.c:2302-2314,
 * filters trilinear/nearest, composites none/min/mean/max/alpha/stddev/
 * percentile/depth; see
 * mchip_server_submit_render) into the w x h image at d_out (device).
 * Pixels equal the reference renderer's fed the same decoded levels. */
int64_t   mc_serve_render_quad(mc_serve *s, const float *grid, int gw, int gh, float x0, float y0, float step,
                               int w, int h, int filter, int comp, float t0, float t1, float dt, void *d_out);
int64_t   mc_serve_submit(mc_serve *s, const mc_box *boxes, int n, long dz, long dy, long dx,
                          void *d_out, size_t batch_stride);          /* ticket >= 0, or -1 */
int       mc_serve_poll(mc_serve *s, int64_t ticket);                 /* 1 done, 0 pending, -1 error */
int       mc_serve_wait(mc_serve *s, int64_t ticket, double *latency_ms);
void      mc_serve_close(mc_serve *s);
/* decode every present block of the LOD into the arena while it has room
 * (before or between requests); loaded receives the number of blocks */
int       mc_serve_preload(mc_serve *s, uint64_t *loaded);
/* re-read the chunks whose offsets changed since the archive generation last
 * seen (submits do this by themselves); returns the number of chunks the
 * store's grids do not cover, or -1 */
int       mc_serve_refresh(mc_serve *s);
/* re-read one chunk of level lod from the archive now, dropping its decoded
 * blocks (mc_cache_invalidate_chunk, src/matter_compressor.h:460-463); 0, 1
 * when the chunk lies outside the store's grid, -1 on failure */
int       mc_serve_invalidate_chunk(mc_serve *s, int lod, int cz, int cy, int cx);
/* change the arena budget while requests run (mc_cache_resize,
 * src/matter_compressor.h:522) */
int       mc_serve_resize(mc_serve *s, uint64_t arena_bytes);
/* the device serving loop behind s (for mchip_server_get_stats, include/mchip.h) */
struct mchip_server *mc_serve_device(mc_serve *s);

#ifdef __cplusplus
}
#endif
#endif
