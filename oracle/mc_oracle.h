/*
 * mc_oracle.h -- CPU restatement of the matter-compressor 16^3-block codec.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product library never links or calls it.
 *
 * Every function restates the reference algorithm (SuperOptimizer/
 * matter-compressor, src/mc_codec.c, src/mc_codec_float.h, src/mc_archive.c,
 * src/mc_archive.h) with the same arithmetic, evaluation order and byte
 * format.  It is compiled with -O2 -std=c11 -fno-fast-math -ffp-contract=off,
 * the reference's own deterministic build mode (CMakeLists.txt:36-37).
 *
 * Pinning (what checks each part; see tests/test_oracle.py):
 *   - every byte of block payloads, chunk blobs, fraction maps and hashes is
 *     compared with the reference compiled from its own sources (oracle/_ref)
 *     over the per-block known-answer set and the committed golden fixtures;
 *   - XXH64 against the published xxHash known answers;
 *   - the DCT against orthonormality / round trip closed forms and a
 *     float64 matrix product;
 *   - the range coder against encode->decode identity on random streams;
 *   - the LOD decimation and the quad render against the reference's
 *     mc_build and mc_render_quad_lod compiled from its sources
 *     (oracle/_ref/libmcref_render.so) and against closed-form cases.
 */
#ifndef MC_ORACLE_H
#define MC_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- tables ------------------------------------------------------------ */
/* Interpolated prior table for quality q (8 classes x 32 slots, u16). */
void     orc_priors_for_q(float q, uint16_t pri[8][32]);
/* Override the q=1 / q=12 endpoint tables (NULL,NULL restores the defaults). */
void     orc_set_priors(const uint16_t *plo, const uint16_t *phi);
/* step[i], rstep[i] for the 4096 raster coefficient positions at quality q. */
void     orc_step_table(float q, float step[4096], float rstep[4096]);
/* 16x16 f32 DCT-II matrix cmf[k][n]. */
void     orc_dct_matrix(float cmf[256]);
/* ascending-L1 scan order (4096 raster indices). */
void     orc_scan_order(uint16_t scan[4096]);
/* quality sanitiser used by the codec context. */
float    orc_sanitize_q(float q);

/* ---- transforms ---------------------------------------------------------- */
void     orc_dct3_fwd(const float in[4096], float out[4096]);
void     orc_dct3_inv(const float in[4096], float out[4096]);

/* ---- range coder exercise (raw bins) ------------------------------------ */
/* kinds[i]: 0 = adaptive bit with context ctx_idx[i] (of 64, p0 init 2048),
 * 1 = bypass bit, 2 = batched bypass of nbits[i] bits (value vals[i]).
 * Returns the encoded length (may exceed cap; bytes beyond cap are dropped). */
size_t   orc_rc_encode_ops(int n, const uint8_t *kinds, const uint8_t *ctx_idx,
                           const uint8_t *nbits, const uint32_t *vals,
                           uint8_t *out, size_t cap);
void     orc_rc_decode_ops(int n, const uint8_t *kinds, const uint8_t *ctx_idx,
                           const uint8_t *nbits, const uint8_t *in, size_t len,
                           uint32_t *vals_out);

/* ---- block codec --------------------------------------------------------- */
/* Encode one 16^3 block (z,y,x raster) at quality q with max error tau.
 * Returns payload length (0 = all-zero block, no payload). out must hold
 * cap >= 4096*4+1024 bytes to never truncate. */
uint32_t orc_enc_block(float q, int tau, const uint8_t vox[4096], uint8_t *out, size_t cap);
/* Encode and also report the quantised levels (raster) and dc. */
uint32_t orc_enc_block_ex(float q, int tau, const uint8_t vox[4096], uint8_t *out, size_t cap,
                          int16_t levels[4096], int *dc_out, int *ncorr_out);
/* Decode one payload into dst; q is sanitised like the codec context does. */
void     orc_dec_block(float q, const uint8_t *p, uint32_t plen, uint8_t dst[4096]);
/* Entropy-decode only: levels (raster), dc, has_corr flag, corrections. */
void     orc_dec_block_levels(float q, const uint8_t *p, uint32_t plen, int16_t levels[4096],
                              int *dc, int *flags, uint32_t *ncorr,
                              uint32_t corr_pos[4096], int32_t corr_delta[4096]);

/* ---- fraction map, hash ---------------------------------------------------- */
/* decode-side deblocking of a nz*ny*nx raster in place (src/mc_codec.c:561-595) */
void     orc_deblock(uint8_t *v, int nz, int ny, int nx, float quality);
uint32_t orc_enc_fracmap(const uint8_t frac[4096], uint8_t *out, size_t cap);
void     orc_dec_fracmap(const uint8_t *in, uint32_t len, uint8_t frac[4096]);
uint64_t orc_xxh64(const void *data, size_t len, uint64_t seed);
uint64_t orc_blob_hash(const uint8_t *blob, uint64_t blob_len);

/* ---- chunk blobs (format v7) ---------------------------------------------- */
/* Encode a dense 256^3 chunk into a v7 blob. Returns blob length, 0 if the
 * chunk is all air.  cap must be large enough (17 MiB is always enough). */
size_t   orc_encode_chunk(float q, int tau, const uint8_t *chunk, uint8_t *out, size_t cap);
/* Decode a v7 blob into a dense 256^3 chunk (absent blocks -> 0). blob_len
 * bounds every payload read like the archive's live cursor does. */
void     orc_decode_chunk(const uint8_t *blob, uint64_t blob_len, uint8_t *out);
/* Decode one block (bz,by,bx) of a v7 blob into dst (absent -> 0). */
void     orc_decode_blob_block(const uint8_t *blob, uint64_t blob_len, int bz, int by, int bx,
                               uint8_t dst[4096]);
/* Total v7 blob length from its own header/bitmap/lens. */
uint64_t orc_blob_len(const uint8_t *blob);

/* ---- multi-threaded CPU baseline ----------------------------------------- */
/* Decode n blobs into n dense chunks with nthreads workers over (chunk,block).*/
void     orc_decode_chunks_mt(int n, const uint8_t *const *blobs, const uint64_t *lens,
                              uint8_t *out, int nthreads);
/* Encode n chunks (q[i], tau[i]) into blobs; out_blobs[i] capacity caps[i];
 * writes blob lengths.  Workers over chunks. */
void     orc_encode_chunks_mt(int n, const uint8_t *chunks, const float *q, const int *tau,
                              uint8_t *const *out_blobs, const size_t *caps, size_t *out_lens,
                              int nthreads);

/* ---- LOD decimation and quad render (mc_oracle_render.c) ----------------- */
/* 2x decimation: mean of the nonzero voxels of each 2x2x2 cell, rounded half
 * up; all-zero cells stay 0; odd dims round up (src/mc_archive.c:61-72).
 * out holds ((dz+1)/2) x ((dy+1)/2) x ((dx+1)/2) voxels. */
void     orc_decimate(const uint8_t *src, int dz, int dy, int dx, uint8_t *out);

enum { ORC_FILTER_NEAREST = 0, ORC_FILTER_TRILINEAR = 1 };
enum { ORC_COMP_NONE = 0, ORC_COMP_MIN = 1, ORC_COMP_MEAN = 2, ORC_COMP_MAX = 3, ORC_COMP_ALPHA = 4,
       ORC_COMP_STDDEV = 5, ORC_COMP_PERCENTILE = 7, ORC_COMP_DEPTH = 8 };

/* One level of a pyramid: a dense z,y,x raster, or (dense == NULL) 16^3
 * blocks at blk_data + blk_off[(bz*gby + by)*gbx + bx] (-1 = absent, reads 0)
 * with gby = ceil(ny/16), gbx = ceil(nx/16). */
typedef struct {
    const uint8_t *dense;
    const int64_t *blk_off;
    const uint8_t *blk_data;
    int nz, ny, nx;
} orc_level;

/* Point and unit normal of pixel (row, j) of a quad render. */
void     orc_quad_point(const float *grid, int gw, int gh, float x0, float y0, float step, int row, int j,
                        float *P, float *N);
int      orc_pick_lod(int nlods, float vox_per_pixel);
float    orc_quad_spacing(const float *grid, int gw, int gh);
/* Ray steps of a [t0, t1] walk at dt (make_cfg, src/mc_render.c:827-835). */
int      orc_render_steps(float t0, float t1, float dt);
/* mc_render_quad_lod for filters NEAREST/TRILINEAR and composites
 * NONE/MIN/MEAN/MAX/ALPHA/STDDEV/PERCENTILE/DEPTH (other render parameters
 * zero: the defaults); returns the level rendered, or -1. */
int      orc_render_quad_lod(int nlods, const orc_level *levels, const float *grid, int gw, int gh, float x0,
                             float y0, float step, int w, int h, int filter, int comp, float t0, float t1,
                             float dt, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
