/* mc_host_internal.h -- shared state of the plain-C host library. */
#ifndef MC_HOST_INTERNAL_H
#define MC_HOST_INTERNAL_H
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>

#include "matter_compressor_amd.h"
#include "mchip.h"

/* v7 .mca format constants (src/mc_archive.h:64-155) */
#define MCH_MAGIC 0x0043434Du
#define MCH_VERSION 7u
#define MCH_HDR 256u
#define MCH_META_END (128u * 1024u)
#define MCH_ROOTOFF 24
#define MCH_TOTLEN 88
#define MCH_PRIOROFF 128
#define MCH_PRIORS_MAGIC 0x53524950u
#define MCH_PRIORS_BYTES (8 + 2 * 8 * 32 * 2)
#define MCH_NODE_BYTES (4096u * 8u)
#define MCH_BLOB_HDR 14
#define MCH_SLOT_ZERO 1ull
#define MCH_HASH_SEED 0x6D636368756E6BULL
#define MCH_CHUNK_BYTES 16777216ull
#define MCH_SLOT_BYTES 17408

typedef struct {
    uint64_t co, tag;       /* chunk offset and stored blob hash */
    uint8_t frac[4096];
} mch_frac_entry;

/* The leading fields keep the order and types of the reference's private
 * struct (src/mc_archive.h:12-25), so that modules compiled against that
 * header (mc_cache.c and mc_volume.c read a->gen) see the same offsets when
 * linked with this library; the fields after gen are this library's own. */
struct mc_archive {
    int fd;
    uint8_t *base;                 /* fixed mmap base */
    _Atomic uint64_t cursor;       /* append EOF */
    _Atomic uint64_t file_len;     /* allocated file size */
    int dim;                       /* nx (legacy single-dimension field) */
    float quality;
    uint32_t block_codec;
    uint64_t reserve;              /* mmap reservation */
    pthread_mutex_t grow_mu;       /* file growth */
    pthread_mutex_t write_mu;      /* index mutation + install */
    _Atomic uint64_t *cov;         /* unused here (always NULL) */
    _Atomic uint64_t gen;          /* bumped on every publish */
    int nx, ny, nz;
    pthread_mutex_t aux_mu;        /* fraction-map cache */
    mch_frac_entry *fr_cache;
    uint32_t fr_cap;
};

mchip_ctx *mch_dev(void);
void mch_dev_error(const char *what);
uint64_t mch_xxh64(const void *data, size_t len, uint64_t seed);
uint64_t mch_blob_len_bounded(const uint8_t *arc, uint64_t off, uint64_t end);
uint64_t mch_resolve(const uint8_t *arc, uint64_t root, int cz, int cy, int cx, uint64_t end);
int mch_block_range(mc_archive *a, uint64_t co, int bi, uint64_t *off, uint32_t *len);

#endif
