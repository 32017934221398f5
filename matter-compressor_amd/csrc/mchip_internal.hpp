// mchip_internal.hpp -- host/device shared records of the HIP codec library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "mchip_device.hpp"

namespace mchip {

constexpr uint32_t kStatusOk = 0;
constexpr uint32_t kStatusFallback = 1;   // decode this block in k_recon, single lane

// One chunk of a decode batch.
struct ChunkDesc {
    uint64_t blob_off;      // byte offset of the blob in the device blob buffer
    uint64_t blob_len;      // readable bytes of the blob
    uint64_t out_off;       // byte offset of the 256^3 output raster
    uint64_t scratch_off;   // word offset of the chunk's coefficient-list slots
    uint32_t qtab;          // index of the chunk's quality table
    uint32_t pad;
};

// Per block of a chunk, written by k_index.  off == 0: absent (decodes to 0).
struct BlkRec {
    uint32_t off;           // payload offset inside the blob
    uint32_t len;           // payload length
    uint32_t slot;          // rank among present blocks
    uint32_t pad;
};

struct ChunkInfo {
    uint32_t np;            // number of present blocks
    uint32_t pay0;          // offset of the first payload inside the blob
};

struct DecodeArgs {
    const uint8_t *blobs;
    const ChunkDesc *cd;
    const QTab *qtabs;
    BlkRec *rec;            // [n*4096]
    uint16_t *plist;        // [n*4096]
    ChunkInfo *cinfo;       // [n]
    uint32_t *scratch;      // coefficient lists
    uint2 *binfo;           // [n*4096]
    uint8_t *out;
    int n;
};

// Scratch words a chunk needs: 8 per payload byte + 32 per block.
inline uint64_t scratch_words_for(uint64_t blob_len) { return 8ull * blob_len + 32ull * 4096ull; }

hipError_t upload_constants(const float *cmf, const uint16_t *band_start, const uint16_t *scan);
void launch_index(const DecodeArgs &a, hipStream_t s);
void launch_entropy(const DecodeArgs &a, hipStream_t s);
void launch_recon(const DecodeArgs &a, hipStream_t s);
void launch_block_single(const uint8_t *base, const uint64_t *off, const uint32_t *len, const uint32_t *qi,
                         const QTab *qtabs, uint8_t *out, int n, hipStream_t s);

}  // namespace mchip
