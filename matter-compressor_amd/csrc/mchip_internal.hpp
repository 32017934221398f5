// mchip_internal.hpp -- host/device shared records of the HIP codec library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "mchip_device.hpp"

namespace mchip {


// One chunk of a decode batch.
struct ChunkDesc {
    uint64_t blob_off;      // byte offset of the blob in the device blob buffer
    uint64_t blob_len;      // readable bytes of the blob
    uint64_t out_off;       // byte offset of the 256^3 output raster
    uint64_t scratch_off;   // word offset of the chunk's coefficient-list slots
    uint32_t qtab;          // index of the chunk's quality table
    uint32_t pad;
};

// Chunk descriptors as the decode kernels see them: a device array, or (for
// launches of at most kInlineDesc chunks, the one-chunk-per-call path) the
// descriptors themselves inside the kernel arguments, so that such a call
// needs no host-to-device copy before its kernels.
constexpr int kInlineDesc = 4;
struct DescRef {
    const ChunkDesc *ptr;   // null: use inl
    ChunkDesc inl[kInlineDesc];
    __device__ __forceinline__ ChunkDesc operator[](int i) const { return ptr ? ptr[i] : inl[i]; }
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
    DescRef cd;
    const QTab *qtabs;
    BlkRec *rec;            // [n*4096]
    uint16_t *plist;        // [n*4096]
    ChunkInfo *cinfo;       // [n]
    uint32_t *scratch;      // coefficient lists
    uint2 *binfo;           // [n*4096]
    uint8_t *out;
    int n;
    int decoder;            // entropy decoder variant (tuning: 0 position, 2 lane, 3 lane in assembly)
    int persist_win;        // payload window words per lane of the persistent launch (16, 32 or 64)
    // persistent entropy launch (lane decoder): nwg workgroups take 64-block
    // groups from order[] (longest payloads first, over all chunks) through
    // the counter at pctl[0]; pctl[1] holds the number of groups
    int persistent;
    int nwg;
    int len_shift;          // k_index lane order: payload-length buckets of 1 << len_shift bytes
    int group_shift;        // k_group_order: first-block length buckets of 1 << group_shift bytes
    int recon_bpw;          // k_recon blocks per workgroup: 1, 2 or 4
    uint32_t *order;        // [n*64]
    uint32_t *pctl;         // [2]
    int slices;             // chunk slices of a batched persistent decode (1: none)
};
constexpr int kMaxSlices = 32;

// ---- encode ----------------------------------------------------------------
constexpr int kSlotBytes = 17408;          // per-block payload cap (src/mc_codec.c:442, :732)
constexpr int kFmapCap = 4096 / 2 + 64;    // fraction-map buffer (src/mc_archive.c:884-887)
constexpr int kFmapStride = 2176;          // per-chunk fraction-map record: u32 length + bytes
constexpr uint32_t kEncSlotOverflow = 1;   // a block payload exceeded kSlotBytes
constexpr uint32_t kEncOutTooSmall = 2;    // blob longer than the caller's capacity

struct EncChunkDesc {
    uint64_t in_off;        // byte offset of the 256^3 raster (16-byte aligned)
    uint64_t out_off;       // byte offset of the blob in the output buffer
    uint64_t out_cap;       // bytes available for the blob
    uint32_t qtab;          // quality table slot (sanitised q)
    int32_t tau;            // max-error bound, 0 = no corrections
    float q_store;          // quality written into the blob header
    uint32_t pad;
};

struct EncChunkResult {
    uint64_t blob_len;      // 0 = all-air chunk (no blob)
    uint32_t status;
    uint32_t npresent;
    uint32_t pay0;
    uint32_t pad;
};

struct EncodeArgs {
    const uint8_t *in;
    const EncChunkDesc *cd;
    const QTab *qtabs;
    uint4 *einfo;           // [n*4096]
    int16_t *levels;        // [n*4096*4096] scan-ordered levels
    uint32_t *corr;         // [n*4096*4096] corrections (pos | delta << 16)
    uint32_t *slots;        // [n*4096*kSlotBytes/4]
    uint32_t *plen;         // [n*4096]
    uint32_t *dst;          // [n*4096] payload offset in the blob
    uint32_t *fmap;         // [n*kFmapStride/4] fraction-map records
    uint8_t *out;
    EncChunkResult *res;    // [n]
    int n;
    const uint32_t *order;  // [n] chunk order of the entropy pass (null: 0..n-1)
    uint16_t *lanes;        // [n*4096] block order of the entropy pass within each chunk, costliest
                            // first (written by k_enc_lane_order; null: raster order)
    int slices;             // chunk slices of the sequence on two streams (1: none)
};

void launch_enc_transform(const EncodeArgs &a, hipStream_t s);
void launch_enc_entropy(const EncodeArgs &a, hipStream_t s);
void launch_enc_assemble(const EncodeArgs &a, hipStream_t s);
void launch_enc_copy(const EncodeArgs &a, hipStream_t s);
void launch_enc_hash(const EncodeArgs &a, hipStream_t s);
// XXH64 (mchip_hash.hip): blob hash + stored-hash check of n blobs at
// base + off[i] (len[i] bytes), and plain XXH64 of n byte ranges.
void launch_blob_hash(const uint8_t *base, const uint64_t *off, const uint64_t *len, int n, uint64_t *out,
                      uint8_t *bad, hipStream_t s);
void launch_xxh64_ranges(const uint8_t *base, const uint64_t *off, const uint64_t *len, const uint64_t *seed, int n,
                         uint64_t *out, hipStream_t s);
void launch_enc_blocks(const EncodeArgs &a, hipStream_t s);

// ---- block jobs (region reads, decoded-block arena) --------------------------
struct BlockJob {
    uint64_t pay_off;       // payload offset in the payload buffer
    uint64_t scratch_off;   // word offset of this job's list slot (8*len + 32 words)
    uint32_t len;           // payload length (0: decoded by the single-lane path)
    uint32_t qtab;
    uint32_t slot;          // destination arena slot (4096 bytes each)
    uint32_t pad;
};
void launch_jobs(const uint8_t *base, const BlockJob *jobs, int n, const QTab *qtabs, uint32_t *scratch,
                 uint2 *binfo, uint8_t *arena, int decoder, hipStream_t s);
void launch_crop(const uint8_t *arena, const int32_t *bmap, const int64_t *origin, int n, int gz, int gy, int gx,
                 int64_t dz, int64_t dy, int64_t dx, uint8_t *out, uint64_t batch_stride, hipStream_t s);

// ---- quad render and LOD decimation (mchip_render.hip) -----------------------
// One quad render at one level (src/mc_render.c:1981-2052, :2255-2314).
struct RenderDesc {
    const float *grid;      // gw*gh*3 control points (z,y,x), device
    int gw, gh;
    float x0, y0, step;     // pixel (i,j) maps to grid (x0 + j*step, y0 + i*step)
    int w, h;
    int filter, comp;       // 0 nearest / 1 trilinear; 0 none, 1 min, 2 mean, 3 max, 4 alpha, 5 stddev,
                            // 7 percentile, 8 depth
    float a_th, a_sc;       // alpha threshold and scale
    float t0, dt;           // first ray offset and step along the normal (level units)
    int nsteps;
    int lod;                // rendered level; > 0: points remapped by P*a + b, clamped at 0
    float a, b;
    int nz, ny, nx;         // voxel dims of the rendered level
};
// A level read from decoded blocks: table[(bz-obz)*gy*gx + ...] = arena slot
// of block (bz,by,bx), -1 absent; blocks outside the gz*gy*gx grid read 0.
struct TableLevel {
    const int32_t *table;
    const uint8_t *arena;
    int obz, oby, obx;
    int gz, gy, gx;
};
// Host side: level choice, step count and descriptor of a render
// (src/mc_render.c:827-835, :2222-2250, :2276-2288).  dims: nz,ny,nx per
// level.  Returns the level, or -1 for an unsupported filter / composite.
int render_describe(RenderDesc &d, int nlods, const int32_t *dims, const float *grid_host, const float *grid_dev,
                    int gw, int gh, float x0, float y0, float step, int w, int h, int filter, int comp, float t0,
                    float t1, float dt);
void launch_render_dense(const RenderDesc &d, const uint8_t *vol, uint8_t *out, hipStream_t s);
void launch_render_table(const RenderDesc &d, const TableLevel &t, uint8_t *out, hipStream_t s);
// Marks (bits over the gz*gy*gx grid of t) every block the render reads.
void launch_render_touch(const RenderDesc &d, const TableLevel &t, uint32_t *bits, hipStream_t s);
// Compacts a bitmap of nbits into block indices; count must start at 0.
void launch_bits_list(const uint32_t *bits, uint32_t nbits, uint32_t *list, uint32_t *count, hipStream_t s);
void launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
// table[upd[i].x] = upd[i].y
void launch_table_set(int32_t *table, const int2 *upd, int n, hipStream_t s);
// 2x decimation of a dense dz*dy*dx raster (src/mc_archive.c:61-72), and of
// a 2x2x2 group of 256^3 chunks (null: all air) into one 256^3 chunk.
void launch_decimate(const uint8_t *src, int dz, int dy, int dx, uint8_t *out, hipStream_t s);
struct ChunkGroup {
    const uint8_t *c[8];    // (z,y,x) order: index (gz*2 + gy)*2 + gx
};
void launch_decimate_group(const ChunkGroup &g, uint8_t *out, hipStream_t s);

// Decode-side deblocking of a dense nz*ny*nx raster in place
// (src/mc_codec.c:561-595); three passes on stream s.
void launch_deblock(uint8_t *v, int nz, int ny, int nx, float quality, hipStream_t s);

// Scratch words a chunk needs: 8 per payload byte + 32 per block.
inline uint64_t scratch_words_for(uint64_t blob_len) { return 8ull * blob_len + 32ull * 4096ull; }

hipError_t upload_constants(const float *cmf, const uint16_t *band_start, const uint16_t *scan);
// The fixed workspace of mchip_decode_chunks_static (a file-scope device
// array of kStaticScratchBytes in the kernels' translation unit): its device
// address on the current device, and a launch that writes n (<= kInlineDesc)
// quality tables passed in the kernel arguments to dst.
constexpr size_t kStaticScratchBytes = 128ull << 20;
hipError_t static_scratch_address(void **p);
void launch_put_qtabs(const QTab *tabs, int n, QTab *dst, hipStream_t s);
void launch_index(const DecodeArgs &a, hipStream_t s);
hipError_t occupancy_persist(int decoder, int win, int *per_cu);   // resident persistent-entropy workgroups per CU
void launch_entropy(const DecodeArgs &a, hipStream_t s);
void launch_recon(const DecodeArgs &a, hipStream_t s);

}  // namespace mchip
