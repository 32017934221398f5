// mchip_serve.hip -- device-resident block store and serving loop (declared
// in include/mchip.h).
//
// The reference serves box batches through mc_cache, an S3-FIFO cache of
// decoded 16^3 blocks in host memory filled by CPU decodes
// (src/mc_cache.c:287-384, src/mc_archive.c:950-1042).  Here the compressed
// payloads stay resident in HBM (mchip_store), decoded blocks live in an HBM
// arena of 4 KiB slots, and each request runs on its own stream:
//
//   admit    (worker thread, host): for every distinct block the request's
//            boxes touch, look up its arena slot; misses take a slot by CLOCK
//            replacement among unpinned slots and become decode jobs; every
//            referenced slot is pinned until the request completes.  A block
//            that another request in flight is still decoding is not decoded
//            again: the stream waits for that request's decode event.
//   decode   k_entropy_jobs + k_recon_jobs write the misses into their slots.
//   assemble k_crop writes the boxes from the arena (absent / outside -> 0).
//
// Render requests (quad surfaces, src/mc_render.c:2302-2314) run a touch
// pass first: the render's own pixel walk, recording every block it reads
// at the chosen level; those blocks are then admitted like a box request's,
// and the render kernel reads them from the arena through a per-block slot
// table (each decode writes its block's slot there before the request's
// decode event).
//
// Requests are admitted in arrival order as lanes (streams) and arena slots
// free up, so requests that arrive while others run overlap on the device.
// The store may hold several levels (an LOD pyramid); box requests read
// level 0.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mchip.h"
#include "mchip_internal.hpp"
#include "mchip_runtime.hpp"

using namespace mchip;

namespace {
// One level of a store: a gz*gy*gx block grid whose block (0,0,0) is level
// block (obz, oby, obx); the level has nz*ny*nx voxels; its blocks have the
// store-wide indices base + (z*gy + y)*gx + x.
struct StoreLevel {
    int gz, gy, gx;
    int obz, oby, obx;
    int nz, ny, nx;
    uint64_t base;
};
}  // namespace

struct mchip_store {
    int device = 0;
    uint8_t *d_pay = nullptr;          // payload heap: bytes [0, nbytes) in use, cap allocated
    uint64_t nbytes = 0, cap = 0;
    std::vector<uint8_t *> old_pay;    // replaced heaps (decodes launched earlier still read them)
    int gz = 0, gy = 0, gx = 0;           // level 0 (box requests)
    std::vector<StoreLevel> lv;
    uint64_t nblocks = 0;
    std::vector<uint64_t> off;
    std::vector<uint32_t> len;
    std::vector<uint32_t> qslot;
};

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t ev_dec = nullptr, ev_done = nullptr;
    void *h_stage = nullptr; size_t cap_h = 0;          // pinned: jobs | origins | block map
    void *d_stage = nullptr; size_t cap_d = 0;          // device copy of the same
    void *d_scratch = nullptr; size_t cap_scratch = 0;  // coefficient lists of the jobs
    void *d_info = nullptr; size_t cap_info = 0;        // per-job block records
    void *d_bits = nullptr; size_t cap_bits = 0;        // render touch bitmap + count
    void *d_list = nullptr; size_t cap_list = 0;        // touched block list
    int64_t ticket = -1;                                // request in flight, -1: free
};

struct Req {
    int64_t ticket = 0;
    int kind = 0;         // 0 boxes, 1 render
    std::vector<int64_t> origin;
    int n = 0;
    int64_t dz = 0, dy = 0, dx = 0;
    // render: control grid (host copy), descriptor, blocks it reads
    std::vector<float> grid;
    RenderDesc rd{};
    bool touched_ok = false;
    std::vector<uint32_t> touched;
    uint8_t *d_out = nullptr;
    uint64_t stride = 0;
    double t_submit = 0.0, t_done = 0.0;
    int lane = -1;
    std::vector<int32_t> pinned;
    int state = 0;        // 0 pending, 1 in flight, 2 done, 3 failed
    bool deferred = false;
    std::string err;
};

// Lane buffers grow by half again when a request needs more.  A replaced
// buffer is not freed while the server runs (hipFree would wait for every
// stream of the device, stalling the other lanes); it is kept on a list and
// freed when the server is destroyed.
struct Graveyard {
    std::vector<void *> dev, host;
};

int grow_device(void **p, size_t *cap, size_t need, Graveyard &g) {
    if (need <= *cap) return 0;
    if (*p) g.dev.push_back(*p);
    *p = nullptr;
    *cap = 0;
    const size_t n = need + need / 2 + 4096;
    MCHIP_RT_CHECK(hipMalloc(p, n));
    *cap = n;
    return 0;
}

int grow_pinned(void **p, size_t *cap, size_t need, Graveyard &g) {
    if (need <= *cap) return 0;
    if (*p) g.host.push_back(*p);
    *p = nullptr;
    *cap = 0;
    const size_t n = need + need / 2 + 4096;
    MCHIP_RT_CHECK(hipHostMalloc(p, n, hipHostMallocDefault));
    *cap = n;
    return 0;
}

// The decoded-block arena: a device address range reserved once and backed
// by physical segments mapped on demand (HIP virtual memory), so that it
// grows while free device memory allows and shrinks on request without
// moving; kernels in flight keep valid addresses.  Without virtual-memory
// support it is one fixed allocation.
constexpr size_t kArenaSeg = 256ull << 20;       // 65536 slots per segment

struct Arena {
    uint8_t *base = nullptr;
    size_t reserved = 0;                         // address range
    size_t mapped = 0;                           // backed bytes
    bool vmm = false;
    int device = 0;
    std::vector<hipMemGenericAllocationHandle_t> seg;

    // budget: the most the arena may hold; initial: bytes backed now.
    int init(int dev, size_t budget, size_t initial) {
        device = dev;
        int vm = 0;
        if (hipDeviceGetAttribute(&vm, hipDeviceAttributeVirtualMemoryManagementSupported, dev) == hipSuccess && vm) {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) != hipSuccess) tot = 0;
            // the range covers the whole device, so that a resize can raise the budget
            const size_t span = (std::max(budget, tot) + kArenaSeg - 1) / kArenaSeg * kArenaSeg;
            void *p = nullptr;
            if (hipMemAddressReserve(&p, span, 0, nullptr, 0) == hipSuccess) {
                base = (uint8_t *)p;
                reserved = span;
                vmm = true;
                return grow(initial);
            }
        }
        MCHIP_RT_CHECK(hipMalloc((void **)&base, budget));
        reserved = mapped = budget;
        return 0;
    }
    // Back the arena up to want bytes (a multiple of kArenaSeg).
    int grow(size_t want) {
        if (!vmm) return want <= mapped ? 0 : rt_fail("arena: a fixed allocation of %zu bytes", mapped);
        if (want > reserved) return rt_fail("arena: %zu bytes exceed the reserved range", want);
        hipMemAllocationProp prop = {};
        prop.type = hipMemAllocationTypePinned;
        prop.location.type = hipMemLocationTypeDevice;
        prop.location.id = device;
        hipMemAccessDesc acc = {};
        acc.location = prop.location;
        acc.flags = hipMemAccessFlagsProtReadWrite;
        while (mapped < want) {
            hipMemGenericAllocationHandle_t h;
            MCHIP_RT_CHECK(hipMemCreate(&h, kArenaSeg, &prop, 0));
            if (hipMemMap(base + mapped, kArenaSeg, 0, h, 0) != hipSuccess) {
                (void)hipMemRelease(h);
                return rt_fail("arena: mapping a segment failed");
            }
            if (hipMemSetAccess(base + mapped, kArenaSeg, &acc, 1) != hipSuccess) {
                (void)hipMemUnmap(base + mapped, kArenaSeg);
                (void)hipMemRelease(h);
                return rt_fail("arena: enabling access to a segment failed");
            }
            seg.push_back(h);
            mapped += kArenaSeg;
        }
        return 0;
    }
    // Release the segments beyond want bytes; no work may use them.
    int shrink(size_t want) {
        while (vmm && mapped > want && !seg.empty()) {
            mapped -= kArenaSeg;
            MCHIP_RT_CHECK(hipMemUnmap(base + mapped, kArenaSeg));
            MCHIP_RT_CHECK(hipMemRelease(seg.back()));
            seg.pop_back();
        }
        return 0;
    }
    void release() {
        if (vmm) {
            (void)shrink(0);
            if (base) (void)hipMemAddressFree(base, reserved);
        } else if (base) {
            (void)hipFree(base);
        }
        base = nullptr;
    }
};

}  // namespace

struct mchip_server {
    mchip_store *st = nullptr;
    Arena arena;
    uint8_t *d_arena = nullptr;      // arena.base
    size_t budget = 0;               // bytes the arena may grow to
    size_t headroom = 2ull << 30;    // device memory the arena leaves free when it grows
    int32_t *d_table = nullptr;      // per block: arena slot, written by the decoding request
    uint32_t nslots = 0;             // usable slots (the backed part of the arena)
    std::vector<int32_t> freelist;   // free slots below nslots
    hipStream_t upd_stream = nullptr;    // uploads of updated payloads
    std::vector<int32_t> slot_of;    // per block: arena slot or -1
    std::vector<int32_t> blk_of;     // per slot: block or -1
    std::vector<int32_t> pins;       // per slot: requests in flight that reference it
    std::vector<uint8_t> refbit;     // per slot: CLOCK reference bit
    std::vector<int64_t> producer;   // per slot: ticket of the request decoding it, -1 when decoded
    std::vector<uint32_t> mark;      // per block: admission attempt that last saw it
    uint32_t gen = 0, hand = 0;
    std::vector<Lane> lanes;
    std::deque<Req *> pending;
    std::map<int64_t, Req *> reqs;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false;
    int64_t next_ticket = 0;
    mchip_server_stats stats{};
    Graveyard grave;

    // More usable slots: up to the budget, from the backed part of the
    // arena, mapping one more segment when the budget and the free device
    // memory (beyond the headroom) allow; the new slots join the free list.
    bool grow_slots() {
        const uint64_t cap_slots = std::min<uint64_t>(budget / kNVox, 0x7FFFFFFFull);
        if (nslots >= cap_slots) return false;
        if ((uint64_t)nslots * kNVox >= arena.mapped) {
            if (!arena.vmm) return false;
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr < kArenaSeg + headroom) return false;
            if (arena.grow(arena.mapped + kArenaSeg)) return false;
            stats.arena_grows++;
        }
        const uint32_t n0 = nslots, n1 = (uint32_t)std::min<uint64_t>(cap_slots, arena.mapped / kNVox);
        blk_of.resize(n1, -1);
        pins.resize(n1, 0);
        refbit.resize(n1, 0);
        producer.resize(n1, -1);
        for (uint32_t x = n1; x-- > n0;) freelist.push_back((int32_t)x);
        nslots = n1;
        stats.arena_slots = nslots;
        stats.arena_mapped_bytes = arena.mapped;
        return true;
    }

    // A free slot: from the free list, from a new segment, or by CLOCK
    // replacement among unpinned slots.  -1: every slot is pinned.
    int evict_one() {
        if (freelist.empty()) grow_slots();
        if (!freelist.empty()) {
            const int32_t x = freelist.back();
            freelist.pop_back();
            return x;
        }
        for (uint32_t k = 0; k < 2 * nslots + 1; ++k) {
            const uint32_t x = hand;
            hand = hand + 1 >= nslots ? 0 : hand + 1;
            if (pins[x] > 0 || blk_of[x] < 0) continue;
            if (refbit[x]) { refbit[x] = 0; continue; }
            slot_of[blk_of[x]] = -1;
            blk_of[x] = -1;
            stats.evictions++;
            return (int)x;
        }
        return -1;
    }

    // Drop one pin of slot x; a slot whose block was detached (updated or
    // retired while in use) becomes free when its last pin goes.
    void unpin_slot(int32_t x) {
        pins[x]--;
        if (pins[x] == 0 && blk_of[x] < 0 && (uint32_t)x < nslots) freelist.push_back(x);
    }

    // Release the request's pins; its decodes are complete.
    void unpin(Req *r) {
        for (int32_t x : r->pinned) {
            if (producer[x] == r->ticket) producer[x] = -1;
            unpin_slot(x);
        }
        r->pinned.clear();
    }
    // Release the request's pins and free the slots it was to decode (not
    // admitted, or its device work failed).
    void rollback(Req *r) {
        for (int32_t x : r->pinned) {
            if (producer[x] == r->ticket) {
                producer[x] = -1;
                if (blk_of[x] >= 0) {
                    slot_of[blk_of[x]] = -1;
                    blk_of[x] = -1;
                }
            }
            unpin_slot(x);
        }
        r->pinned.clear();
    }

    // Detach block id from its slot: later requests decode it again; the
    // slot is free now, or when the requests holding it finish.
    void detach(uint64_t id) {
        const int32_t x = slot_of[id];
        if (x < 0) return;
        slot_of[id] = -1;
        if (blk_of[x] == (int32_t)id) {
            blk_of[x] = -1;
            if (pins[x] == 0 && (uint32_t)x < nslots) freelist.push_back(x);
        }
    }

    // The blocks of one admission attempt.
    struct Admit {
        std::vector<BlockJob> jobs;
        std::vector<int2> upd;       // (block, slot) table entries of the jobs
        std::vector<int> deps;       // lanes whose decode events the request waits for
        uint64_t sw = 0, refs = 0, hits = 0;
    };

    // Arena slot of block id for request r: a resident block is pinned (and,
    // while another request in flight still decodes it, that request's lane
    // becomes a dependency); a missing block takes a slot by CLOCK and
    // becomes a decode job.  -1: no slot can be freed now.
    int32_t take(Req *r, uint64_t id, Admit &ad) {
        const mchip_store &S = *st;
        if (mark[id] == gen) return slot_of[id];
        mark[id] = gen;
        ad.refs++;
        int32_t s = slot_of[id];
        if (s >= 0) {
            ad.hits++;
            if (producer[s] >= 0 && producer[s] != r->ticket) {
                auto it = reqs.find(producer[s]);
                if (it != reqs.end() && it->second->lane >= 0) {
                    const int dl = it->second->lane;
                    bool have = false;
                    for (int x : ad.deps) have |= x == dl;
                    if (!have) ad.deps.push_back(dl);
                }
            }
        } else {
            s = evict_one();
            if (s < 0) return -1;
            blk_of[s] = (int32_t)id;
            slot_of[id] = s;
            producer[s] = r->ticket;
            ad.jobs.push_back(BlockJob{S.off[id], ad.sw, S.len[id], S.qslot[id], (uint32_t)s, 0});
            ad.upd.push_back(make_int2((int)id, s));
            ad.sw += 8ull * S.len[id] + 32ull;
        }
        pins[s]++;
        refbit[s] = 1;
        r->pinned.push_back(s);
        return s;
    }

    // Render touch pass on lane L: the blocks (level-local indices) the
    // render reads, into r->touched.  Synchronous on the lane's stream.
    int touch(Req *r, Lane &L) {
        const StoreLevel &lv = st->lv[r->rd.lod];
        const uint64_t nb = (uint64_t)lv.gz * lv.gy * lv.gx, words = (nb + 31) / 32;
        const size_t gb = r->grid.size() * sizeof(float);
        if (grow_pinned(&L.h_stage, &L.cap_h, std::max<size_t>(gb, 4 * nb + 16), grave) ||
            grow_device(&L.d_stage, &L.cap_d, gb, grave) ||
            grow_device(&L.d_bits, &L.cap_bits, 4 * words + 16, grave) ||
            grow_device(&L.d_list, &L.cap_list, 4 * nb + 16, grave))
            return -1;
        memcpy(L.h_stage, r->grid.data(), gb);
        MCHIP_RT_CHECK(hipMemcpyAsync(L.d_stage, L.h_stage, gb, hipMemcpyHostToDevice, L.stream));
        uint32_t *bits = (uint32_t *)L.d_bits;
        launch_fill_u32(bits, 0, words + 1, L.stream);
        RenderDesc d = r->rd;
        d.grid = (const float *)L.d_stage;
        const TableLevel t{nullptr, nullptr, lv.obz, lv.oby, lv.obx, lv.gz, lv.gy, lv.gx};
        launch_render_touch(d, t, bits, L.stream);
        launch_bits_list(bits, (uint32_t)nb, (uint32_t *)L.d_list, bits + words, L.stream);
        MCHIP_RT_CHECK(hipGetLastError());
        MCHIP_RT_CHECK(hipMemcpyAsync(L.h_stage, bits + words, 4, hipMemcpyDeviceToHost, L.stream));
        MCHIP_RT_CHECK(hipStreamSynchronize(L.stream));
        const uint32_t n = *(const uint32_t *)L.h_stage;
        if (n > nb) return rt_fail("render touch pass: %u blocks of %llu", n, (unsigned long long)nb);
        if (n) {
            MCHIP_RT_CHECK(hipMemcpyAsync(L.h_stage, L.d_list, 4ull * n, hipMemcpyDeviceToHost, L.stream));
            MCHIP_RT_CHECK(hipStreamSynchronize(L.stream));
        }
        r->touched.assign((const uint32_t *)L.h_stage, (const uint32_t *)L.h_stage + n);
        std::sort(r->touched.begin(), r->touched.end());
        r->touched_ok = true;
        return 0;
    }

    // 1 admitted and issued, 0 not enough free slots or memory now (r->err
    // says which), -1 error (r->err).
    int try_admit(Req *r, int li) {
        Lane &L = lanes[li];
        r->err.clear();
        const mchip_store &S = *st;
        if (r->kind == 1 && !r->touched_ok && touch(r, L)) {
            r->err = std::string("render touch pass failed: ") + mchip_last_error();
            return -1;
        }
        if (++gen == 0) { std::fill(mark.begin(), mark.end(), 0u); gen = 1; }
        Admit ad;
        bool ok = true;
        // boxes: the touched block grid of each box, as arena slots (-1 zeros)
        int gzb = 0, gyb = 0, gxb = 0;
        size_t cells_per = 0, cells = 0;
        std::vector<int32_t> bmap;
        if (r->kind == 0) {
            gzb = (int)((r->dz + 15) >> 4) + 1; gyb = (int)((r->dy + 15) >> 4) + 1; gxb = (int)((r->dx + 15) >> 4) + 1;
            cells_per = (size_t)gzb * gyb * gxb;
            cells = cells_per * r->n;
            bmap.assign(cells, -1);
            for (int b = 0; b < r->n && ok; ++b) {
                const int64_t z0 = r->origin[3 * b], y0 = r->origin[3 * b + 1], x0 = r->origin[3 * b + 2];
                const int64_t bz0 = z0 >> 4, by0 = y0 >> 4, bx0 = x0 >> 4;
                const int64_t nbz = ((z0 + r->dz - 1) >> 4) - bz0 + 1, nby = ((y0 + r->dy - 1) >> 4) - by0 + 1,
                              nbx = ((x0 + r->dx - 1) >> 4) - bx0 + 1;
                for (int64_t iz = 0; iz < nbz && ok; ++iz)
                    for (int64_t iy = 0; iy < nby && ok; ++iy)
                        for (int64_t ix = 0; ix < nbx; ++ix) {
                            const int64_t bz = bz0 + iz, by = by0 + iy, bx = bx0 + ix;
                            if (bz < 0 || by < 0 || bx < 0 || bz >= S.gz || by >= S.gy || bx >= S.gx) continue;
                            const uint64_t id = (uint64_t)((bz * S.gy + by) * S.gx + bx);
                            if (!S.len[id]) continue;
                            const int32_t s = take(r, id, ad);
                            if (s < 0) { ok = false; break; }
                            bmap[(size_t)b * cells_per + (size_t)((iz * gyb + iy) * gxb + ix)] = s;
                        }
            }
        } else {
            const uint64_t base = S.lv[r->rd.lod].base;
            for (uint32_t loc : r->touched) {
                const uint64_t id = base + loc;
                if (!S.len[id]) continue;
                if (take(r, id, ad) < 0) { ok = false; break; }
            }
        }
        if (!ok) {                         // not enough unpinned slots now
            rollback(r);
            char buf[200];
            snprintf(buf, sizeof buf,
                     "serving request: the arena's %u slots cannot hold the request (arena budget %zu bytes, "
                     "device memory left free %zu bytes)", nslots, budget, headroom);
            r->err = buf;
            return 0;
        }
        // staging: [grid (render)] [jobs] [table entries] [origins | block map (boxes)]
        auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
        const size_t gb = r->kind == 1 ? r->grid.size() * sizeof(float) : 0;
        const size_t jb = sizeof(BlockJob) * ad.jobs.size(), ub = sizeof(int2) * ad.upd.size();
        const size_t ob = r->kind == 0 ? 24 * (size_t)r->n : 0, mb = 4 * cells;
        const size_t o_job = al(gb), o_upd = al(o_job + jb), o_org = al(o_upd + ub), o_map = al(o_org + ob);
        const size_t total = o_map + mb;
        auto fail_req = [&](void) {
            r->err = "serving request: device allocation or launch failed";
            rollback(r);
            return -1;
        };
        if (grow_pinned(&L.h_stage, &L.cap_h, total, grave) || grow_device(&L.d_stage, &L.cap_d, total, grave) ||
            grow_device(&L.d_scratch, &L.cap_scratch, sizeof(uint32_t) * (ad.sw + 64), grave) ||
            grow_device(&L.d_info, &L.cap_info, sizeof(uint2) * (ad.jobs.size() + 1), grave)) {
            // memory is short: wait for requests in flight to finish (the
            // caller fails the request only when none is in flight)
            rollback(r);
            r->err = std::string("serving request: device or pinned memory is short: ") + mchip_last_error();
            return 0;
        }
        uint8_t *h = (uint8_t *)L.h_stage;
        if (gb) memcpy(h, r->grid.data(), gb);
        if (jb) memcpy(h + o_job, ad.jobs.data(), jb);
        if (ub) memcpy(h + o_upd, ad.upd.data(), ub);
        if (ob) memcpy(h + o_org, r->origin.data(), ob);
        if (mb) memcpy(h + o_map, bmap.data(), mb);
        uint8_t *d = (uint8_t *)L.d_stage;
        if (hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, L.stream) != hipSuccess) return fail_req();
        for (int dl : ad.deps)
            if (hipStreamWaitEvent(L.stream, lanes[dl].ev_dec, 0) != hipSuccess) return fail_req();
        if (!ad.jobs.empty()) {
            launch_jobs(S.d_pay, (const BlockJob *)(d + o_job), (int)ad.jobs.size(), rt_qtabs(),
                        (uint32_t *)L.d_scratch, (uint2 *)L.d_info, d_arena, rt_decoder(), L.stream);
            launch_table_set(d_table, (const int2 *)(d + o_upd), (int)ad.upd.size(), L.stream);
        }
        if (hipEventRecord(L.ev_dec, L.stream) != hipSuccess) return fail_req();
        if (r->kind == 0) {
            launch_crop(d_arena, (const int32_t *)(d + o_map), (const int64_t *)(d + o_org), r->n, gzb, gyb, gxb,
                        r->dz, r->dy, r->dx, r->d_out, r->stride, L.stream);
        } else {
            const StoreLevel &lv = S.lv[r->rd.lod];
            RenderDesc rd = r->rd;
            rd.grid = (const float *)d;
            const TableLevel t{d_table + lv.base, d_arena, lv.obz, lv.oby, lv.obx, lv.gz, lv.gy, lv.gx};
            launch_render_table(rd, t, r->d_out, L.stream);
        }
        if (hipGetLastError() != hipSuccess || hipEventRecord(L.ev_done, L.stream) != hipSuccess) return fail_req();
        L.ticket = r->ticket;
        r->lane = li;
        stats.block_refs += ad.refs;
        stats.hits += ad.hits;
        stats.misses += ad.jobs.size();
        return 1;
    }

    void retire() {
        for (Lane &L : lanes) {
            if (L.ticket < 0) continue;
            const hipError_t e = hipEventQuery(L.ev_done);
            if (e == hipErrorNotReady) continue;
            Req *r = reqs[L.ticket];
            r->t_done = now_ms();
            r->lane = -1;
            if (e == hipSuccess) {
                unpin(r);
                r->state = 2;
            } else {
                rollback(r);
                r->state = 3;
                r->err = std::string("serving request: ") + hipGetErrorString(e);
            }
            stats.completed++;
            L.ticket = -1;
            cv_done.notify_all();
        }
    }

    void run() {
        (void)hipSetDevice(st->device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            retire();
            while (!pending.empty()) {
                int li = -1, busy = 0;
                for (int i = 0; i < (int)lanes.size(); ++i) {
                    if (lanes[i].ticket < 0) { if (li < 0) li = i; } else busy++;
                }
                Req *r = pending.front();
                if (li < 0) {
                    if (!r->deferred) { r->deferred = true; stats.deferred++; }
                    break;
                }
                const int rc = try_admit(r, li);
                if (rc == 1) {
                    r->state = 1;
                    pending.pop_front();
                } else if (rc == 0 && busy > 0) {
                    if (!r->deferred) { r->deferred = true; stats.deferred++; }
                    break;                      // wait for requests in flight to release slots
                } else {
                    if (rc == 0 && r->err.empty())
                        r->err = "serving request: the arena holds fewer slots than one request needs";
                    r->state = 3;
                    r->t_done = now_ms();
                    pending.pop_front();
                    cv_done.notify_all();
                }
            }
            bool busy = false;
            for (Lane &L : lanes) busy |= L.ticket >= 0;
            if (!busy && pending.empty()) {
                if (stop) break;
                cv_work.wait(lk);
                continue;
            }
            // requests in flight: poll their completion events (the worker
            // thread spins with yields so completions are seen within
            // microseconds)
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        }
    }
};

extern "C" {

mchip_store *mchip_store_create_levels(const uint8_t *payload, uint64_t nbytes, int nlevels, const int32_t *geom,
                                       const uint64_t *off, const uint32_t *len, const float *q) {
    if (nlevels < 1 || nlevels > 16 || !geom || !off || !len || !q || (nbytes && !payload)) {
        rt_fail("mchip_store_create: bad argument");
        return nullptr;
    }
    auto *st = new mchip_store;
    uint64_t nb = 0;
    for (int l = 0; l < nlevels; ++l) {
        const int32_t *g = geom + 9 * l;
        if (g[0] <= 0 || g[1] <= 0 || g[2] <= 0 || g[6] <= 0 || g[7] <= 0 || g[8] <= 0) {
            delete st;
            rt_fail("mchip_store_create: level %d has an empty grid or volume", l);
            return nullptr;
        }
        st->lv.push_back(StoreLevel{g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], g[8], nb});
        nb += (uint64_t)g[0] * g[1] * g[2];
    }
    if (nb > (1ull << 30)) {
        delete st;
        rt_fail("mchip_store_create: %llu blocks", (unsigned long long)nb);
        return nullptr;
    }
    for (uint64_t i = 0; i < nb; ++i)
        if (len[i] && (off[i] > nbytes || len[i] > nbytes - off[i])) {
            delete st;
            rt_fail("mchip_store_create: block %llu lies outside the payload bytes", (unsigned long long)i);
            return nullptr;
        }
    st->gz = st->lv[0].gz; st->gy = st->lv[0].gy; st->gx = st->lv[0].gx;
    st->nblocks = nb;
    st->off.assign(off, off + nb);
    st->len.assign(len, len + nb);
    st->qslot.assign(nb, 0);
    std::map<uint32_t, uint32_t> qs;              // q bit pattern -> table slot
    for (uint64_t i = 0; i < nb; ++i) {
        if (!len[i]) continue;
        uint32_t key;
        memcpy(&key, &q[i], 4);
        auto it = qs.find(key);
        if (it == qs.end()) {
            uint32_t slot;
            if (rt_qtab_slot(q[i], &slot)) { delete st; return nullptr; }
            it = qs.emplace(key, slot).first;
        }
        st->qslot[i] = it->second;
    }
    const uint64_t cap = nbytes + nbytes / 4 + (64ull << 20);       // room for updated chunks
    if (hipGetDevice(&st->device) != hipSuccess ||
        hipMalloc((void **)&st->d_pay, cap) != hipSuccess ||
        (nbytes && hipMemcpy(st->d_pay, payload, nbytes, hipMemcpyHostToDevice) != hipSuccess)) {
        rt_fail("mchip_store_create: device allocation or upload of %llu bytes failed",
                (unsigned long long)nbytes);
        if (st->d_pay) (void)hipFree(st->d_pay);
        delete st;
        return nullptr;
    }
    st->nbytes = (nbytes + 255) & ~255ull;
    st->cap = cap;
    return st;
}

mchip_store *mchip_store_create(const uint8_t *payload, uint64_t nbytes, const uint64_t *off, const uint32_t *len,
                                const float *q, int gz, int gy, int gx) {
    if (gz <= 0 || gy <= 0 || gx <= 0 || gz > (1 << 26) || gy > (1 << 26) || gx > (1 << 26)) {
        rt_fail("mchip_store_create: bad argument");
        return nullptr;
    }
    const int32_t geom[9] = {gz, gy, gx, 0, 0, 0, 16 * gz, 16 * gy, 16 * gx};
    return mchip_store_create_levels(payload, nbytes, 1, geom, off, len, q);
}

void mchip_store_destroy(mchip_store *st) {
    if (!st) return;
    if (st->d_pay) (void)hipFree(st->d_pay);
    for (uint8_t *p : st->old_pay) (void)hipFree(p);
    delete st;
}

uint64_t mchip_store_device_bytes(const mchip_store *st) { return st ? st->nbytes : 0; }

mchip_server *mchip_server_create(mchip_store *st, uint64_t arena_bytes, int lanes) {
    if (!st || lanes <= 0 || lanes > 64) { rt_fail("mchip_server_create: bad argument"); return nullptr; }
    (void)hipSetDevice(st->device);
    if (!arena_bytes) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) { rt_fail("hipMemGetInfo failed"); return nullptr; }
        arena_bytes = fr / 2;
    }
    const uint64_t ns = arena_bytes / kNVox;
    if (ns < 1 || ns > 0x7FFFFFFFull) { rt_fail("mchip_server_create: arena of %llu bytes", (unsigned long long)arena_bytes); return nullptr; }
    auto *sv = new mchip_server;
    sv->st = st;
    sv->budget = ns * kNVox;
    // back one segment now; the arena grows by segments as requests need
    // slots, while the budget and the free device memory allow
    const size_t initial = kArenaSeg;
    if (sv->arena.init(st->device, sv->budget, initial)) {
        rt_fail("mchip_server_create: arena of %llu bytes: %s", (unsigned long long)sv->budget, mchip_last_error());
        delete sv;
        return nullptr;
    }
    sv->d_arena = sv->arena.base;
    const size_t nb = (size_t)st->nblocks;
    if (hipMalloc((void **)&sv->d_table, 4 * nb + 4) != hipSuccess ||
        hipMemset(sv->d_table, 0xFF, 4 * nb + 4) != hipSuccess ||
        hipStreamCreateWithFlags(&sv->upd_stream, hipStreamNonBlocking) != hipSuccess) {
        rt_fail("mchip_server_create: slot table allocation failed");
        sv->arena.release();
        if (sv->d_table) (void)hipFree(sv->d_table);
        delete sv;
        return nullptr;
    }
    sv->slot_of.assign(nb, -1);
    sv->mark.assign(nb, 0u);
    sv->grow_slots();
    sv->stats.arena_budget_bytes = sv->budget;
    sv->lanes.resize(lanes);
    for (Lane &L : sv->lanes) {
        // initial lane buffers for a batch of 64 boxes of 128^3 (64 * 9^3
        // touched blocks) of typical payloads; they grow when needed
        const size_t jobs = 64 * 729;
        if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&L.ev_dec, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&L.ev_done, hipEventDisableTiming) != hipSuccess ||
            grow_pinned(&L.h_stage, &L.cap_h, jobs * (sizeof(BlockJob) + 4) + 64 * 24, sv->grave) ||
            grow_device(&L.d_stage, &L.cap_d, jobs * (sizeof(BlockJob) + 4) + 64 * 24, sv->grave) ||
            grow_device(&L.d_scratch, &L.cap_scratch, jobs * (8ull * 320 + 32) * 4, sv->grave) ||
            grow_device(&L.d_info, &L.cap_info, jobs * sizeof(uint2), sv->grave)) {
            rt_fail("mchip_server_create: lane stream, event or buffer creation failed");
            mchip_server_destroy(sv);
            return nullptr;
        }
    }
    sv->worker = std::thread([sv] { sv->run(); });
    return sv;
}

int64_t mchip_server_submit(mchip_server *sv, const int64_t *origins, int n, int64_t dz, int64_t dy, int64_t dx,
                            uint8_t *d_out, uint64_t batch_stride) {
    if (!sv || !origins || n <= 0 || dz <= 0 || dy <= 0 || dx <= 0 || !d_out)
        return rt_fail("mchip_server_submit: bad argument");
    if (batch_stride < (uint64_t)(dz * dy * dx) && n > 1)
        return rt_fail("mchip_server_submit: batch_stride smaller than a box");
    auto *r = new Req;
    r->origin.assign(origins, origins + 3 * (size_t)n);
    r->n = n; r->dz = dz; r->dy = dy; r->dx = dx;
    r->d_out = d_out;
    r->stride = batch_stride;
    std::lock_guard<std::mutex> lk(sv->mu);
    r->ticket = sv->next_ticket++;
    r->t_submit = now_ms();
    sv->reqs[r->ticket] = r;
    sv->pending.push_back(r);
    sv->stats.requests++;
    sv->stats.boxes += (uint64_t)n;
    sv->cv_work.notify_one();
    return r->ticket;
}

int64_t mchip_server_submit_render(mchip_server *sv, const float *grid, int gw, int gh, float x0, float y0,
                                   float step, int w, int h, int filter, int comp, float t0, float t1, float dt,
                                   uint8_t *d_out) {
    if (!sv || !grid || gw < 1 || gh < 1 || w <= 0 || h <= 0 || !d_out || (int64_t)w * h > (1ll << 30))
        return rt_fail("mchip_server_submit_render: bad argument");
    auto *r = new Req;
    r->kind = 1;
    r->grid.assign(grid, grid + 3 * (size_t)gw * gh);
    std::vector<int32_t> dims;
    for (const StoreLevel &l : sv->st->lv) { dims.push_back(l.nz); dims.push_back(l.ny); dims.push_back(l.nx); }
    if (render_describe(r->rd, (int)sv->st->lv.size(), dims.data(), r->grid.data(), nullptr, gw, gh, x0, y0, step,
                        w, h, filter, comp, t0, t1, dt) < 0) {
        delete r;
        return rt_fail("mchip_server_submit_render: filter %d / composite %d not supported", filter, comp);
    }
    r->d_out = d_out;
    std::lock_guard<std::mutex> lk(sv->mu);
    r->ticket = sv->next_ticket++;
    r->t_submit = now_ms();
    sv->reqs[r->ticket] = r;
    sv->pending.push_back(r);
    sv->stats.requests++;
    sv->stats.renders++;
    sv->cv_work.notify_one();
    return r->ticket;
}

int mchip_server_render_level(mchip_server *sv, int64_t ticket) {
    if (!sv) return rt_fail("null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end() || it->second->kind != 1) return rt_fail("not a pending render ticket");
    return it->second->rd.lod;
}

int mchip_server_poll(mchip_server *sv, int64_t ticket) {
    if (!sv) return rt_fail("null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return rt_fail("unknown ticket %lld", (long long)ticket);
    if (it->second->state == 3) return rt_fail("%s", it->second->err.c_str());
    return it->second->state == 2 ? 1 : 0;
}

int mchip_server_wait(mchip_server *sv, int64_t ticket, double *latency_ms) {
    if (!sv) return rt_fail("null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return rt_fail("unknown ticket %lld", (long long)ticket);
    Req *r = it->second;
    sv->cv_done.wait(lk, [r] { return r->state >= 2; });
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    const int rc = r->state == 2 ? 0 : rt_fail("%s", r->err.c_str());
    sv->reqs.erase(it);
    delete r;
    return rc;
}

int mchip_server_preload(mchip_server *sv, uint64_t *loaded) {
    if (!sv) return rt_fail("null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv_done.wait(lk, [sv] {                   // idle: nothing pending or in flight
        if (!sv->pending.empty()) return false;
        for (Lane &L : sv->lanes) if (L.ticket >= 0) return false;
        return true;
    });
    const mchip_store &S = *sv->st;
    Lane &L = sv->lanes[0];
    std::vector<BlockJob> jobs;
    std::vector<int32_t> pinned;
    uint64_t n = 0;
    auto unpin_all = [&]() {
        for (int32_t s : pinned) sv->unpin_slot(s);
    };
    auto drop_jobs = [&]() {                      // jobs not decoded: their slots hold nothing
        for (const BlockJob &j : jobs) {
            sv->slot_of[sv->blk_of[j.slot]] = -1;
            sv->blk_of[j.slot] = -1;
        }
        jobs.clear();
    };
    auto flush = [&]() -> int {
        if (jobs.empty()) return 0;
        uint64_t sw = 0;
        std::vector<int2> upd;
        for (BlockJob &j : jobs) {
            j.scratch_off = sw;
            sw += 8ull * j.len + 32ull;
            upd.push_back(make_int2(sv->blk_of[j.slot], (int)j.slot));
        }
        const size_t jb = sizeof(BlockJob) * jobs.size(), o_upd = (jb + 255) & ~size_t(255);
        const size_t total = o_upd + sizeof(int2) * upd.size();
        if (grow_pinned(&L.h_stage, &L.cap_h, total, sv->grave) ||
            grow_device(&L.d_stage, &L.cap_d, total, sv->grave) ||
            grow_device(&L.d_scratch, &L.cap_scratch, sizeof(uint32_t) * (sw + 64), sv->grave) ||
            grow_device(&L.d_info, &L.cap_info, sizeof(uint2) * (jobs.size() + 1), sv->grave))
            return -1;
        memcpy(L.h_stage, jobs.data(), jb);
        memcpy((uint8_t *)L.h_stage + o_upd, upd.data(), sizeof(int2) * upd.size());
        MCHIP_RT_CHECK(hipMemcpyAsync(L.d_stage, L.h_stage, total, hipMemcpyHostToDevice, L.stream));
        launch_jobs(S.d_pay, (const BlockJob *)L.d_stage, (int)jobs.size(), rt_qtabs(), (uint32_t *)L.d_scratch,
                    (uint2 *)L.d_info, sv->d_arena, rt_decoder(), L.stream);
        launch_table_set(sv->d_table, (const int2 *)((uint8_t *)L.d_stage + o_upd), (int)upd.size(), L.stream);
        MCHIP_RT_CHECK(hipGetLastError());
        MCHIP_RT_CHECK(hipStreamSynchronize(L.stream));
        n += jobs.size();
        jobs.clear();
        return 0;
    };
    const size_t nb = (size_t)S.nblocks;
    for (size_t id = 0; id < nb; ++id) {
        if (!S.len[id] || sv->slot_of[id] >= 0) continue;
        const int s = sv->evict_one();
        if (s < 0) break;                         // the arena is full of pinned blocks
        sv->blk_of[s] = (int32_t)id;
        sv->slot_of[id] = s;
        sv->pins[s]++;                            // pinned until the preload ends
        pinned.push_back(s);
        jobs.push_back(BlockJob{S.off[id], 0, S.len[id], S.qslot[id], (uint32_t)s, 0});
        if (jobs.size() == 65536 && flush()) { drop_jobs(); unpin_all(); return -1; }
    }
    const int rc = flush();
    if (rc) drop_jobs();
    unpin_all();
    if (rc) return -1;
    sv->stats.misses += n;
    if (loaded) *loaded = n;
    return 0;
}

int mchip_server_get_stats(mchip_server *sv, mchip_server_stats *out) {
    if (!sv || !out) return rt_fail("null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    *out = sv->stats;
    return 0;
}

int mchip_server_resize(mchip_server *sv, uint64_t arena_bytes) {
    if (!sv) return rt_fail("null server");
    if (arena_bytes < (uint64_t)kNVox || arena_bytes / kNVox > 0x7FFFFFFFull)
        return rt_fail("mchip_server_resize: arena of %llu bytes", (unsigned long long)arena_bytes);
    std::unique_lock<std::mutex> lk(sv->mu);
    if (!sv->arena.vmm && arena_bytes > sv->arena.mapped)
        return rt_fail("mchip_server_resize: the arena is one fixed allocation of %zu bytes (no virtual memory "
                       "support)", sv->arena.mapped);
    const uint32_t ns = (uint32_t)(arena_bytes / kNVox);
    sv->budget = (size_t)ns * kNVox;
    sv->stats.arena_budget_bytes = sv->budget;
    if (ns >= sv->nslots) return 0;                // grows as requests need slots
    // retire the slots from ns on: no new request gets one, their blocks
    // are decoded again elsewhere when needed
    const uint32_t old = sv->nslots;
    sv->nslots = ns;
    if (sv->hand >= ns) sv->hand = 0;
    sv->freelist.erase(std::remove_if(sv->freelist.begin(), sv->freelist.end(),
                                      [ns](int32_t x) { return (uint32_t)x >= ns; }),
                       sv->freelist.end());
    for (uint32_t x = ns; x < old; ++x)
        if (sv->blk_of[x] >= 0) {
            sv->slot_of[sv->blk_of[x]] = -1;
            sv->blk_of[x] = -1;
        }
    sv->cv_done.wait(lk, [sv, ns, old] {
        for (uint32_t x = ns; x < old; ++x)
            if (sv->pins[x] > 0) return false;
        return true;
    });
    sv->blk_of.resize(ns);
    sv->pins.resize(ns);
    sv->refbit.resize(ns);
    sv->producer.resize(ns);
    sv->stats.arena_slots = ns;
    // segments wholly beyond the new size go back to the device
    if (sv->arena.vmm) {
        const size_t keep = ((size_t)ns * kNVox + kArenaSeg - 1) / kArenaSeg * kArenaSeg;
        if (keep < sv->arena.mapped) {
            MCHIP_RT_CHECK(hipDeviceSynchronize());
            if (sv->arena.shrink(keep)) return -1;
        }
        sv->stats.arena_mapped_bytes = sv->arena.mapped;
    }
    return 0;
}

int mchip_server_update_blocks(mchip_server *sv, const uint8_t *payload, uint64_t nbytes, int n, const uint64_t *ids,
                               const uint64_t *off, const uint32_t *len, const float *q) {
    if (!sv) return rt_fail("null server");
    if (n < 0 || (n && (!ids || !off || !len || !q)) || (nbytes && !payload))
        return rt_fail("mchip_server_update_blocks: bad argument");
    mchip_store &S = *sv->st;
    for (int i = 0; i < n; ++i) {
        if (ids[i] >= S.nblocks) return rt_fail("mchip_server_update_blocks: block %llu of %llu",
                                                (unsigned long long)ids[i], (unsigned long long)S.nblocks);
        if (len[i] && (off[i] > nbytes || len[i] > nbytes - off[i]))
            return rt_fail("mchip_server_update_blocks: block %d lies outside the payload bytes", i);
    }
    std::vector<uint32_t> qs(n, 0);
    for (int i = 0; i < n; ++i)
        if (len[i] && rt_qtab_slot(q[i], &qs[i])) return -1;
    std::lock_guard<std::mutex> lk(sv->mu);
    const uint64_t need = (nbytes + 255) & ~255ull;
    if (S.nbytes + need > S.cap) {
        // a larger heap; decodes launched earlier keep reading the old one
        const uint64_t cap = std::max<uint64_t>(2 * S.cap, S.nbytes + need + (64ull << 20));
        uint8_t *p = nullptr;
        MCHIP_RT_CHECK(hipMalloc((void **)&p, cap));
        if (hipMemcpyAsync(p, S.d_pay, S.nbytes, hipMemcpyDeviceToDevice, sv->upd_stream) != hipSuccess ||
            hipStreamSynchronize(sv->upd_stream) != hipSuccess) {
            (void)hipFree(p);
            return rt_fail("mchip_server_update_blocks: moving the payload heap failed");
        }
        S.old_pay.push_back(S.d_pay);
        S.d_pay = p;
        S.cap = cap;
    }
    const uint64_t at = S.nbytes;
    if (nbytes) {
        MCHIP_RT_CHECK(hipMemcpyAsync(S.d_pay + at, payload, nbytes, hipMemcpyHostToDevice, sv->upd_stream));
        MCHIP_RT_CHECK(hipStreamSynchronize(sv->upd_stream));
    }
    S.nbytes += need;
    for (int i = 0; i < n; ++i) {
        const uint64_t id = ids[i];
        S.off[id] = len[i] ? at + off[i] : 0;
        S.len[id] = len[i];
        S.qslot[id] = qs[i];
        sv->detach(id);
    }
    sv->stats.updated_blocks += (uint64_t)n;
    return 0;
}

void mchip_server_destroy(mchip_server *sv) {
    if (!sv) return;
    if (sv->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->stop = true;
            sv->cv_work.notify_one();
        }
        sv->worker.join();
    }
    for (Lane &L : sv->lanes) {
        if (L.stream) (void)hipStreamSynchronize(L.stream);
        if (L.h_stage) (void)hipHostFree(L.h_stage);
        if (L.d_stage) (void)hipFree(L.d_stage);
        if (L.d_scratch) (void)hipFree(L.d_scratch);
        if (L.d_info) (void)hipFree(L.d_info);
        if (L.d_bits) (void)hipFree(L.d_bits);
        if (L.d_list) (void)hipFree(L.d_list);
        if (L.ev_dec) (void)hipEventDestroy(L.ev_dec);
        if (L.ev_done) (void)hipEventDestroy(L.ev_done);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    for (auto &kv : sv->reqs) delete kv.second;
    for (void *p : sv->grave.dev) (void)hipFree(p);
    for (void *p : sv->grave.host) (void)hipHostFree(p);
    sv->arena.release();
    if (sv->d_table) (void)hipFree(sv->d_table);
    if (sv->upd_stream) (void)hipStreamDestroy(sv->upd_stream);
    delete sv;
}

}  // extern "C"
