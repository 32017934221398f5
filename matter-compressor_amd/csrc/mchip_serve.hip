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
// Requests are admitted in arrival order as lanes (streams) and arena slots
// free up, so requests that arrive while others run overlap on the device.
#include <hip/hip_runtime.h>

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

struct mchip_store {
    int device = 0;
    uint8_t *d_pay = nullptr;
    uint64_t nbytes = 0;
    int gz = 0, gy = 0, gx = 0;
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
    int64_t ticket = -1;                                // request in flight, -1: free
};

struct Req {
    int64_t ticket = 0;
    std::vector<int64_t> origin;
    int n = 0;
    int64_t dz = 0, dy = 0, dx = 0;
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

}  // namespace

struct mchip_server {
    mchip_store *st = nullptr;
    uint8_t *d_arena = nullptr;
    uint32_t nslots = 0;
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

    int evict_one() {
        for (uint32_t k = 0; k < 2 * nslots + 1; ++k) {
            const uint32_t s = hand;
            hand = hand + 1 == nslots ? 0 : hand + 1;
            if (pins[s] > 0) continue;
            if (blk_of[s] < 0) return (int)s;
            if (refbit[s]) { refbit[s] = 0; continue; }
            slot_of[blk_of[s]] = -1;
            blk_of[s] = -1;
            stats.evictions++;
            return (int)s;
        }
        return -1;
    }

    // Release the request's pins; its decodes are complete.
    void unpin(Req *r) {
        for (int32_t s : r->pinned) {
            pins[s]--;
            if (producer[s] == r->ticket) producer[s] = -1;
        }
        r->pinned.clear();
    }
    // Release the request's pins and free the slots it was to decode (not
    // admitted, or its device work failed).
    void rollback(Req *r) {
        for (int32_t s : r->pinned) {
            pins[s]--;
            if (producer[s] == r->ticket) {
                producer[s] = -1;
                slot_of[blk_of[s]] = -1;
                blk_of[s] = -1;
            }
        }
        r->pinned.clear();
    }

    // 1 admitted and issued, 0 not enough free slots now, -1 error (r->err).
    int try_admit(Req *r, int li) {
        Lane &L = lanes[li];
        const mchip_store &S = *st;
        const int gzb = (int)((r->dz + 15) >> 4) + 1, gyb = (int)((r->dy + 15) >> 4) + 1,
                  gxb = (int)((r->dx + 15) >> 4) + 1;
        const size_t cells_per = (size_t)gzb * gyb * gxb, cells = cells_per * r->n;
        if (++gen == 0) { std::fill(mark.begin(), mark.end(), 0u); gen = 1; }
        std::vector<BlockJob> jobs;
        std::vector<int32_t> bmap(cells, -1);
        std::vector<int> deps;
        uint64_t sw = 0, refs = 0, hits = 0;
        bool ok = true;
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
                        const int64_t id = (bz * S.gy + by) * S.gx + bx;
                        if (!S.len[id]) continue;
                        int32_t s;
                        if (mark[id] == gen) {
                            s = slot_of[id];
                        } else {
                            mark[id] = gen;
                            refs++;
                            s = slot_of[id];
                            if (s >= 0) {
                                hits++;
                                if (producer[s] >= 0 && producer[s] != r->ticket) {
                                    auto it = reqs.find(producer[s]);
                                    if (it != reqs.end() && it->second->lane >= 0) {
                                        const int dl = it->second->lane;
                                        bool have = false;
                                        for (int x : deps) have |= x == dl;
                                        if (!have) deps.push_back(dl);
                                    }
                                }
                            } else {
                                s = evict_one();
                                if (s < 0) { ok = false; break; }
                                blk_of[s] = (int32_t)id;
                                slot_of[id] = s;
                                producer[s] = r->ticket;
                                jobs.push_back(BlockJob{S.off[id], sw, S.len[id], S.qslot[id], (uint32_t)s, 0});
                                sw += 8ull * S.len[id] + 32ull;
                            }
                            pins[s]++;
                            refbit[s] = 1;
                            r->pinned.push_back(s);
                        }
                        bmap[(size_t)b * cells_per + (size_t)((iz * gyb + iy) * gxb + ix)] = s;
                    }
        }
        if (!ok) {                         // not enough unpinned slots now
            rollback(r);
            return 0;
        }
        const size_t jb = sizeof(BlockJob) * jobs.size(), ob = 24 * (size_t)r->n, mb = 4 * cells;
        const size_t o_org = (jb + 255) & ~size_t(255), o_map = (o_org + ob + 255) & ~size_t(255);
        const size_t total = o_map + mb;
        auto fail_req = [&](void) {
            r->err = "serving request: device allocation or launch failed";
            rollback(r);
            return -1;
        };
        if (grow_pinned(&L.h_stage, &L.cap_h, total, grave) || grow_device(&L.d_stage, &L.cap_d, total, grave) ||
            grow_device(&L.d_scratch, &L.cap_scratch, sizeof(uint32_t) * (sw + 64), grave) ||
            grow_device(&L.d_info, &L.cap_info, sizeof(uint2) * (jobs.size() + 1), grave))
            return fail_req();
        uint8_t *h = (uint8_t *)L.h_stage;
        if (jb) memcpy(h, jobs.data(), jb);
        memcpy(h + o_org, r->origin.data(), ob);
        memcpy(h + o_map, bmap.data(), mb);
        uint8_t *d = (uint8_t *)L.d_stage;
        if (hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, L.stream) != hipSuccess) return fail_req();
        for (int dl : deps)
            if (hipStreamWaitEvent(L.stream, lanes[dl].ev_dec, 0) != hipSuccess) return fail_req();
        if (!jobs.empty())
            launch_jobs(S.d_pay, (const BlockJob *)d, (int)jobs.size(), rt_qtabs(), (uint32_t *)L.d_scratch,
                        (uint2 *)L.d_info, d_arena, rt_decoder(), L.stream);
        if (hipEventRecord(L.ev_dec, L.stream) != hipSuccess) return fail_req();
        launch_crop(d_arena, (const int32_t *)(d + o_map), (const int64_t *)(d + o_org), r->n, gzb, gyb, gxb,
                    r->dz, r->dy, r->dx, r->d_out, r->stride, L.stream);
        if (hipGetLastError() != hipSuccess || hipEventRecord(L.ev_done, L.stream) != hipSuccess) return fail_req();
        L.ticket = r->ticket;
        r->lane = li;
        stats.block_refs += refs;
        stats.hits += hits;
        stats.misses += jobs.size();
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
                    if (rc == 0) r->err = "serving request: the arena holds fewer slots than one request needs";
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

mchip_store *mchip_store_create(const uint8_t *payload, uint64_t nbytes, const uint64_t *off, const uint32_t *len,
                                const float *q, int gz, int gy, int gx) {
    if (gz <= 0 || gy <= 0 || gx <= 0 || !off || !len || !q || (nbytes && !payload)) {
        rt_fail("mchip_store_create: bad argument");
        return nullptr;
    }
    const size_t nb = (size_t)gz * gy * gx;
    for (size_t i = 0; i < nb; ++i)
        if (len[i] && (off[i] > nbytes || len[i] > nbytes - off[i])) {
            rt_fail("mchip_store_create: block %zu lies outside the payload bytes", i);
            return nullptr;
        }
    auto *st = new mchip_store;
    st->gz = gz; st->gy = gy; st->gx = gx;
    st->off.assign(off, off + nb);
    st->len.assign(len, len + nb);
    st->qslot.assign(nb, 0);
    std::map<uint32_t, uint32_t> qs;              // q bit pattern -> table slot
    for (size_t i = 0; i < nb; ++i) {
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
    if (hipGetDevice(&st->device) != hipSuccess ||
        hipMalloc((void **)&st->d_pay, nbytes + 256) != hipSuccess ||
        (nbytes && hipMemcpy(st->d_pay, payload, nbytes, hipMemcpyHostToDevice) != hipSuccess)) {
        rt_fail("mchip_store_create: device allocation or upload of %llu bytes failed",
                (unsigned long long)nbytes);
        if (st->d_pay) (void)hipFree(st->d_pay);
        delete st;
        return nullptr;
    }
    st->nbytes = nbytes;
    return st;
}

void mchip_store_destroy(mchip_store *st) {
    if (!st) return;
    if (st->d_pay) (void)hipFree(st->d_pay);
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
    sv->nslots = (uint32_t)ns;
    if (hipMalloc((void **)&sv->d_arena, ns * kNVox) != hipSuccess) {
        rt_fail("mchip_server_create: arena allocation of %llu bytes failed", (unsigned long long)(ns * kNVox));
        delete sv;
        return nullptr;
    }
    const size_t nb = (size_t)st->gz * st->gy * st->gx;
    sv->slot_of.assign(nb, -1);
    sv->mark.assign(nb, 0u);
    sv->blk_of.assign(ns, -1);
    sv->pins.assign(ns, 0);
    sv->refbit.assign(ns, 0);
    sv->producer.assign(ns, -1);
    sv->stats.arena_slots = ns;
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
        for (int32_t s : pinned) sv->pins[s]--;
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
        for (BlockJob &j : jobs) { j.scratch_off = sw; sw += 8ull * j.len + 32ull; }
        const size_t jb = sizeof(BlockJob) * jobs.size();
        if (grow_pinned(&L.h_stage, &L.cap_h, jb, sv->grave) || grow_device(&L.d_stage, &L.cap_d, jb, sv->grave) ||
            grow_device(&L.d_scratch, &L.cap_scratch, sizeof(uint32_t) * (sw + 64), sv->grave) ||
            grow_device(&L.d_info, &L.cap_info, sizeof(uint2) * (jobs.size() + 1), sv->grave))
            return -1;
        memcpy(L.h_stage, jobs.data(), jb);
        MCHIP_RT_CHECK(hipMemcpyAsync(L.d_stage, L.h_stage, jb, hipMemcpyHostToDevice, L.stream));
        launch_jobs(S.d_pay, (const BlockJob *)L.d_stage, (int)jobs.size(), rt_qtabs(), (uint32_t *)L.d_scratch,
                    (uint2 *)L.d_info, sv->d_arena, rt_decoder(), L.stream);
        MCHIP_RT_CHECK(hipGetLastError());
        MCHIP_RT_CHECK(hipStreamSynchronize(L.stream));
        n += jobs.size();
        jobs.clear();
        return 0;
    };
    const size_t nb = (size_t)S.gz * S.gy * S.gx;
    for (size_t id = 0; id < nb; ++id) {
        if (!S.len[id] || sv->slot_of[id] >= 0) continue;
        const int s = sv->evict_one();
        if (s < 0) break;                         // the arena is full of pinned blocks
        if (sv->blk_of[s] >= 0) { sv->slot_of[sv->blk_of[s]] = -1; }
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
        if (L.ev_dec) (void)hipEventDestroy(L.ev_dec);
        if (L.ev_done) (void)hipEventDestroy(L.ev_done);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    for (auto &kv : sv->reqs) delete kv.second;
    for (void *p : sv->grave.dev) (void)hipFree(p);
    for (void *p : sv->grave.host) (void)hipHostFree(p);
    if (sv->d_arena) (void)hipFree(sv->d_arena);
    delete sv;
}

}  // extern "C"
