// mchip_api.hip -- C ABI of the HIP codec engine (declared in include/mchip.h):
// device selection, constant tables, per-quality tables, contexts (stream +
// growable workspace), launches and kernel timing.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mchip.h"
#include "mchip_internal.hpp"
#include "mchip_runtime.hpp"

using namespace mchip;

namespace {

thread_local std::string t_err;

int fail(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
    return -1;
}
#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) return fail("%s failed: %s", #x, hipGetErrorString(e_));   \
    } while (0)

// Trained prior endpoint tables (q=1, q=12), P(bit==0)/4096, class order
// sig, mag, eob, mask, mask-u, mask-a, flag, dc; unlisted slots are 2048.
// Format constants: identical numbers to src/mc_codec.c:121-140.
const uint16_t kLoHead[8][24] = {
    {115, 110, 112, 144, 3866, 3522, 3330, 948, 3939, 3723, 3483, 2383,
     4007, 3795, 3575, 3130, 4042, 3912, 3805, 3686, 4064, 4064, 4028, 4064},
    {1996, 1234, 948, 743, 595, 539, 441, 421, 333, 358, 297, 305},
    {32, 32, 32, 32, 32, 32, 32, 32, 32, 92, 285, 2295, 4064},
    {4025, 3064, 2995, 827, 2107, 473, 576, 50},
    {250, 1713, 2961, 2991},
    {3936, 32},
    {1023, 4064, 4064, 4064},
    {3873, 366, 1820, 2056, 2040, 2055, 2041, 2045},
};
const int kLoN[8] = {24, 12, 13, 8, 4, 2, 4, 8};
const uint16_t kHiHead[8][12] = {
    {456, 335, 535, 1133, 3950, 3755, 3590, 2892, 4064, 4051, 3990, 3686},
    {2121, 1410, 1081, 863, 746, 631, 569, 508, 465, 431, 399, 370},
    {59, 32, 32, 32, 32, 63, 92, 315, 2273, 4064, 4064},
    {4025, 3064, 2995, 827, 2107, 473, 576, 50},
    {250, 1713, 2961, 2991},
    {3936, 32},
    {1023, 4064, 4064, 4064},
    {3873, 366, 1820, 2056, 2040, 2055, 2041, 2045},
};
const int kHiN[8] = {12, 12, 11, 8, 4, 2, 4, 8};

// Quality tables are never overwritten: a slot, once handed out, keeps its
// table for the life of the process, because launches in flight, captured
// graphs and serving stores hold slot numbers.  The device array grows by
// doubling (the replaced array stays allocated for the same reason).
constexpr uint32_t kInitQTabs = 4096;
constexpr uint32_t kMaxQTabs = 1u << 20;     // 504 MiB of tables

// Launch-shape tuning knobs (mchip_set_tuning); process-wide, read as one
// snapshot per call (tuning()).
struct Tuning {
    int decoder = 3;          // entropy decoder: 0 per scan position, 2 branch-free lane decoder, 3 its assembly loop
                              // over an LDS payload window
    int enc_batch = 0;        // chunks per encode launch sequence (device workspace ~165 MiB each);
                              // 0: sized from the free device memory (enc_batch_for)
    int persistent = 1;       // batched decode (>= persist_min chunks): persistent entropy launch
    int persist_min = 2;
    int persist_wg = 0;       // workgroups of the persistent launch (0: occupancy x CUs)
    int persist_win = 32;     // payload window words per lane of the persistent launch (16, 32, 64)
    int len_shift = 3;        // k_index lane order: length buckets of 1 << len_shift bytes (2..5)
    int group_shift = 6;      // k_group_order: first-block length buckets of 1 << group_shift bytes (4..8)
    int recon_bpw = 1;        // k_recon blocks per workgroup (1, 2, 4)
    int enc_slices = 0;       // encode sequence: chunk slices on two streams (0: one slice)
    int enc_order = 1;        // encode entropy pass: 1 heaviest chunk first, 0 in sequence order
    int enc_lanes = 0;        // encode entropy pass: 1 blocks of each chunk costliest first, 0 raster order
    int slices = 0;           // batched persistent decode: chunk slices whose reconstruction overlaps the
                              // next slice's entropy pass on a second stream (0: 2 from 16 chunks, else 1)
    int enc_keep_mib = 2048;  // encode workspace a context keeps after a call (MiB); more is freed
    int merge = 1;            // concurrent one-chunk host calls join one batched decode
    int merge_inflight = 2;   // merged batches in flight at most (1..4)
    int merge_min = 8;        // queued calls that start a batch while another is in flight
};
std::mutex g_tune_mu;
Tuning g_tune;
Tuning tuning() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return g_tune;
}

// Decode launch shapes chosen by mchip_autotune_decode, per workload
// signature: chunk count and mean blob length as bit lengths, and whether
// the qualities differ.  Guarded by g_tune_mu.
std::map<uint32_t, Tuning> g_tuned;

uint32_t workload_signature(int n, const uint64_t *len, const float *q) {
    uint64_t tot = 0;
    bool mixed = false;
    for (int i = 0; i < n; ++i) {
        tot += len[i];
        mixed |= q[i] != q[0];
    }
    const uint64_t mean = n > 0 ? tot / (uint64_t)n : 0;
    auto bits = [](uint64_t x) { uint32_t b = 0; while (x) { b++; x >>= 1; } return b; };
    return bits((uint64_t)n) | (bits(mean >> 10) << 8) | ((mixed ? 1u : 0u) << 16);
}

// The tuning a decode of this batch uses: the autotuned one for its
// signature, else the process tuning.
Tuning tuning_for(int n, const uint64_t *len, const float *q) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!g_tuned.empty()) {
        auto it = g_tuned.find(workload_signature(n, len, q));
        if (it != g_tuned.end()) {
            Tuning t = it->second;
            t.enc_batch = g_tune.enc_batch;     // only the decode launch shapes come from the cache
            t.merge = g_tune.merge;
            t.merge_inflight = g_tune.merge_inflight;
            t.merge_min = g_tune.merge_min;
            return t;
        }
    }
    return g_tune;
}

// Encode shapes chosen by mchip_autotune_encode, per encode signature: chunk
// count as a bit length, whether the qualities differ and whether any chunk
// has corrections.  Guarded by g_tune_mu.
std::map<uint32_t, Tuning> g_tuned_enc;

uint32_t encode_signature(int n, const float *q, const int *tau) {
    bool mixed = false, corr = false;
    for (int i = 0; i < n; ++i) {
        mixed |= q[i] != q[0];
        corr |= tau && tau[i] > 0;
    }
    uint32_t b = 0;
    for (uint32_t x = (uint32_t)n; x; x >>= 1) b++;
    return b | ((mixed ? 1u : 0u) << 8) | ((corr ? 1u : 0u) << 9);
}

// The tuning an encode of these chunks uses: the process tuning with the
// autotuned encode shapes for its signature, if any.
Tuning tuning_for_encode(int n, const float *q, const int *tau) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    Tuning t = g_tune;
    if (!g_tuned_enc.empty()) {
        auto it = g_tuned_enc.find(encode_signature(n, q, tau));
        if (it != g_tuned_enc.end()) {
            t.enc_batch = it->second.enc_batch;
            t.enc_slices = it->second.enc_slices;
            t.enc_order = it->second.enc_order;
            t.enc_lanes = it->second.enc_lanes;
        }
    }
    return t;
}

// One prior configuration: the q=1 and q=12 endpoint tables.
struct PriorSet {
    uint16_t lo[8][32], hi[8][32];
};

constexpr int kMaxDevices = 64;

struct Global {
    std::mutex mu;
    int device = -1;
    bool ready = false;
    void *static_ws[kMaxDevices] = {};      // device address of the fixed workspace, per device
    std::vector<PriorSet> priors;           // every configuration seen; [0] = trained defaults
    uint32_t cur = 0;                       // index of the process priors in priors[]
    std::map<uint64_t, uint32_t> qidx;      // (prior set << 32 | sanitised-q bits) -> slot
    std::vector<QTab> tabs;                 // host copy of every slot handed out
    std::atomic<QTab *> d_qtabs{nullptr};   // device array of cap slots
    uint32_t cap = 0;
    std::vector<QTab *> retired;            // replaced device arrays (still read by older launches)
};
Global &G() {
    static Global g;
    return g;
}

void default_priors(uint16_t lo[8][32], uint16_t hi[8][32]) {
    for (int c = 0; c < 8; ++c)
        for (int s = 0; s < 32; ++s) {
            lo[c][s] = s < kLoN[c] ? kLoHead[c][s] : 2048;
            hi[c][s] = s < kHiN[c] ? kHiHead[c][s] : 2048;
        }
}

float sanitize_q(float q) { return (!(q > 0.0f) || q > 1024.0f) ? 8.0f : q; }

// Host evaluation of the per-quality tables with the reference's float
// expressions (src/mc_codec.c:156-167 priors, :541-552 steps).
QTab make_qtab(float q, const uint16_t lo[8][32], const uint16_t hi[8][32]) {
    QTab t;
    memset(&t, 0, sizeof t);
    for (int l = 0; l < 46; ++l) {
        float w = powf(1.0f + (float)l, 0.65f);
        t.step[l] = q * w;
        t.rstep[l] = 1.0f / t.step[l];
    }
    const float flo = 0.0f, fhi = 3.585f;
    float w = (q <= 1.0f) ? 0.0f : ((float)(log((double)q) / log(2.0)) - flo) / (fhi - flo);
    if (w < 0.0f) w = 0.0f;
    if (w > 1.0f) w = 1.0f;
    uint16_t pri[8][32];
    for (int c = 0; c < 8; ++c)
        for (int s = 0; s < 32; ++s) {
            int a = lo[c][s], b = hi[c][s];
            float v = (float)a + (float)(b - a) * w;
            v = v + 0.5f;
            pri[c][s] = (uint16_t)v;
        }
    memcpy(t.sig, pri[0], sizeof t.sig);
    memcpy(t.mag, pri[1], sizeof t.mag);
    memcpy(t.eob, pri[2], sizeof t.eob);
    memcpy(t.flag, pri[6], sizeof t.flag);
    memcpy(t.dc, pri[7], sizeof t.dc);
    return t;
}

int init_locked(Global &g, int device) {
    if (g.ready) {
        if (device < 0 || device == g.device) return 0;
        return fail("the library already runs on device %d (one device per process)", g.device);
    }
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev <= 0) return fail("no HIP device");
    if (device < 0) device = 0;
    if (device >= ndev) return fail("device %d out of range (%d devices)", device, ndev);
    HIPCHK(hipSetDevice(device));
    // DCT-II matrix in double rounded to float (src/mc_codec_float.h:70-78)
    float cmf[256];
    for (int k = 0; k < 16; ++k) {
        double ck = (k == 0) ? sqrt(1.0 / 16) : sqrt(2.0 / 16);
        for (int n = 0; n < 16; ++n)
            cmf[k * 16 + n] = (float)(ck * cos(3.14159265358979323846 * (2.0 * n + 1.0) * k / 32.0));
    }
    // ascending-L1 scan, ties by raster index (src/mc_codec.c:283-304):
    // counting sort over the 46 L1 values keeps raster order within a value.
    uint16_t scan[kNVox];
    int count[47] = {0};
    for (int i = 0; i < kNVox; ++i) count[(i >> 8) + ((i >> 4) & 15) + (i & 15) + 1]++;
    for (int l = 1; l < 47; ++l) count[l] += count[l - 1];
    uint16_t band_start[9];
    for (int b = 0; b < 8; ++b) band_start[b] = (uint16_t)count[6 * b];   // band = min(7, L1/6)
    band_start[8] = kNVox;
    for (int i = 0; i < kNVox; ++i) scan[count[(i >> 8) + ((i >> 4) & 15) + (i & 15)]++] = (uint16_t)i;
    HIPCHK(upload_constants(cmf, band_start, scan));
    if (device >= kMaxDevices) return fail("device %d: at most %d devices", device, kMaxDevices);
    if (!g.static_ws[device]) {
        void *ws = nullptr;
        HIPCHK(static_scratch_address(&ws));
        g.static_ws[device] = ws;
    }
    QTab *d = nullptr;
    HIPCHK(hipMalloc((void **)&d, sizeof(QTab) * kInitQTabs));
    g.d_qtabs.store(d, std::memory_order_release);
    g.cap = kInitQTabs;
    PriorSet def;
    default_priors(def.lo, def.hi);
    g.priors.assign(1, def);
    g.cur = 0;
    g.device = device;
    g.ready = true;
    return 0;
}

// Slot of the quality table for q under the process priors (built and
// uploaded on first use).  may_grow = false refuses to allocate a larger
// device array (callers that must not allocate).
int qtab_slot(float q, uint32_t *slot, bool may_grow = true) {
    Global &g = G();
    std::lock_guard<std::mutex> lk(g.mu);
    if (init_locked(g, -1)) return -1;
    float qs = sanitize_q(q);
    uint32_t bits;
    memcpy(&bits, &qs, 4);
    const uint64_t key = ((uint64_t)g.cur << 32) | bits;
    auto it = g.qidx.find(key);
    if (it != g.qidx.end()) { *slot = it->second; return 0; }
    const uint32_t s = (uint32_t)g.tabs.size();
    if (s >= kMaxQTabs) return fail("more than %u distinct (priors, quality) tables", kMaxQTabs);
    const QTab t = make_qtab(qs, g.priors[g.cur].lo, g.priors[g.cur].hi);
    if (s >= g.cap) {
        if (!may_grow) return fail("the quality-table array is full (%u tables) and this call may not allocate", g.cap);
        const uint32_t ncap = std::min(kMaxQTabs, 2 * g.cap);
        QTab *nd = nullptr;
        HIPCHK(hipMalloc((void **)&nd, sizeof(QTab) * ncap));
        HIPCHK(hipMemcpy(nd, g.tabs.data(), sizeof(QTab) * s, hipMemcpyHostToDevice));
        g.retired.push_back(g.d_qtabs.load(std::memory_order_relaxed));
        g.d_qtabs.store(nd, std::memory_order_release);
        g.cap = ncap;
    }
    HIPCHK(hipMemcpy(g.d_qtabs.load(std::memory_order_relaxed) + s, &t, sizeof t, hipMemcpyHostToDevice));
    g.tabs.push_back(t);
    g.qidx[key] = s;
    *slot = s;
    return 0;
}

// The device table array; holds every slot handed out before the call.
const QTab *qtabs_now() { return G().d_qtabs.load(std::memory_order_acquire); }

}  // namespace

namespace mchip {
int rt_fail(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
    return -1;
}
int rt_qtab_slot(float q, uint32_t *slot) { return qtab_slot(q, slot); }
const QTab *rt_qtabs() { return qtabs_now(); }
int rt_decoder() { return tuning().decoder; }
}  // namespace mchip

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct TimedLaunch {
    const char *name;
    hipEvent_t a, b;
};

// One half of the double-buffered host-blob verify: a pinned slab the host
// threads gather blobs into, its device copy, and the per-blob offsets,
// lengths and results.
struct VerifySlab {
    uint8_t *h = nullptr, *d = nullptr; size_t cap = 0;
    uint8_t *hm = nullptr, *dm = nullptr; size_t capm = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
    int i0 = 0, i1 = 0;
};

struct mchip_ctx {
    hipStream_t stream = nullptr;
    // growable device buffers
    void *d_cd = nullptr; size_t cap_cd = 0;
    void *d_rec = nullptr; size_t cap_rec = 0;
    void *d_plist = nullptr; size_t cap_plist = 0;
    void *d_cinfo = nullptr; size_t cap_cinfo = 0;
    void *d_binfo = nullptr; size_t cap_binfo = 0;
    void *d_scratch = nullptr; size_t cap_scratch = 0;
    void *d_stage_in = nullptr; size_t cap_stage_in = 0;
    void *d_stage_out = nullptr; size_t cap_stage_out = 0;
    void *d_aux = nullptr; size_t cap_aux = 0;
    void *d_order = nullptr; size_t cap_order = 0;      // persistent decode: group order + counters
    // second stream of sliced batched decodes, created on first use
    hipStream_t stream_b = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // encode workspace
    void *d_ecd = nullptr; size_t cap_ecd = 0;
    void *d_einfo = nullptr; size_t cap_einfo = 0;
    void *d_levels = nullptr; size_t cap_levels = 0;
    void *d_corr = nullptr; size_t cap_corr = 0;
    void *d_slots = nullptr; size_t cap_slots = 0;
    void *d_plen = nullptr; size_t cap_plen = 0;
    void *d_dst = nullptr; size_t cap_dst = 0;
    void *d_res = nullptr; size_t cap_res = 0;
    void *d_fmap = nullptr; size_t cap_fmap = 0;
    void *d_lanes = nullptr; size_t cap_lanes = 0;
    void *h_res = nullptr; size_t cap_h_res = 0;
    // block jobs / crops
    void *d_jobs = nullptr; size_t cap_jobs = 0;
    void *d_jscratch = nullptr; size_t cap_jscratch = 0;
    void *d_jinfo = nullptr; size_t cap_jinfo = 0;
    void *d_bmap = nullptr; size_t cap_bmap = 0;
    void *d_arena = nullptr; size_t cap_arena = 0;
    void *h_crop = nullptr; size_t cap_h_crop = 0;
    hipEvent_t ev_crop = nullptr;
    bool crop_pending = false;
    // pinned descriptor staging, reusable once ev_desc has completed
    void *h_desc = nullptr; size_t cap_h_desc = 0;
    hipEvent_t ev_desc = nullptr;
    bool desc_pending = false;
    // host-blob verify
    VerifySlab vslab[2];
    // timing
    bool timing = false;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> free_events;
    std::map<std::string, std::pair<double, long>> stats;
};

namespace {

int ensure(mchip_ctx *c, void **p, size_t *cap, size_t need) {
    if (need <= *cap) return 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (*p) HIPCHK(hipFree(*p));
    *p = nullptr;
    size_t n = need + need / 4 + 256;
    hipError_t e = hipMalloc(p, n);
    if (e != hipSuccess) { *cap = 0; *p = nullptr; return fail("hipMalloc(%zu) failed: %s", n, hipGetErrorString(e)); }
    *cap = n;
    return 0;
}

int desc_buffer(mchip_ctx *c, size_t need, void **h) {
    if (c->desc_pending) {
        HIPCHK(hipEventSynchronize(c->ev_desc));
        c->desc_pending = false;
    }
    if (need > c->cap_h_desc) {
        if (c->h_desc) HIPCHK(hipHostFree(c->h_desc));
        c->h_desc = nullptr;
        size_t n = need + need / 2 + 4096;
        HIPCHK(hipHostMalloc(&c->h_desc, n, hipHostMallocDefault));
        c->cap_h_desc = n;
    }
    *h = c->h_desc;
    return 0;
}

hipEvent_t get_event(mchip_ctx *c) {
    if (!c->free_events.empty()) {
        hipEvent_t e = c->free_events.back();
        c->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

template <class F>
void timed(mchip_ctx *c, const char *name, F &&launch) {
    if (!c->timing) { launch(); return; }
    TimedLaunch t{name, get_event(c), get_event(c)};
    (void)hipEventRecord(t.a, c->stream);
    launch();
    (void)hipEventRecord(t.b, c->stream);
    c->pending.push_back(t);
}

int collect_timing(mchip_ctx *c) {
    for (auto &t : c->pending) {
        HIPCHK(hipEventSynchronize(t.b));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
        auto &s = c->stats[t.name];
        s.first += ms;
        s.second += 1;
        c->free_events.push_back(t.a);
        c->free_events.push_back(t.b);
    }
    c->pending.clear();
    return 0;
}

// Workgroups of the persistent entropy launch: as many as the device keeps
// resident at once (occupancy per CU x CUs), per decoder and window size.
int persistent_grid(int decoder, int win) {
    static std::atomic<int> cached[2][3];              // lane / assembly lane decoder; window 16, 32, 64
    std::atomic<int> &slot = cached[decoder == 3 ? 1 : 0][win == 16 ? 0 : (win == 64 ? 2 : 1)];
    const int have = slot.load(std::memory_order_relaxed);
    if (have) return have;
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 1024;
    if (occupancy_persist(decoder, win, &per) != hipSuccess || per <= 0) per = 4;
    slot.store(cus * per, std::memory_order_relaxed);
    return cus * per;
}

}  // namespace

// Descriptors, workspace and launch arguments of a decode of n resident
// chunk blobs (everything but the launches).
// out_at (optional) gives each chunk's byte offset in d_out; without it
// chunk i is written at i * out_stride.
static int prepare_decode(mchip_ctx *c, const uint8_t *d_blobs, const uint64_t *blob_off, const uint64_t *blob_len,
                          const float *q, int n, uint8_t *d_out, uint64_t out_stride, const Tuning &tn,
                          DecodeArgs &a, const uint64_t *out_at = nullptr) {
    if (!c) return fail("null context");
    if (n <= 0) return 0;
    if (!d_blobs || !blob_off || !blob_len || !q || !d_out) return fail("null argument");
    if (out_stride < (uint64_t)16777216) return fail("out_stride %llu < 256^3", (unsigned long long)out_stride);
    // descriptors: inside the kernel arguments for small calls, else staged
    // through pinned memory and copied on the stream
    const bool inl = n <= kInlineDesc;
    ChunkDesc local[kInlineDesc];
    ChunkDesc *cd = local;
    if (!inl) {
        void *h = nullptr;
        if (desc_buffer(c, sizeof(ChunkDesc) * (size_t)n, &h)) return -1;
        cd = (ChunkDesc *)h;
    }
    uint64_t sw = 0;
    for (int i = 0; i < n; ++i) {
        uint32_t slot;
        if (qtab_slot(q[i], &slot)) return -1;
        if (blob_len[i] > 0xFFFFFFF0ull) return fail("blob %d too long", i);
        cd[i].blob_off = blob_off[i];
        cd[i].blob_len = blob_len[i];
        cd[i].out_off = out_at ? out_at[i] : (uint64_t)i * out_stride;
        cd[i].scratch_off = sw;
        cd[i].qtab = slot;
        cd[i].pad = 0;
        sw += scratch_words_for(blob_len[i]);
    }
    size_t nb = (size_t)n * kNVox;
    if ((!inl && ensure(c, &c->d_cd, &c->cap_cd, sizeof(ChunkDesc) * n)) ||
        ensure(c, &c->d_rec, &c->cap_rec, sizeof(BlkRec) * nb) ||
        ensure(c, &c->d_plist, &c->cap_plist, sizeof(uint16_t) * nb) ||
        ensure(c, &c->d_cinfo, &c->cap_cinfo, sizeof(ChunkInfo) * n) ||
        ensure(c, &c->d_binfo, &c->cap_binfo, sizeof(uint2) * nb) ||
        ensure(c, &c->d_scratch, &c->cap_scratch, sizeof(uint32_t) * sw))
        return -1;
    DescRef dref{};
    if (inl) {
        dref.ptr = nullptr;
        for (int i = 0; i < n; ++i) dref.inl[i] = cd[i];
    } else {
        HIPCHK(hipMemcpyAsync(c->d_cd, cd, sizeof(ChunkDesc) * n, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipEventRecord(c->ev_desc, c->stream));
        c->desc_pending = true;
        dref.ptr = (const ChunkDesc *)c->d_cd;
    }
    a.blobs = d_blobs;
    a.cd = dref;
    a.qtabs = qtabs_now();
    a.rec = (BlkRec *)c->d_rec;
    a.plist = (uint16_t *)c->d_plist;
    a.cinfo = (ChunkInfo *)c->d_cinfo;
    a.scratch = (uint32_t *)c->d_scratch;
    a.binfo = (uint2 *)c->d_binfo;
    a.out = d_out;
    a.n = n;
    a.decoder = tn.decoder;
    a.persistent = tn.persistent && n >= tn.persist_min && a.decoder >= 2;
    a.persist_win = tn.persist_win;
    a.len_shift = tn.len_shift;
    a.group_shift = tn.group_shift;
    a.recon_bpw = tn.recon_bpw;
    a.nwg = 0;
    a.order = nullptr;
    a.pctl = nullptr;
    a.slices = 1;
    if (a.persistent) {
        if (ensure(c, &c->d_order, &c->cap_order, sizeof(uint32_t) * ((size_t)n * 64 + 2 * kMaxSlices))) return -1;
        a.order = (uint32_t *)c->d_order;
        a.pctl = a.order + (size_t)n * 64;
        a.nwg = tn.persist_wg > 0 ? tn.persist_wg : persistent_grid(tn.decoder, tn.persist_win);
        if (a.nwg > n * 64) a.nwg = n * 64;
        // slices need descriptors in device memory (a slice addresses its own)
        // (automatic: two slices from 16 chunks; more slices lengthen the
        // persistent launches' tails more than the overlap gains)
        const int want = tn.slices > 0 ? tn.slices : (n >= 16 ? 2 : 1);
        const int ns = std::max(1, std::min(want, std::min(kMaxSlices, n)));
        if (!inl && ns > 1) {
            if (!c->stream_b && (hipStreamCreateWithFlags(&c->stream_b, hipStreamNonBlocking) != hipSuccess ||
                                 hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
                                 hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess))
                return fail("second stream creation failed");
            a.slices = ns;
        }
    }
    return 0;
}

// Chunks [n*i/ns, n*(i+1)/ns) of a as a decode of their own: per-chunk
// records, descriptors and the group order are addressed from the slice's
// first chunk; workspace and output offsets in the descriptors are absolute.
static DecodeArgs slice_of(const DecodeArgs &a, int i, int ns) {
    const int c0 = (int)((int64_t)a.n * i / ns), c1 = (int)((int64_t)a.n * (i + 1) / ns);
    DecodeArgs s = a;
    s.n = c1 - c0;
    s.cd.ptr = a.cd.ptr + c0;
    s.rec = a.rec + (size_t)c0 * kNVox;
    s.plist = a.plist + (size_t)c0 * kNVox;
    s.cinfo = a.cinfo + c0;
    s.binfo = a.binfo + (size_t)c0 * kNVox;
    s.order = a.order + (size_t)c0 * 64;
    s.pctl = a.order + (size_t)a.n * 64 + 2 * i;
    if (s.nwg > s.n * 64) s.nwg = s.n * 64;
    s.slices = 1;
    return s;
}

// Index all chunks, then per chunk slice the entropy pass and the
// reconstruction.  Slices alternate between the context's stream and its
// second stream, so a slice's reconstruction (and the tail of its persistent
// entropy launch) overlaps the next slice's entropy pass; the context's
// stream waits for the second at the end, so completion on it means the
// whole decode is done (also inside a captured graph: fork and join events).
static void launch_decode(mchip_ctx *c, const DecodeArgs &a) {
    timed(c, "index", [&] { launch_index(a, c->stream); });
    if (a.slices <= 1) {
        timed(c, "entropy", [&] { launch_entropy(a, c->stream); });
        timed(c, "recon", [&] { launch_recon(a, c->stream); });
        return;
    }
    timed(c, "entropy_recon", [&] {
        (void)hipEventRecord(c->ev_fork, c->stream);
        (void)hipStreamWaitEvent(c->stream_b, c->ev_fork, 0);
        for (int i = 0; i < a.slices; ++i) {
            const DecodeArgs s = slice_of(a, i, a.slices);
            hipStream_t st = (i & 1) ? c->stream_b : c->stream;
            launch_entropy(s, st);
            launch_recon(s, st);
        }
        (void)hipEventRecord(c->ev_join, c->stream_b);
        (void)hipStreamWaitEvent(c->stream, c->ev_join, 0);
    });
}


extern "C" {

int mchip_init(int device) {
    Global &g = G();
    std::lock_guard<std::mutex> lk(g.mu);
    return init_locked(g, device);
}

const char *mchip_last_error(void) { return t_err.c_str(); }

int mchip_set_priors(const uint16_t *plo, const uint16_t *phi) {
    Global &g = G();
    std::lock_guard<std::mutex> lk(g.mu);
    if (init_locked(g, -1)) return -1;
    PriorSet ps;
    if (plo && phi) {
        memcpy(ps.lo, plo, sizeof ps.lo);
        memcpy(ps.hi, phi, sizeof ps.hi);
    } else {
        default_priors(ps.lo, ps.hi);
    }
    // a configuration seen before keeps its index (and so its quality-table
    // slots); the tables of other configurations stay valid for their holders
    uint32_t id = (uint32_t)g.priors.size();
    for (uint32_t i = 0; i < g.priors.size(); ++i)
        if (!memcmp(&g.priors[i], &ps, sizeof ps)) { id = i; break; }
    if (id == g.priors.size()) g.priors.push_back(ps);
    g.cur = id;
    return 0;
}

int mchip_set_tuning(const char *key, int value) {
    if (!key) return fail("null key");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!strcmp(key, "persistent")) { g_tune.persistent = value != 0; return 0; }
    if (!strcmp(key, "persist_min")) { g_tune.persist_min = value < 1 ? 1 : value; return 0; }
    if (!strcmp(key, "persist_wg")) {
        if (value < 0 || value > 1 << 20) return fail("persist_wg out of range");
        g_tune.persist_wg = value;
        return 0;
    }
    if (!strcmp(key, "persist_win")) {
        if (value != 16 && value != 32 && value != 64) return fail("persist_win must be 16, 32 or 64");
        g_tune.persist_win = value;
        return 0;
    }
    if (!strcmp(key, "enc_batch")) {
        if (value < 0 || value > 1024) return fail("enc_batch must be in [0, 1024] (0: automatic)");
        g_tune.enc_batch = value;
        return 0;
    }
    if (!strcmp(key, "decoder")) {
        if (value != 0 && value != 2 && value != 3) return fail("decoder must be 0, 2 or 3");
        g_tune.decoder = value;
        return 0;
    }
    if (!strcmp(key, "merge")) { g_tune.merge = value != 0; return 0; }
    if (!strcmp(key, "len_shift")) {
        if (value < 2 || value > 5) return fail("len_shift must be in [2, 5]");
        g_tune.len_shift = value;
        return 0;
    }
    if (!strcmp(key, "group_shift")) {
        if (value < 4 || value > 8) return fail("group_shift must be in [4, 8]");
        g_tune.group_shift = value;
        return 0;
    }
    if (!strcmp(key, "recon_bpw")) {
        if (value != 1 && value != 2 && value != 4) return fail("recon_bpw must be 1, 2 or 4");
        g_tune.recon_bpw = value;
        return 0;
    }
    if (!strcmp(key, "enc_slices")) {
        if (value < 0 || value > kMaxSlices) return fail("enc_slices must be 0..%d", kMaxSlices);
        g_tune.enc_slices = value;
        return 0;
    }
    if (!strcmp(key, "slices")) {
        if (value < 0 || value > kMaxSlices) return fail("slices must be 0..%d", kMaxSlices);
        g_tune.slices = value;
        return 0;
    }
    if (!strcmp(key, "tuned_clear")) { g_tuned.clear(); g_tuned_enc.clear(); return 0; }
    if (!strcmp(key, "enc_order")) {
        if (value != 0 && value != 1) return fail("enc_order must be 0 or 1");
        g_tune.enc_order = value;
        return 0;
    }
    if (!strcmp(key, "enc_lanes")) {
        if (value != 0 && value != 1) return fail("enc_lanes must be 0 or 1");
        g_tune.enc_lanes = value;
        return 0;
    }
    if (!strcmp(key, "enc_keep_mib")) {
        if (value < 0) return fail("enc_keep_mib must be >= 0");
        g_tune.enc_keep_mib = value;
        return 0;
    }
    if (!strcmp(key, "merge_inflight")) {
        if (value < 1 || value > 4) return fail("merge_inflight must be in [1, 4]");
        g_tune.merge_inflight = value;
        return 0;
    }
    if (!strcmp(key, "merge_min")) {
        if (value < 1 || value > 32) return fail("merge_min must be in [1, 32]");
        g_tune.merge_min = value;
        return 0;
    }
    return fail("unknown tuning key '%s'", key);
}

mchip_ctx *mchip_ctx_create(void) {
    if (mchip_init(-1)) return nullptr;
    mchip_ctx *c = new mchip_ctx();
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_desc, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_crop, hipEventDisableTiming) != hipSuccess) {
        fail("stream/event creation failed");
        delete c;
        return nullptr;
    }
    return c;
}

void mchip_ctx_destroy(mchip_ctx *c) {
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    collect_timing(c);
    void *bufs[] = {c->d_cd, c->d_rec, c->d_plist, c->d_cinfo, c->d_binfo, c->d_scratch,
                    c->d_stage_in, c->d_stage_out, c->d_aux, c->d_ecd, c->d_einfo, c->d_levels,
                    c->d_corr, c->d_slots, c->d_plen, c->d_dst, c->d_res, c->d_fmap, c->d_lanes,
                    c->d_jobs, c->d_jscratch, c->d_jinfo, c->d_bmap, c->d_arena, c->d_order};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (c->h_desc) (void)hipHostFree(c->h_desc);
    if (c->h_res) (void)hipHostFree(c->h_res);
    if (c->h_crop) (void)hipHostFree(c->h_crop);
    if (c->ev_crop) (void)hipEventDestroy(c->ev_crop);
    for (VerifySlab &v : c->vslab) {
        if (v.h) (void)hipHostFree(v.h);
        if (v.hm) (void)hipHostFree(v.hm);
        if (v.d) (void)hipFree(v.d);
        if (v.dm) (void)hipFree(v.dm);
        if (v.ev) (void)hipEventDestroy(v.ev);
    }
    for (hipEvent_t e : c->free_events) (void)hipEventDestroy(e);
    (void)hipEventDestroy(c->ev_desc);
    if (c->stream_b) {
        (void)hipStreamSynchronize(c->stream_b);
        (void)hipStreamDestroy(c->stream_b);
        (void)hipEventDestroy(c->ev_fork);
        (void)hipEventDestroy(c->ev_join);
    }
    (void)hipStreamDestroy(c->stream);
    delete c;
}

void *mchip_ctx_stream(mchip_ctx *c) { return c ? (void *)c->stream : nullptr; }

int mchip_ctx_sync(mchip_ctx *c) {
    if (!c) return fail("null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return collect_timing(c);
}

void *mchip_malloc(size_t bytes) {
    if (mchip_init(-1)) return nullptr;
    void *p = nullptr;
    if (hipMalloc(&p, bytes + 256) != hipSuccess) { fail("hipMalloc(%zu) failed", bytes); return nullptr; }
    return p;
}
void mchip_free(void *p) { if (p) (void)hipFree(p); }

int mchip_mem_info(size_t *free_bytes, size_t *total_bytes) {
    if (mchip_init(-1)) return -1;
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return 0;
}

int mchip_h2d(mchip_ctx *c, void *dst, const void *src, size_t n) {
    HIPCHK(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}
int mchip_d2h(mchip_ctx *c, void *dst, const void *src, size_t n) {
    HIPCHK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}
int mchip_d2d(mchip_ctx *c, void *dst, const void *src, size_t n) {
    HIPCHK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}
int mchip_memset(mchip_ctx *c, void *dst, int v, size_t n) {
    HIPCHK(hipMemsetAsync(dst, v, n, c->stream));
    return 0;
}

int mchip_decode_chunks(mchip_ctx *c, const uint8_t *d_blobs, const uint64_t *blob_off,
                        const uint64_t *blob_len, const float *q, int n, uint8_t *d_out,
                        uint64_t out_stride) {
    if (!c) return fail("null context");
    if (n <= 0) return 0;
    DecodeArgs a;
    if (prepare_decode(c, d_blobs, blob_off, blob_len, q, n, d_out, out_stride, tuning_for(n, blob_len, q), a))
        return -1;
    launch_decode(c, a);
    HIPCHK(hipGetLastError());
    return 0;
}

int mchip_decode_chunks_static(const uint8_t *d_blobs, const uint64_t *blob_off, const uint64_t *blob_len,
                               const float *q, int n, uint8_t *d_out, uint64_t out_stride, void *stream) {
    if (n <= 0) return 0;
    if (n > kInlineDesc) return fail("mchip_decode_chunks_static: %d chunks (at most %d per call)", n, kInlineDesc);
    if (!d_blobs || !blob_off || !blob_len || !q || !d_out) return fail("null argument");
    if (out_stride < (uint64_t)16777216) return fail("out_stride %llu < 256^3", (unsigned long long)out_stride);
    Global &g = G();
    PriorSet pri;
    uint8_t *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g.mu);
        if (!g.ready) return fail("mchip_decode_chunks_static: call mchip_init first");
        pri = g.priors[g.cur];
        base = (uint8_t *)g.static_ws[g.device];
    }
    if (!base) return fail("mchip_decode_chunks_static: no fixed workspace on device %d", g.device);
    // workspace layout inside the fixed array, every part 256-byte aligned
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t nb = (size_t)n * kNVox;
    const size_t o_rec = 0, o_plist = al(o_rec + sizeof(BlkRec) * nb), o_cinfo = al(o_plist + sizeof(uint16_t) * nb),
                 o_binfo = al(o_cinfo + sizeof(ChunkInfo) * n), o_order = al(o_binfo + sizeof(uint2) * nb),
                 o_qt = al(o_order + sizeof(uint32_t) * ((size_t)n * 64 + 64)),
                 o_lists = al(o_qt + sizeof(QTab) * n);
    DescRef dref{};
    dref.ptr = nullptr;
    QTab tabs[kInlineDesc];
    uint64_t sw = 0;
    for (int i = 0; i < n; ++i) {
        if (blob_len[i] > 0xFFFFFFF0ull) return fail("blob %d too long", i);
        tabs[i] = make_qtab(sanitize_q(q[i]), pri.lo, pri.hi);
        dref.inl[i] = ChunkDesc{blob_off[i], blob_len[i], (uint64_t)i * out_stride, sw, (uint32_t)i, 0};
        sw += scratch_words_for(blob_len[i]);
    }
    const size_t need = o_lists + sizeof(uint32_t) * sw;
    if (need > kStaticScratchBytes)
        return fail("mchip_decode_chunks_static: these blobs need %zu bytes of workspace; the fixed workspace "
                    "holds %zu", need, kStaticScratchBytes);
    const Tuning tn = tuning_for(n, blob_len, q);
    DecodeArgs a;
    a.blobs = d_blobs;
    a.cd = dref;
    a.qtabs = (const QTab *)(base + o_qt);
    a.rec = (BlkRec *)(base + o_rec);
    a.plist = (uint16_t *)(base + o_plist);
    a.cinfo = (ChunkInfo *)(base + o_cinfo);
    a.binfo = (uint2 *)(base + o_binfo);
    a.scratch = (uint32_t *)(base + o_lists);
    a.out = d_out;
    a.n = n;
    a.decoder = tn.decoder;
    a.persistent = tn.persistent && n >= tn.persist_min && a.decoder >= 2;
    a.persist_win = tn.persist_win;
    a.len_shift = tn.len_shift;
    a.group_shift = tn.group_shift;
    a.recon_bpw = tn.recon_bpw;
    a.order = (uint32_t *)(base + o_order);
    a.pctl = a.order + (size_t)n * 64;
    a.nwg = 0;
    if (a.persistent) {
        a.nwg = tn.persist_wg > 0 ? tn.persist_wg : persistent_grid(tn.decoder, tn.persist_win);
        if (a.nwg > n * 64) a.nwg = n * 64;
    }
    hipStream_t s = (hipStream_t)stream;
    launch_put_qtabs(tabs, n, (QTab *)(base + o_qt), s);
    launch_index(a, s);
    launch_entropy(a, s);
    launch_recon(a, s);
    HIPCHK(hipGetLastError());
    return 0;
}

struct mchip_graph {
    mchip_ctx *ctx = nullptr;       // private context: its workspace belongs to the graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t stream = nullptr;   // the stream replays are launched on
    int n = 0;                      // chunks of an encode graph (0: a decode graph)
};

mchip_graph *mchip_decode_graph_create(mchip_ctx *caller, const uint8_t *d_blobs, const uint64_t *blob_off,
                                       const uint64_t *blob_len, const float *q, int n, uint8_t *d_out,
                                       uint64_t out_stride) {
    if (!caller || n <= 0) { fail("mchip_decode_graph_create: bad argument"); return nullptr; }
    auto *g = new mchip_graph;
    g->stream = caller->stream;
    g->ctx = mchip_ctx_create();
    DecodeArgs a;
    bool ok = g->ctx &&
              !prepare_decode(g->ctx, d_blobs, blob_off, blob_len, q, n, d_out, out_stride,
                              tuning_for(n, blob_len, q), a) &&
              hipStreamSynchronize(g->ctx->stream) == hipSuccess;
    if (ok) {
        // the launches only (no copies, no memsets): k_group_order resets the
        // persistent launch's counter on every replay
        const bool timing = g->ctx->timing;
        g->ctx->timing = false;
        ok = hipStreamBeginCapture(g->ctx->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            launch_decode(g->ctx, a);
            ok = hipStreamEndCapture(g->ctx->stream, &g->graph) == hipSuccess && g->graph &&
                 hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) == hipSuccess;
        }
        g->ctx->timing = timing;
    }
    if (!ok) {
        fail("mchip_decode_graph_create: capture failed");
        mchip_graph_destroy(g);
        return nullptr;
    }
    return g;
}

int mchip_graph_launch(mchip_graph *g) {
    if (!g || !g->exec) return fail("null graph");
    HIPCHK(hipGraphLaunch(g->exec, g->stream));
    return 0;
}

void mchip_graph_destroy(mchip_graph *g) {
    if (!g) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->ctx) mchip_ctx_destroy(g->ctx);
    delete g;
}

int mchip_autotune_decode(mchip_ctx *c, const uint8_t *d_blobs, const uint64_t *blob_off, const uint64_t *blob_len,
                          const float *q, int n, uint8_t *d_out, uint64_t out_stride, char *report, size_t cap) {
    if (!c || n <= 0 || !blob_len || !q) return fail("mchip_autotune_decode: bad argument");
    // Coordinate descent over the decode launch shapes, one knob at a time
    // from the batch's current tuning, each candidate timed (best of two)
    // with a private tuning; the winner is cached for this batch's signature
    // (chunk count, mean blob length, mixed qualities), so concurrent
    // decodes never see a partial change.
    Tuning best = tuning_for(n, blob_len, q);
    best.persist_min = 1;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    std::vector<std::string> entries;          // one JSON object per candidate timed
    auto time_of = [&](const Tuning &tn, float *out_ms) -> int {
        float t = 1e30f;
        for (int rep_i = 0; rep_i < 2; ++rep_i) {
            DecodeArgs a;
            HIPCHK(hipEventRecord(e0, c->stream));
            if (prepare_decode(c, d_blobs, blob_off, blob_len, q, n, d_out, out_stride, tn, a)) return -1;
            launch_decode(c, a);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(e1, c->stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < t) t = ms;
        }
        char buf[256];
        snprintf(buf, sizeof buf,
                 "{\"decoder\": %d, \"persistent\": %d, \"persist_win\": %d, \"persist_wg\": %d, "
                 "\"recon_bpw\": %d, \"len_shift\": %d, \"group_shift\": %d, \"slices\": %d, \"ms\": %.3f",
                 tn.decoder, tn.persistent, tn.persist_win, tn.persist_wg, tn.recon_bpw, tn.len_shift,
                 tn.group_shift, tn.slices, t);
        entries.push_back(buf);
        *out_ms = t;
        return 0;
    };
    float best_ms;
    if (time_of(best, &best_ms)) return -1;
    size_t best_at = 0;
    enum Knob { kDecoder, kPersistent, kPersistWin, kPersistWg, kReconBpw, kLenShift, kGroupShift, kSlices };
    struct Step { Knob k; std::vector<int> vals; };
    // kPersistWg candidates are quarters of the resident grid of the window size in use
    std::vector<Step> steps = {
        {kDecoder, {3, 2, 0}},
        {kPersistent, n >= 2 ? std::vector<int>{1, 0} : std::vector<int>{0}},
        {kPersistWin, {32, 16, 64}},
        {kPersistWg, {4, 3, 2, 1}},
        {kReconBpw, {1, 2, 4}},
        {kLenShift, {2, 3, 4}},
        {kGroupShift, {5, 6, 7}},
        {kSlices, {1, 2, 4, 8}},
    };
    // passes over the knobs until one changes nothing (at most three)
    for (int pass = 0; pass < 3; ++pass) {
        bool changed = false;
        for (const Step &st : steps) {
            for (int v : st.vals) {
                Tuning tn = best;
                switch (st.k) {
                case kDecoder: if (tn.decoder == v) continue; tn.decoder = v; break;
                case kPersistent: if (tn.persistent == v) continue; tn.persistent = v; break;
                case kPersistWin:
                    if (!tn.persistent || tn.decoder < 2 || n < 2 || tn.persist_win == v) continue;
                    tn.persist_win = v;
                    tn.persist_wg = 0;
                    break;
                case kPersistWg: {
                    if (!tn.persistent || tn.decoder < 2 || n < 2) continue;
                    const int full = persistent_grid(tn.decoder, tn.persist_win);
                    const int wg = v == 4 ? 0 : full * v / 4;
                    if (tn.persist_wg == wg) continue;
                    tn.persist_wg = wg;
                    break;
                }
                case kReconBpw: if (tn.recon_bpw == v) continue; tn.recon_bpw = v; break;
                case kLenShift: if (tn.len_shift == v) continue; tn.len_shift = v; break;
                case kGroupShift:
                    if (!tn.persistent || tn.decoder < 2 || n < 2 || tn.group_shift == v) continue;
                    tn.group_shift = v;
                    break;
                case kSlices:
                    if (!tn.persistent || tn.decoder < 2 || n <= kInlineDesc || v > n || tn.slices == v) continue;
                    tn.slices = v;
                    break;
                }
                float ms;
                if (time_of(tn, &ms)) return -1;
                if (ms < best_ms * 0.99f) { best_ms = ms; best = tn; best_at = entries.size() - 1; changed = true; }
            }
        }
        if (!changed) break;
    }
    std::string rep = "[";
    for (size_t i = 0; i < entries.size(); ++i)
        rep += (i ? ", " : "") + entries[i] + (i == best_at ? ", \"chosen\": true}" : "}");
    rep += "]";
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[workload_signature(n, blob_len, q)] = best;
    }
    if (report && cap) snprintf(report, cap, "%s", rep.c_str());
    return 0;
}

// ---------------------------------------------------------------------------
// Merged decode of concurrent one-chunk host calls (mchip_decode_chunk_merged).
//
// Each caller uploads its blob into a slot of a shared device blob arena on
// its own stream and queues a request; one worker thread issues every queued
// request as one batched decode (one index, entropy and reconstruction
// launch sequence over all of them) on its own stream, after the callers'
// uploads, and waits for those kernels before it takes the next batch, so
// the requests that arrive while a batch runs form the next one.  Each
// caller then copies its chunk out of the shared output arena on its own
// stream.  Without this, concurrent callers' kernels run one per hardware
// queue at a time (HIP's default is 4 queues, one held by the null stream).
// ---------------------------------------------------------------------------
namespace {

constexpr int kMergeSlots = 32;
constexpr size_t kMergeBlobCap = 4ull << 20;       // larger blobs take the caller's own decode
constexpr uint64_t kChunkBytes = 16777216ull;

struct MergeBatch {                                // decode event of one batch, shared by its callers
    hipEvent_t ev = nullptr;
    ~MergeBatch() { if (ev) (void)hipEventDestroy(ev); }
};

struct MergeReq {
    int slot = -1;
    uint64_t len = 0;
    float q = 8.0f;
    hipEvent_t ev_in = nullptr;                    // the caller's upload
    std::shared_ptr<MergeBatch> batch;             // set when the batch is issued
    bool issued = false;
    int rc = 0;
    std::string err;
};

constexpr int kMergeLanes = 4;                     // batches in flight at most

struct Merger {
    std::mutex mu;
    std::condition_variable cv_work, cv_issued, cv_slot;
    std::vector<int> free_slots;
    std::deque<MergeReq *> pending;
    uint8_t *d_blob = nullptr, *d_out = nullptr;   // kMergeSlots slots each
    mchip_ctx *ctx[kMergeLanes] = {};              // one context (stream + workspace) per batch in flight
    int state = 0;                                 // 0 not started, 1 running, -1 unavailable
    uint64_t batches = 0, chunks = 0;

    // Issue queued requests as batched decodes.  A batch starts when no
    // batch is in flight, or when at least merge_min requests wait and a
    // context is free (at most merge_inflight batches in flight); otherwise
    // arrivals wait for a batch to finish and join the next one.
    void run() {
        (void)hipSetDevice(G().device);
        struct Flight { int lane; std::shared_ptr<MergeBatch> b; };
        std::deque<Flight> flight;
        bool busy[kMergeLanes] = {};
        std::vector<MergeReq *> b;
        std::vector<uint64_t> off, len, oat;
        std::vector<float> q;
        for (;;) {
            while (!flight.empty() && hipEventQuery(flight.front().b->ev) != hipErrorNotReady) {
                busy[flight.front().lane] = false;
                flight.pop_front();
            }
            const Tuning tn = tuning();
            const int maxf = std::max(1, std::min(kMergeLanes, tn.merge_inflight));
            int lane = -1;
            for (int l = 0; l < maxf && lane < 0; ++l)
                if (!busy[l]) lane = l;
            {
                std::unique_lock<std::mutex> lk(mu);
                if (flight.empty()) {
                    cv_work.wait(lk, [&] { return !pending.empty(); });
                } else if (pending.empty() || lane < 0 || (int)pending.size() < tn.merge_min) {
                    // a batch is in flight: wait a little for arrivals or for it to finish
                    cv_work.wait_for(lk, std::chrono::microseconds(50));
                    continue;
                }
                b.assign(pending.begin(), pending.end());
                pending.clear();
            }
            if (lane < 0) lane = 0;
            mchip_ctx *c = ctx[lane];
            const int n = (int)b.size();
            off.resize(n); len.resize(n); oat.resize(n); q.resize(n);
            for (int i = 0; i < n; ++i) {
                off[i] = (uint64_t)b[i]->slot * kMergeBlobCap;
                len[i] = b[i]->len;
                oat[i] = (uint64_t)b[i]->slot * kChunkBytes;
                q[i] = b[i]->q;
            }
            auto batch = std::make_shared<MergeBatch>();
            int rc = hipEventCreateWithFlags(&batch->ev, hipEventDisableTiming) == hipSuccess ? 0
                         : fail("merged decode: event creation failed");
            for (int i = 0; i < n && !rc; ++i)
                if (hipStreamWaitEvent(c->stream, b[i]->ev_in, 0) != hipSuccess)
                    rc = fail("merged decode: stream wait failed");
            DecodeArgs a;
            if (!rc) rc = prepare_decode(c, d_blob, off.data(), len.data(), q.data(), n, d_out, kChunkBytes,
                                         tuning_for(n, len.data(), q.data()), a, oat.data());
            if (!rc) {
                launch_decode(c, a);
                if (hipGetLastError() != hipSuccess || hipEventRecord(batch->ev, c->stream) != hipSuccess)
                    rc = fail("merged decode: launch failed");
            }
            const std::string err = rc ? mchip_last_error() : std::string();
            {
                std::lock_guard<std::mutex> lk(mu);
                for (MergeReq *r : b) {
                    r->batch = batch;
                    r->rc = rc;
                    r->err = err;
                    r->issued = true;
                }
                batches++;
                chunks += (uint64_t)n;
            }
            cv_issued.notify_all();
            if (!rc) {
                busy[lane] = true;
                flight.push_back(Flight{lane, batch});
            }
        }
    }
};

// The process-wide merger, started on first use (its worker lives until the
// process ends).  Null when the device buffers cannot be had.
Merger *merger() {
    static Merger *m = new Merger;
    std::lock_guard<std::mutex> lk(m->mu);
    if (m->state == 0) {
        m->state = -1;
        bool ok = hipMalloc((void **)&m->d_blob, kMergeSlots * kMergeBlobCap) == hipSuccess &&
                  hipMalloc((void **)&m->d_out, kMergeSlots * kChunkBytes) == hipSuccess;
        for (int l = 0; l < kMergeLanes && ok; ++l) ok = (m->ctx[l] = mchip_ctx_create()) != nullptr;
        if (!ok) {
            if (m->d_blob) (void)hipFree(m->d_blob);
            if (m->d_out) (void)hipFree(m->d_out);
            m->d_blob = m->d_out = nullptr;
            return nullptr;
        }
        for (int i = kMergeSlots - 1; i >= 0; --i) m->free_slots.push_back(i);
        std::thread([] { m->run(); }).detach();
        m->state = 1;
    }
    return m->state == 1 ? m : nullptr;
}

}  // namespace

int mchip_decode_chunk_merged(mchip_ctx *c, const uint8_t *blob, uint64_t blob_len, uint8_t *out) {
    if (!c) return fail("null context");
    if (!blob || !out) return fail("null argument");
    Merger *m = (tuning().merge && blob_len <= kMergeBlobCap) ? merger() : nullptr;
    if (!m) return mchip_decode_chunks_host(c, &blob, &blob_len, 1, out, kChunkBytes);
    MergeReq r;
    r.len = blob_len;
    if (blob_len >= 4) memcpy(&r.q, blob, 4);
    {
        std::unique_lock<std::mutex> lk(m->mu);
        m->cv_slot.wait(lk, [&] { return !m->free_slots.empty(); });
        r.slot = m->free_slots.back();
        m->free_slots.pop_back();
    }
    auto release = [&] {
        {
            std::lock_guard<std::mutex> lk(m->mu);
            m->free_slots.push_back(r.slot);
        }
        m->cv_slot.notify_one();
    };
    r.ev_in = get_event(c);
    if (!r.ev_in ||
        (blob_len && hipMemcpyAsync(m->d_blob + (size_t)r.slot * kMergeBlobCap, blob, blob_len,
                                    hipMemcpyHostToDevice, c->stream) != hipSuccess) ||
        hipEventRecord(r.ev_in, c->stream) != hipSuccess) {
        if (r.ev_in) c->free_events.push_back(r.ev_in);
        release();
        return fail("merged decode: upload failed");
    }
    {
        std::unique_lock<std::mutex> lk(m->mu);
        m->pending.push_back(&r);
        m->cv_work.notify_one();
        m->cv_issued.wait(lk, [&] { return r.issued; });
    }
    c->free_events.push_back(r.ev_in);
    int rc = r.rc;
    if (!rc) {
        if (hipStreamWaitEvent(c->stream, r.batch->ev, 0) != hipSuccess ||
            hipMemcpyAsync(out, m->d_out + (size_t)r.slot * kChunkBytes, kChunkBytes, hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail("merged decode: copy-out failed");
    } else {
        (void)hipStreamSynchronize(c->stream);
        fail("%s", r.err.c_str());
    }
    r.batch.reset();
    release();
    return rc;
}

int mchip_merge_stats(uint64_t *batches, uint64_t *chunks) {
    Merger *m = merger();
    if (!m) return fail("merged decode unavailable");
    std::lock_guard<std::mutex> lk(m->mu);
    if (batches) *batches = m->batches;
    if (chunks) *chunks = m->chunks;
    return 0;
}

int mchip_decode_chunks_host(mchip_ctx *c, const uint8_t *const *blobs, const uint64_t *blob_len, int n,
                             uint8_t *out, uint64_t out_stride) {
    if (!c) return fail("null context");
    if (n <= 0) return 0;
    std::vector<uint64_t> off(n);
    std::vector<float> q(n);
    uint64_t tot = 0;
    for (int i = 0; i < n; ++i) {
        off[i] = tot;
        tot += (blob_len[i] + 255) & ~255ull;
        float qq = 8.0f;
        if (blob_len[i] >= 4) memcpy(&qq, blobs[i], 4);
        q[i] = qq;
    }
    if (ensure(c, &c->d_stage_in, &c->cap_stage_in, tot + 256) ||
        ensure(c, &c->d_stage_out, &c->cap_stage_out, (size_t)n * 16777216ull))
        return -1;
    uint8_t *din = (uint8_t *)c->d_stage_in, *dout = (uint8_t *)c->d_stage_out;
    for (int i = 0; i < n; ++i)
        if (blob_len[i]) HIPCHK(hipMemcpyAsync(din + off[i], blobs[i], blob_len[i], hipMemcpyHostToDevice, c->stream));
    if (mchip_decode_chunks(c, din, off.data(), blob_len, q.data(), n, dout, 16777216ull)) return -1;
    for (int i = 0; i < n; ++i)
        HIPCHK(hipMemcpyAsync(out + (size_t)i * out_stride, dout + (size_t)i * 16777216ull, 16777216ull,
                              hipMemcpyDeviceToHost, c->stream));
    return mchip_ctx_sync(c);
}

int mchip_decode_blocks_host(mchip_ctx *c, const uint8_t *const *payloads, const uint32_t *lens,
                             const float *q, int n, uint8_t *out) {
    if (!c) return fail("null context");
    if (n <= 0) return 0;
    // Present blocks go through the block-job path (the lane decoder, then
    // the reconstruction; a payload it cannot list, an empty one included,
    // falls back to the single-lane decode there); absent blocks are zeros.
    // Job records and payloads are staged together in the pinned descriptor
    // buffer: [BlockJob x n][payloads, 16-byte aligned].
    uint64_t tot = 0;
    for (int i = 0; i < n; ++i) if (payloads[i]) tot += ((uint64_t)lens[i] + 15) & ~15ull;
    const size_t meta = sizeof(BlockJob) * (size_t)n;
    const size_t pay_at = (meta + 255) & ~size_t(255);
    void *h = nullptr;
    if (desc_buffer(c, pay_at + tot, &h)) return -1;
    BlockJob *jb = (BlockJob *)h;
    uint8_t *hpay = (uint8_t *)h + pay_at;
    int nj = 0;
    uint64_t at = 0, sw = 0;
    for (int i = 0; i < n; ++i) {
        if (!payloads[i]) continue;
        uint32_t qs;
        if (qtab_slot(q[i], &qs)) return -1;
        jb[nj++] = BlockJob{at, sw, lens[i], qs, (uint32_t)i, 0};
        if (lens[i]) memcpy(hpay + at, payloads[i], lens[i]);
        at += ((uint64_t)lens[i] + 15) & ~15ull;
        sw += 8ull * lens[i] + 32ull;
    }
    if (ensure(c, &c->d_stage_in, &c->cap_stage_in, tot + 256) ||
        ensure(c, &c->d_stage_out, &c->cap_stage_out, (size_t)n * kNVox) ||
        (nj && (ensure(c, &c->d_jobs, &c->cap_jobs, sizeof(BlockJob) * nj) ||
                ensure(c, &c->d_jscratch, &c->cap_jscratch, sizeof(uint32_t) * sw) ||
                ensure(c, &c->d_jinfo, &c->cap_jinfo, sizeof(uint2) * nj))))
        return -1;
    uint8_t *din = (uint8_t *)c->d_stage_in;
    if (tot) HIPCHK(hipMemcpyAsync(din, hpay, tot, hipMemcpyHostToDevice, c->stream));
    if (nj) HIPCHK(hipMemcpyAsync(c->d_jobs, jb, sizeof(BlockJob) * nj, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->ev_desc, c->stream));
    c->desc_pending = true;
    if (nj < n) HIPCHK(hipMemsetAsync(c->d_stage_out, 0, (size_t)n * kNVox, c->stream));
    if (nj) {
        timed(c, "blocks_host", [&] {
            launch_jobs(din, (const BlockJob *)c->d_jobs, nj, qtabs_now(), (uint32_t *)c->d_jscratch,
                        (uint2 *)c->d_jinfo, (uint8_t *)c->d_stage_out, tuning().decoder, c->stream);
        });
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemcpyAsync(out, c->d_stage_out, (size_t)n * kNVox, hipMemcpyDeviceToHost, c->stream));
    return mchip_ctx_sync(c);
}

// Blocks per launch of the independent-block encode.
constexpr int kEncBatch = 32;

// Chunks per encode launch sequence.  Each chunk of the sequence needs about
// 165 MiB of workspace (levels 32 MiB, corrections 64 MiB, payload slots
// 68 MiB); larger sequences keep the device busy through each chunk's tail
// (measured on 256 mixed-quality chunks: 26 GB/s at 64 per sequence, 37 at
// 128, 49 at 256).  Automatic: a quarter of the free device memory, from 16
// to 256 chunks.
static int enc_batch_for(const Tuning &tn) {
    const int fixed = tn.enc_batch;
    if (fixed > 0) return fixed;
    constexpr size_t kPerChunk = (size_t)kNVox * kNVox * (sizeof(int16_t) + sizeof(uint32_t)) +
                                 (size_t)kNVox * kSlotBytes + (size_t)kNVox * 32 + kFmapStride;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 64;
    size_t eb = free_b / 4 / kPerChunk;
    return (int)(eb < 16 ? 16 : (eb > 256 ? 256 : eb));
}

}  // extern "C"

// Descriptors, workspace and launch arguments of an encode of m chunks at
// d_in + i*in_stride (everything but the launches); q, tau, q_store,
// out_off and out_cap are indexed from 0.
static int prepare_encode(mchip_ctx *c, const uint8_t *d_in, uint64_t in_stride, const float *q, const int *tau,
                          const float *q_store, int m, uint8_t *d_out, const uint64_t *out_off,
                          const uint64_t *out_cap, const Tuning &tn, EncodeArgs &a) {
    void *h = nullptr;
    const size_t order_at = sizeof(EncChunkDesc) * (size_t)m;
    if (desc_buffer(c, order_at + sizeof(uint32_t) * (size_t)m, &h)) return -1;
    EncChunkDesc *cd = (EncChunkDesc *)h;
    // Chunk slices of the sequence (run on two streams, see launch_encode)
    const int want = tn.enc_slices > 0 ? tn.enc_slices : 1;
    const int ns = std::max(1, std::min(want, std::min(kMaxSlices, m)));
    if (ns > 1 && !c->stream_b &&
        (hipStreamCreateWithFlags(&c->stream_b, hipStreamNonBlocking) != hipSuccess ||
         hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
         hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess))
        return fail("second stream creation failed");
    // Within a slice the entropy pass takes the chunks heaviest first (lowest
    // q, then the larger max-error bound with its corrections), so that the
    // longest waves of a mixed-quality sequence start in the first dispatch
    // round; ord[] holds each slice's order relative to its first chunk.
    uint32_t *ord = (uint32_t *)((uint8_t *)h + order_at);
    auto qkey = [&](uint32_t i) { return q[i] > 0.0f && q[i] <= 1024.0f ? q[i] : 1024.0f; };   // NaN-free
    for (int si = 0; si < ns; ++si) {
        const int c0 = (int)((int64_t)m * si / ns), c1 = (int)((int64_t)m * (si + 1) / ns);
        for (int i = c0; i < c1; ++i) ord[i] = (uint32_t)i;
        if (tn.enc_order)
            std::stable_sort(ord + c0, ord + c1, [&](uint32_t x, uint32_t y) {
                const float kx = qkey(x), ky = qkey(y);
                const int tx = tau ? tau[x] : 0, ty = tau ? tau[y] : 0;
                return kx < ky || (kx == ky && tx > ty);
            });
        for (int i = c0; i < c1; ++i) ord[i] -= (uint32_t)c0;
    }
    for (int i = 0; i < m; ++i) {
        uint32_t slot;
        if (qtab_slot(q[i], &slot)) return -1;
        cd[i].in_off = (uint64_t)i * in_stride;
        cd[i].out_off = out_off[i];
        cd[i].out_cap = out_cap[i];
        cd[i].qtab = slot;
        cd[i].tau = tau ? (tau[i] < 0 ? 0 : tau[i]) : 0;
        cd[i].q_store = q_store ? q_store[i] : q[i];
        cd[i].pad = 0;
    }
    const size_t nb = (size_t)m * kNVox;
    if (ensure(c, &c->d_ecd, &c->cap_ecd, order_at + sizeof(uint32_t) * (size_t)m) ||
        ensure(c, &c->d_einfo, &c->cap_einfo, sizeof(uint4) * nb) ||
        ensure(c, &c->d_levels, &c->cap_levels, sizeof(int16_t) * nb * kNVox) ||
        ensure(c, &c->d_corr, &c->cap_corr, sizeof(uint32_t) * nb * kNVox) ||
        ensure(c, &c->d_slots, &c->cap_slots, (size_t)kSlotBytes * nb) ||
        ensure(c, &c->d_plen, &c->cap_plen, sizeof(uint32_t) * nb) ||
        ensure(c, &c->d_dst, &c->cap_dst, sizeof(uint32_t) * nb) ||
        ensure(c, &c->d_res, &c->cap_res, sizeof(EncChunkResult) * m) ||
        ensure(c, &c->d_fmap, &c->cap_fmap, (size_t)kFmapStride * m) ||
        (tn.enc_lanes && ensure(c, &c->d_lanes, &c->cap_lanes, sizeof(uint16_t) * nb)))
        return -1;
    HIPCHK(hipMemcpyAsync(c->d_ecd, cd, order_at + sizeof(uint32_t) * (size_t)m, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->ev_desc, c->stream));
    c->desc_pending = true;
    a.in = d_in;
    a.cd = (const EncChunkDesc *)c->d_ecd;
    a.order = (const uint32_t *)((const uint8_t *)c->d_ecd + order_at);
    a.qtabs = qtabs_now();
    a.einfo = (uint4 *)c->d_einfo;
    a.levels = (int16_t *)c->d_levels;
    a.corr = (uint32_t *)c->d_corr;
    a.slots = (uint32_t *)c->d_slots;
    a.plen = (uint32_t *)c->d_plen;
    a.dst = (uint32_t *)c->d_dst;
    a.out = d_out;
    a.res = (EncChunkResult *)c->d_res;
    a.fmap = (uint32_t *)c->d_fmap;
    a.lanes = tn.enc_lanes ? (uint16_t *)c->d_lanes : nullptr;
    a.n = m;
    a.slices = ns;
    return 0;
}

// Chunks [n*i/ns, n*(i+1)/ns) of an encode sequence as a sequence of their
// own (per-chunk and per-block arrays addressed from the slice's first chunk;
// input and output offsets in the descriptors are absolute).
static EncodeArgs enc_slice_of(const EncodeArgs &a, int i, int ns) {
    const int c0 = (int)((int64_t)a.n * i / ns), c1 = (int)((int64_t)a.n * (i + 1) / ns);
    const size_t b0 = (size_t)c0 * kNVox;
    EncodeArgs s = a;
    s.n = c1 - c0;
    s.cd = a.cd + c0;
    s.einfo = a.einfo + b0;
    s.levels = a.levels + b0 * kNVox;
    s.corr = a.corr + b0 * kNVox;
    s.slots = a.slots + b0 * (kSlotBytes / 4);
    s.plen = a.plen + b0;
    s.dst = a.dst + b0;
    s.fmap = a.fmap + (size_t)c0 * (kFmapStride / 4);
    s.res = a.res + c0;
    s.order = a.order + c0;
    s.lanes = a.lanes ? a.lanes + b0 : nullptr;
    s.slices = 1;
    return s;
}

// The five launches of a sequence; with slices, slice i runs on the
// context's stream (even i) or its second stream (odd i), so that one slice's
// latency-bound passes (the blob hash is one wave per chunk) run beside the
// other slice's transform and entropy passes.  The context's stream waits for
// the second at the end (fork and join events, also inside a captured graph).
static void launch_encode(mchip_ctx *c, const EncodeArgs &a) {
    if (a.slices <= 1) {
        timed(c, "enc_transform", [&] { launch_enc_transform(a, c->stream); });
        timed(c, "enc_entropy", [&] { launch_enc_entropy(a, c->stream); });
        timed(c, "enc_assemble", [&] { launch_enc_assemble(a, c->stream); });
        timed(c, "enc_copy", [&] { launch_enc_copy(a, c->stream); });
        timed(c, "enc_hash", [&] { launch_enc_hash(a, c->stream); });
        return;
    }
    timed(c, "enc_sliced", [&] {
        (void)hipEventRecord(c->ev_fork, c->stream);
        (void)hipStreamWaitEvent(c->stream_b, c->ev_fork, 0);
        for (int i = 0; i < a.slices; ++i) {
            const EncodeArgs s = enc_slice_of(a, i, a.slices);
            hipStream_t st = (i & 1) ? c->stream_b : c->stream;
            launch_enc_transform(s, st);
            launch_enc_entropy(s, st);
            launch_enc_assemble(s, st);
            launch_enc_copy(s, st);
            launch_enc_hash(s, st);
        }
        (void)hipEventRecord(c->ev_join, c->stream_b);
        (void)hipStreamWaitEvent(c->stream, c->ev_join, 0);
    });
}

// Frees the context's encode workspace above the retention limit (tuning
// "enc_keep_mib"), once the stream is idle.
static int trim_encode_workspace(mchip_ctx *c) {
    const size_t keep = (size_t)tuning().enc_keep_mib << 20;
    const size_t held = c->cap_levels + c->cap_corr + c->cap_slots;
    if (held <= keep) return 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    void **bufs[] = {&c->d_levels, &c->d_corr, &c->d_slots};
    size_t *caps[] = {&c->cap_levels, &c->cap_corr, &c->cap_slots};
    for (int i = 0; i < 3; ++i) {
        if (*bufs[i]) HIPCHK(hipFree(*bufs[i]));
        *bufs[i] = nullptr;
        *caps[i] = 0;
    }
    return 0;
}

// The encode of mchip_encode_chunks with a given tuning (arguments checked).
static int encode_sequences(mchip_ctx *c, const uint8_t *d_in, uint64_t in_stride, const float *q, const int *tau,
                            const float *q_store, int n, uint8_t *d_out, const uint64_t *out_off,
                            const uint64_t *out_cap, uint64_t *blob_len, int *status, const Tuning &tn) {
    const int eb = enc_batch_for(tn);
    for (int b0 = 0; b0 < n; b0 += eb) {
        const int m = n - b0 < eb ? n - b0 : eb;
        if ((size_t)m > c->cap_h_res) {
            if (c->h_res) HIPCHK(hipHostFree(c->h_res));
            c->h_res = nullptr;
            HIPCHK(hipHostMalloc(&c->h_res, sizeof(EncChunkResult) * m, hipHostMallocDefault));
            c->cap_h_res = m;
        }
        EncodeArgs a;
        if (prepare_encode(c, d_in + (size_t)b0 * in_stride, in_stride, q + b0, tau ? tau + b0 : nullptr,
                           q_store ? q_store + b0 : nullptr, m, d_out, out_off + b0, out_cap + b0, tn, a))
            return -1;
        launch_encode(c, a);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(c->h_res, c->d_res, sizeof(EncChunkResult) * m, hipMemcpyDeviceToHost, c->stream));
        if (mchip_ctx_sync(c)) return -1;
        const EncChunkResult *r = (const EncChunkResult *)c->h_res;
        for (int i = 0; i < m; ++i) {
            blob_len[b0 + i] = r[i].blob_len;
            if (status) status[b0 + i] = (int)r[i].status;
        }
    }
    return trim_encode_workspace(c);
}

extern "C" {

int mchip_encode_chunks(mchip_ctx *c, const uint8_t *d_in, uint64_t in_stride, const float *q, const int *tau,
                        const float *q_store, int n, uint8_t *d_out, const uint64_t *out_off,
                        const uint64_t *out_cap, uint64_t *blob_len, int *status) {
    if (!c) return fail("null context");
    if (n <= 0) return 0;
    if (!d_in || !q || !d_out || !out_off || !out_cap || !blob_len) return fail("null argument");
    if (in_stride % 16 || ((uintptr_t)d_in) % 16) return fail("input rasters must be 16-byte aligned");
    return encode_sequences(c, d_in, in_stride, q, tau, q_store, n, d_out, out_off, out_cap, blob_len, status,
                            tuning_for_encode(n, q, tau));
}

int mchip_autotune_encode(mchip_ctx *c, const uint8_t *d_in, uint64_t in_stride, const float *q, const int *tau,
                          const float *q_store, int n, uint8_t *d_out, const uint64_t *out_off,
                          const uint64_t *out_cap, char *report, size_t cap) {
    if (!c || n <= 0 || !d_in || !q || !d_out || !out_off || !out_cap || in_stride % 16 || ((uintptr_t)d_in) % 16)
        return fail("mchip_autotune_encode: bad argument");
    // Coordinate descent over the encode shapes (chunks per launch sequence,
    // entropy-pass chunk order and block order within a chunk, chunk
    // slices), each candidate timed (best of
    // two whole encodes of the caller's chunks, which rewrite the same
    // blobs) with a private tuning; the winner is cached for this encode's
    // signature (chunk count, mixed qualities, corrections).
    Tuning best = tuning_for_encode(n, q, tau);
    std::vector<uint64_t> lens(n);
    std::vector<std::string> entries;          // one JSON object per candidate timed
    auto time_of = [&](const Tuning &tn, float *out_ms) -> int {
        double t = 1e30;
        for (int rep_i = 0; rep_i < 2; ++rep_i) {
            const auto t0 = std::chrono::steady_clock::now();
            if (encode_sequences(c, d_in, in_stride, q, tau, q_store, n, d_out, out_off, out_cap, lens.data(),
                                 nullptr, tn))
                return -1;
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (ms < t) t = ms;
        }
        char buf[192];
        snprintf(buf, sizeof buf,
                 "{\"enc_batch\": %d, \"enc_order\": %d, \"enc_lanes\": %d, \"enc_slices\": %d, \"ms\": %.3f",
                 enc_batch_for(tn), tn.enc_order, tn.enc_lanes, tn.enc_slices > 0 ? tn.enc_slices : 1, t);
        entries.push_back(buf);
        *out_ms = (float)t;
        return 0;
    };
    float best_ms;
    if (time_of(best, &best_ms)) return -1;
    size_t best_at = 0;
    const int full = enc_batch_for(best);
    enum Knob { kOrder, kLanes, kBatch, kSlices };
    struct Step { Knob k; std::vector<int> vals; };
    const std::vector<Step> steps = {
        {kOrder, {1, 0}},
        {kLanes, {0, 1}},
        {kBatch, {0, std::max(1, full / 2), std::max(1, full / 4)}},
        {kSlices, {1, 2}},
    };
    // passes over the knobs until one changes nothing (at most three)
    for (int pass = 0; pass < 3; ++pass) {
        bool changed = false;
        for (const Step &st : steps) {
            for (int v : st.vals) {
                Tuning tn = best;
                switch (st.k) {
                case kOrder: if (tn.enc_order == v) continue; tn.enc_order = v; break;
                case kLanes: if (tn.enc_lanes == v) continue; tn.enc_lanes = v; break;
                case kBatch:
                    if (tn.enc_batch == v || (v > 0 && v >= std::min(n, full))) continue;
                    tn.enc_batch = v;
                    break;
                case kSlices:
                    if ((tn.enc_slices > 0 ? tn.enc_slices : 1) == v || v > std::min(n, enc_batch_for(tn))) continue;
                    tn.enc_slices = v;
                    break;
                }
                float ms;
                if (time_of(tn, &ms)) return -1;
                if (ms < best_ms * 0.99f) { best_ms = ms; best = tn; best_at = entries.size() - 1; changed = true; }
            }
        }
        if (!changed) break;
    }
    std::string rep = "[";
    for (size_t i = 0; i < entries.size(); ++i)
        rep += (i ? ", " : "") + entries[i] + (i == best_at ? ", \"chosen\": true}" : "}");
    rep += "]";
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned_enc[encode_signature(n, q, tau)] = best;
    }
    if (report && cap) snprintf(report, cap, "%s", rep.c_str());
    return 0;
}

mchip_graph *mchip_encode_graph_create(mchip_ctx *caller, const uint8_t *d_in, uint64_t in_stride, const float *q,
                                       const int *tau, const float *q_store, int n, uint8_t *d_out,
                                       const uint64_t *out_off, const uint64_t *out_cap) {
    if (!caller || n <= 0 || !d_in || !q || !d_out || !out_off || !out_cap || in_stride % 16 ||
        ((uintptr_t)d_in) % 16) {
        fail("mchip_encode_graph_create: bad argument");
        return nullptr;
    }
    const Tuning tn = tuning_for_encode(n, q, tau);
    if (n > enc_batch_for(tn)) {
        fail("mchip_encode_graph_create: %d chunks exceed one encode sequence (%d)", n, enc_batch_for(tn));
        return nullptr;
    }
    auto *g = new mchip_graph;
    g->stream = caller->stream;
    g->ctx = mchip_ctx_create();
    g->n = n;
    EncodeArgs a;
    bool ok = g->ctx && !prepare_encode(g->ctx, d_in, in_stride, q, tau, q_store, n, d_out, out_off, out_cap, tn, a) &&
              hipStreamSynchronize(g->ctx->stream) == hipSuccess;
    if (ok) {
        // the five launches only: descriptors stay in the graph's workspace
        // and every kernel rewrites what it reads later, so replays need no
        // copies or memsets
        const bool timing = g->ctx->timing;
        g->ctx->timing = false;
        ok = hipStreamBeginCapture(g->ctx->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            launch_encode(g->ctx, a);
            ok = hipStreamEndCapture(g->ctx->stream, &g->graph) == hipSuccess && g->graph &&
                 hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) == hipSuccess;
        }
        g->ctx->timing = timing;
    }
    if (!ok) {
        fail("mchip_encode_graph_create: capture failed");
        mchip_graph_destroy(g);
        return nullptr;
    }
    return g;
}

int mchip_encode_graph_results(mchip_graph *g, uint64_t *blob_len, int *status) {
    if (!g || !g->exec || g->n <= 0 || !blob_len) return fail("mchip_encode_graph_results: bad argument");
    mchip_ctx *c = g->ctx;
    std::vector<EncChunkResult> r(g->n);
    HIPCHK(hipStreamSynchronize(g->stream));
    HIPCHK(hipMemcpyAsync(r.data(), c->d_res, sizeof(EncChunkResult) * g->n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < g->n; ++i) {
        blob_len[i] = r[i].blob_len;
        if (status) status[i] = (int)r[i].status;
    }
    return 0;
}

int mchip_encode_chunks_host(mchip_ctx *c, const uint8_t *chunks, const float *q, const int *tau,
                             const float *q_store, int n, uint8_t *const *blobs, const uint64_t *caps,
                             uint64_t *blob_len, int *status) {
    if (!c) return fail("null context");
    if (n <= 0) return 0;
    const uint64_t CH = 16777216ull;
    std::vector<uint64_t> off(n), cap(n);
    uint64_t tot = 0;
    for (int i = 0; i < n; ++i) {
        off[i] = tot;
        cap[i] = caps[i];
        tot += (caps[i] + 255) & ~255ull;
    }
    std::vector<int> st(n, 0);
    const int eb = enc_batch_for(tuning_for_encode(n, q, tau));
    for (int b0 = 0; b0 < n; b0 += eb) {
        const int m = n - b0 < eb ? n - b0 : eb;
        uint64_t base = off[b0], span = 0;
        for (int i = 0; i < m; ++i) span = off[b0 + i] + ((caps[b0 + i] + 255) & ~255ull) - base;
        if (ensure(c, &c->d_stage_in, &c->cap_stage_in, (size_t)m * CH + 256) ||
            ensure(c, &c->d_stage_out, &c->cap_stage_out, span + 256))
            return -1;
        HIPCHK(hipMemcpyAsync(c->d_stage_in, chunks + (size_t)b0 * CH, (size_t)m * CH, hipMemcpyHostToDevice,
                              c->stream));
        std::vector<uint64_t> ro(m);
        for (int i = 0; i < m; ++i) ro[i] = off[b0 + i] - base;
        if (mchip_encode_chunks(c, (const uint8_t *)c->d_stage_in, CH, q + b0, tau ? tau + b0 : nullptr,
                                q_store ? q_store + b0 : nullptr, m, (uint8_t *)c->d_stage_out, ro.data(),
                                cap.data() + b0, blob_len + b0, st.data() + b0))
            return -1;
        for (int i = 0; i < m; ++i)
            if (st[b0 + i] == 0 && blob_len[b0 + i])
                HIPCHK(hipMemcpyAsync(blobs[b0 + i], (const uint8_t *)c->d_stage_out + ro[i], blob_len[b0 + i],
                                      hipMemcpyDeviceToHost, c->stream));
        if (mchip_ctx_sync(c)) return -1;
    }
    if (status)
        for (int i = 0; i < n; ++i) status[i] = st[i];
    return 0;
}

int mchip_encode_blocks_host(mchip_ctx *c, const uint8_t *vox, const float *q, const int *tau, int n,
                             uint8_t *const *payloads, uint32_t *lens, int *status) {
    if (!c) return fail("null context");
    if (n <= 0) return 0;
    if (!vox || !q || !payloads || !lens) return fail("null argument");
    const int kMax = kEncBatch * kNVox;
    for (int b0 = 0; b0 < n; b0 += kMax) {
        const int m = n - b0 < kMax ? n - b0 : kMax;
        void *h = nullptr;
        if (desc_buffer(c, sizeof(EncChunkDesc) * (size_t)m, &h)) return -1;
        EncChunkDesc *cd = (EncChunkDesc *)h;
        for (int i = 0; i < m; ++i) {
            uint32_t slot;
            if (qtab_slot(q[b0 + i], &slot)) return -1;
            cd[i] = EncChunkDesc{0, 0, 0, slot, tau ? (tau[b0 + i] < 0 ? 0 : tau[b0 + i]) : 0, q[b0 + i], 0};
        }
        if (ensure(c, &c->d_ecd, &c->cap_ecd, sizeof(EncChunkDesc) * m) ||
            ensure(c, &c->d_einfo, &c->cap_einfo, sizeof(uint4) * m) ||
            ensure(c, &c->d_levels, &c->cap_levels, sizeof(int16_t) * (size_t)m * kNVox) ||
            ensure(c, &c->d_corr, &c->cap_corr, sizeof(uint32_t) * (size_t)m * kNVox) ||
            ensure(c, &c->d_slots, &c->cap_slots, (size_t)kSlotBytes * m) ||
            ensure(c, &c->d_plen, &c->cap_plen, sizeof(uint32_t) * m) ||
            ensure(c, &c->d_stage_in, &c->cap_stage_in, (size_t)m * kNVox + 256))
            return -1;
        HIPCHK(hipMemcpyAsync(c->d_stage_in, vox + (size_t)b0 * kNVox, (size_t)m * kNVox, hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipMemcpyAsync(c->d_ecd, cd, sizeof(EncChunkDesc) * m, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipEventRecord(c->ev_desc, c->stream));
        c->desc_pending = true;
        EncodeArgs a{};
        a.in = (const uint8_t *)c->d_stage_in;
        a.cd = (const EncChunkDesc *)c->d_ecd;
        a.qtabs = qtabs_now();
        a.einfo = (uint4 *)c->d_einfo;
        a.levels = (int16_t *)c->d_levels;
        a.corr = (uint32_t *)c->d_corr;
        a.slots = (uint32_t *)c->d_slots;
        a.plen = (uint32_t *)c->d_plen;
        a.n = m;
        timed(c, "enc_blocks", [&] { launch_enc_blocks(a, c->stream); });
        HIPCHK(hipGetLastError());
        std::vector<uint32_t> pl(m);
        HIPCHK(hipMemcpyAsync(pl.data(), c->d_plen, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, c->stream));
        if (mchip_ctx_sync(c)) return -1;
        for (int i = 0; i < m; ++i) {
            const uint32_t len = pl[i];
            lens[b0 + i] = len;
            if (status) status[b0 + i] = len > (uint32_t)kSlotBytes ? 1 : 0;
            if (len && len <= (uint32_t)kSlotBytes)
                HIPCHK(hipMemcpyAsync(payloads[b0 + i], (const uint8_t *)c->d_slots + (size_t)i * kSlotBytes, len,
                                      hipMemcpyDeviceToHost, c->stream));
        }
        if (mchip_ctx_sync(c)) return -1;
    }
    return 0;
}

int mchip_decode_block_jobs(mchip_ctx *c, const uint8_t *d_pay, const uint64_t *pay_off, const uint32_t *pay_len,
                            const float *q, const uint32_t *slot, int n, uint8_t *d_arena) {
    if (!c) return fail("null context");
    if (n <= 0) return 0;
    if (!d_pay || !pay_off || !pay_len || !q || !slot || !d_arena) return fail("null argument");
    void *h = nullptr;
    if (desc_buffer(c, sizeof(BlockJob) * (size_t)n, &h)) return -1;
    BlockJob *jb = (BlockJob *)h;
    uint64_t sw = 0;
    for (int i = 0; i < n; ++i) {
        uint32_t qs;
        if (qtab_slot(q[i], &qs)) return -1;
        jb[i] = BlockJob{pay_off[i], sw, pay_len[i], qs, slot[i], 0};
        sw += 8ull * pay_len[i] + 32ull;
    }
    if (ensure(c, &c->d_jobs, &c->cap_jobs, sizeof(BlockJob) * n) ||
        ensure(c, &c->d_jscratch, &c->cap_jscratch, sizeof(uint32_t) * sw) ||
        ensure(c, &c->d_jinfo, &c->cap_jinfo, sizeof(uint2) * n))
        return -1;
    HIPCHK(hipMemcpyAsync(c->d_jobs, jb, sizeof(BlockJob) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->ev_desc, c->stream));
    c->desc_pending = true;
    timed(c, "block_jobs", [&] {
        launch_jobs(d_pay, (const BlockJob *)c->d_jobs, n, qtabs_now(), (uint32_t *)c->d_jscratch,
                    (uint2 *)c->d_jinfo, d_arena, tuning().decoder, c->stream);
    });
    HIPCHK(hipGetLastError());
    return 0;
}

int mchip_crop_boxes(mchip_ctx *c, const uint8_t *d_arena, const int32_t *bmap, const int64_t *origin, int n, int gz,
                     int gy, int gx, int64_t dz, int64_t dy, int64_t dx, uint8_t *d_out, uint64_t batch_stride) {
    if (!c) return fail("null context");
    if (n <= 0) return 0;
    if (!d_arena || !bmap || !origin || !d_out || gz <= 0 || gy <= 0 || gx <= 0 || dz <= 0 || dy <= 0 || dx <= 0)
        return fail("bad argument");
    const size_t cells = (size_t)n * gz * gy * gx;
    const size_t bytes = cells * 4 + (size_t)n * 24;
    if (c->crop_pending) { HIPCHK(hipEventSynchronize(c->ev_crop)); c->crop_pending = false; }
    if (bytes > c->cap_h_crop) {
        if (c->h_crop) HIPCHK(hipHostFree(c->h_crop));
        c->h_crop = nullptr;
        HIPCHK(hipHostMalloc(&c->h_crop, bytes + bytes / 2 + 4096, hipHostMallocDefault));
        c->cap_h_crop = bytes + bytes / 2 + 4096;
    }
    memcpy(c->h_crop, origin, (size_t)n * 24);
    memcpy((uint8_t *)c->h_crop + (size_t)n * 24, bmap, cells * 4);
    if (ensure(c, &c->d_bmap, &c->cap_bmap, bytes)) return -1;
    HIPCHK(hipMemcpyAsync(c->d_bmap, c->h_crop, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->ev_crop, c->stream));
    c->crop_pending = true;
    const int64_t *dorig = (const int64_t *)c->d_bmap;
    const int32_t *dbm = (const int32_t *)((const uint8_t *)c->d_bmap + (size_t)n * 24);
    timed(c, "crop", [&] {
        launch_crop(d_arena, dbm, dorig, n, gz, gy, gx, dz, dy, dx, d_out, batch_stride, c->stream);
    });
    HIPCHK(hipGetLastError());
    return 0;
}

int mchip_read_regions_host(mchip_ctx *c, int nblk, const uint8_t *const *payloads, const uint32_t *lens,
                            const float *q, const int32_t *bmap, const int64_t *origin, int n, int gz, int gy,
                            int gx, int64_t dz, int64_t dy, int64_t dx, uint8_t *out, uint64_t batch_stride) {
    if (!c) return fail("null context");
    if (n <= 0) return 0;
    std::vector<uint64_t> po(nblk > 0 ? nblk : 1);
    std::vector<uint32_t> sl(nblk > 0 ? nblk : 1);
    uint64_t tot = 0;
    for (int i = 0; i < nblk; ++i) { po[i] = tot; sl[i] = (uint32_t)i; tot += ((uint64_t)lens[i] + 15) & ~15ull; }
    const uint64_t box = (uint64_t)dz * dy * dx;
    if (ensure(c, &c->d_stage_in, &c->cap_stage_in, tot + 256) ||
        ensure(c, &c->d_arena, &c->cap_arena, (size_t)(nblk > 0 ? nblk : 1) * kNVox) ||
        ensure(c, &c->d_stage_out, &c->cap_stage_out, box * n))
        return -1;
    uint8_t *din = (uint8_t *)c->d_stage_in;
    // gather payloads into one pinned upload
    if (tot) {
        void *h = nullptr;
        if (desc_buffer(c, tot, &h)) return -1;
        for (int i = 0; i < nblk; ++i) memcpy((uint8_t *)h + po[i], payloads[i], lens[i]);
        HIPCHK(hipMemcpyAsync(din, h, tot, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipEventRecord(c->ev_desc, c->stream));
        c->desc_pending = true;
    }
    if (nblk > 0 && mchip_decode_block_jobs(c, din, po.data(), lens, q, sl.data(), nblk, (uint8_t *)c->d_arena))
        return -1;
    if (mchip_crop_boxes(c, (const uint8_t *)c->d_arena, bmap, origin, n, gz, gy, gx, dz, dy, dx,
                         (uint8_t *)c->d_stage_out, box))
        return -1;
    if (batch_stride == box) {
        HIPCHK(hipMemcpyAsync(out, c->d_stage_out, box * n, hipMemcpyDeviceToHost, c->stream));
    } else {
        for (int i = 0; i < n; ++i)
            HIPCHK(hipMemcpyAsync(out + (size_t)i * batch_stride, (const uint8_t *)c->d_stage_out + (size_t)i * box,
                                  box, hipMemcpyDeviceToHost, c->stream));
    }
    return mchip_ctx_sync(c);
}

int mchip_xxh64_ranges(mchip_ctx *c, const uint8_t *d_base, const uint64_t *off, const uint64_t *len,
                       const uint64_t *seed, int n, uint64_t *out) {
    if (!c) return fail("null context");
    if (n <= 0) return 0;
    if (!d_base || !off || !len || !seed || !out) return fail("null argument");
    const size_t m = 8 * (size_t)n;
    void *h = nullptr;
    if (desc_buffer(c, 4 * m, &h) || ensure(c, &c->d_aux, &c->cap_aux, 4 * m)) return -1;
    uint8_t *hb = (uint8_t *)h, *db = (uint8_t *)c->d_aux;
    memcpy(hb, off, m);
    memcpy(hb + m, len, m);
    memcpy(hb + 2 * m, seed, m);
    HIPCHK(hipMemcpyAsync(db, hb, 3 * m, hipMemcpyHostToDevice, c->stream));
    timed(c, "xxh64", [&] {
        launch_xxh64_ranges(d_base, (const uint64_t *)db, (const uint64_t *)(db + m), (const uint64_t *)(db + 2 * m),
                            n, (uint64_t *)(db + 3 * m), c->stream);
    });
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hb + 3 * m, db + 3 * m, m, hipMemcpyDeviceToHost, c->stream));
    if (mchip_ctx_sync(c)) return -1;
    memcpy(out, hb + 3 * m, m);
    return 0;
}

long mchip_verify_blobs(mchip_ctx *c, const uint8_t *d_base, const uint64_t *off, const uint64_t *len, int n,
                        uint64_t *hash_out, uint8_t *bad) {
    if (!c) return fail("null context");
    if (n <= 0) return 0;
    if (!d_base || !off || !len) return fail("null argument");
    const size_t m = 8 * (size_t)n;
    const size_t bytes = 3 * m + (size_t)n;
    void *h = nullptr;
    if (desc_buffer(c, bytes, &h) || ensure(c, &c->d_aux, &c->cap_aux, bytes)) return -1;
    uint8_t *hb = (uint8_t *)h, *db = (uint8_t *)c->d_aux;
    memcpy(hb, off, m);
    memcpy(hb + m, len, m);
    HIPCHK(hipMemcpyAsync(db, hb, 2 * m, hipMemcpyHostToDevice, c->stream));
    timed(c, "blob_hash", [&] {
        launch_blob_hash(d_base, (const uint64_t *)db, (const uint64_t *)(db + m), n, (uint64_t *)(db + 2 * m),
                         db + 3 * m, c->stream);
    });
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hb + 2 * m, db + 2 * m, m + (size_t)n, hipMemcpyDeviceToHost, c->stream));
    if (mchip_ctx_sync(c)) return -1;
    long nbad = 0;
    for (int i = 0; i < n; ++i) nbad += hb[3 * m + i];
    if (hash_out) memcpy(hash_out, hb + 2 * m, m);
    if (bad) memcpy(bad, hb + 3 * m, (size_t)n);
    return nbad;
}

long mchip_verify_blobs_host(mchip_ctx *c, const uint8_t *const *blobs, const uint64_t *len, int n,
                             uint64_t *hash_out, uint8_t *bad) {
    if (!c) return fail("null context");
    if (n <= 0) return 0;
    if (!blobs || !len) return fail("null argument");
    auto r16 = [](uint64_t x) -> uint64_t { return (x + 15) & ~(uint64_t)15; };
    uint64_t biggest = 0;
    for (int i = 0; i < n; ++i) biggest = std::max(biggest, r16(len[i]));
    const size_t cap = std::max<size_t>(256ull << 20, biggest);
    const size_t capm = 25 * (size_t)n + 256;
    for (VerifySlab &v : c->vslab) {
        if (v.pending) { HIPCHK(hipEventSynchronize(v.ev)); v.pending = false; }
        if (!v.ev) HIPCHK(hipEventCreateWithFlags(&v.ev, hipEventDisableTiming));
        if (v.cap < cap) {
            if (v.h) HIPCHK(hipHostFree(v.h));
            if (v.d) HIPCHK(hipFree(v.d));
            v.h = v.d = nullptr;
            v.cap = 0;
            HIPCHK(hipHostMalloc((void **)&v.h, cap, hipHostMallocDefault));
            HIPCHK(hipMalloc((void **)&v.d, cap));
            v.cap = cap;
        }
        if (v.capm < capm) {
            if (v.hm) HIPCHK(hipHostFree(v.hm));
            if (v.dm) HIPCHK(hipFree(v.dm));
            v.hm = v.dm = nullptr;
            v.capm = 0;
            HIPCHK(hipHostMalloc((void **)&v.hm, capm, hipHostMallocDefault));
            HIPCHK(hipMalloc((void **)&v.dm, capm));
            v.capm = capm;
        }
    }
    const int nthr = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    long nbad = 0;
    auto harvest = [&](VerifySlab &v) -> int {
        HIPCHK(hipEventSynchronize(v.ev));
        v.pending = false;
        const int k = v.i1 - v.i0;
        const uint64_t *hh = (const uint64_t *)(v.hm + 16 * (size_t)k);
        const uint8_t *hbad = v.hm + 24 * (size_t)k;
        for (int j = 0; j < k; ++j) {
            nbad += hbad[j];
            if (hash_out) hash_out[v.i0 + j] = hh[j];
            if (bad) bad[v.i0 + j] = hbad[j];
        }
        return 0;
    };
    int i = 0, turn = 0;
    while (i < n) {
        VerifySlab &v = c->vslab[turn++ & 1];
        if (v.pending && harvest(v)) return -1;
        // pack the next blobs into this slab
        v.i0 = i;
        std::vector<uint64_t> at;
        size_t pos = 0;
        while (i < n && pos + r16(len[i]) <= v.cap) { at.push_back(pos); pos += r16(len[i]); ++i; }
        v.i1 = i;
        const int k = v.i1 - v.i0;
        // gather: pieces of at most 4 MiB, taken by host threads in turn
        struct Piece { const uint8_t *src; uint8_t *dst; size_t n; };
        std::vector<Piece> pieces;
        for (int j = 0; j < k; ++j)
            for (uint64_t o = 0; o < len[v.i0 + j]; o += 4ull << 20)
                pieces.push_back({blobs[v.i0 + j] + o, v.h + at[j] + o,
                                  (size_t)std::min<uint64_t>(4ull << 20, len[v.i0 + j] - o)});
        std::atomic<size_t> next{0};
        auto work = [&] {
            for (size_t p; (p = next.fetch_add(1)) < pieces.size();) memcpy(pieces[p].dst, pieces[p].src, pieces[p].n);
        };
        std::vector<std::thread> pool;
        const int t = (int)std::min<size_t>((size_t)nthr, pieces.size());
        for (int w = 1; w < t; ++w) pool.emplace_back(work);
        work();
        for (std::thread &th : pool) th.join();
        uint64_t *mo = (uint64_t *)v.hm, *ml = mo + k;
        for (int j = 0; j < k; ++j) { mo[j] = at[j]; ml[j] = len[v.i0 + j]; }
        HIPCHK(hipMemcpyAsync(v.d, v.h, pos, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(v.dm, v.hm, 16 * (size_t)k, hipMemcpyHostToDevice, c->stream));
        timed(c, "blob_hash", [&] {
            launch_blob_hash(v.d, (const uint64_t *)v.dm, (const uint64_t *)(v.dm + 8 * (size_t)k), k,
                             (uint64_t *)(v.dm + 16 * (size_t)k), v.dm + 24 * (size_t)k, c->stream);
        });
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(v.hm + 16 * (size_t)k, v.dm + 16 * (size_t)k, 9 * (size_t)k, hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipEventRecord(v.ev, c->stream));
        v.pending = true;
    }
    for (int w = 0; w < 2; ++w) {
        VerifySlab &v = c->vslab[turn++ & 1];
        if (v.pending && harvest(v)) return -1;
    }
    return nbad;
}

int mchip_render_quad(mchip_ctx *c, int nlods, const uint8_t *const *d_levels, const int32_t *dims,
                      const float *grid, int gw, int gh, float x0, float y0, float step, int w, int h, int filter,
                      int comp, float t0, float t1, float dt, uint8_t *d_out, int *lod_out) {
    if (!c) return fail("null context");
    if (nlods < 1 || nlods > 16 || !d_levels || !dims || !grid || gw < 1 || gh < 1 || w <= 0 || h <= 0 || !d_out ||
        (int64_t)w * h > (1ll << 30))
        return fail("mchip_render_quad: bad argument");
    for (int l = 0; l < nlods; ++l)
        if (!d_levels[l] || dims[3 * l] <= 0 || dims[3 * l + 1] <= 0 || dims[3 * l + 2] <= 0)
            return fail("mchip_render_quad: level %d is empty", l);
    const size_t gb = 12 * (size_t)gw * gh;
    void *hg = nullptr;
    if (desc_buffer(c, gb, &hg) || ensure(c, &c->d_aux, &c->cap_aux, gb)) return -1;
    RenderDesc d;
    const int L = render_describe(d, nlods, dims, grid, (const float *)c->d_aux, gw, gh, x0, y0, step, w, h, filter,
                                  comp, t0, t1, dt);
    if (L < 0) return fail("mchip_render_quad: filter %d / composite %d not supported", filter, comp);
    memcpy(hg, grid, gb);
    HIPCHK(hipMemcpyAsync(c->d_aux, hg, gb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->ev_desc, c->stream));
    c->desc_pending = true;
    timed(c, "render", [&] { launch_render_dense(d, d_levels[L], d_out, c->stream); });
    HIPCHK(hipGetLastError());
    if (lod_out) *lod_out = L;
    return 0;
}

int mchip_decimate(mchip_ctx *c, const uint8_t *d_src, int dz, int dy, int dx, uint8_t *d_out) {
    if (!c) return fail("null context");
    if (!d_src || !d_out || dz <= 0 || dy <= 0 || dx <= 0) return fail("mchip_decimate: bad argument");
    timed(c, "decimate", [&] { launch_decimate(d_src, dz, dy, dx, d_out, c->stream); });
    HIPCHK(hipGetLastError());
    return 0;
}

int mchip_decimate_chunks(mchip_ctx *c, const uint8_t *const *d_src, uint8_t *d_out) {
    if (!c) return fail("null context");
    if (!d_src || !d_out) return fail("mchip_decimate_chunks: bad argument");
    ChunkGroup g;
    for (int k = 0; k < 8; ++k) g.c[k] = d_src[k];
    timed(c, "decimate", [&] { launch_decimate_group(g, d_out, c->stream); });
    HIPCHK(hipGetLastError());
    return 0;
}

int mchip_deblock(mchip_ctx *c, uint8_t *d_vol, int nz, int ny, int nx, float quality) {
    if (!c) return fail("null context");
    if (!d_vol || nz <= 0 || ny <= 0 || nx <= 0) return fail("mchip_deblock: bad argument");
    timed(c, "deblock", [&] { launch_deblock(d_vol, nz, ny, nx, quality, c->stream); });
    HIPCHK(hipGetLastError());
    return 0;
}

int mchip_deblock_host(mchip_ctx *c, uint8_t *vol, int nz, int ny, int nx, float quality) {
    if (!c) return fail("null context");
    if (!vol || nz <= 0 || ny <= 0 || nx <= 0) return fail("mchip_deblock_host: bad argument");
    const size_t n = (size_t)nz * ny * nx;
    if (ensure(c, &c->d_stage_out, &c->cap_stage_out, n)) return -1;
    HIPCHK(hipMemcpyAsync(c->d_stage_out, vol, n, hipMemcpyHostToDevice, c->stream));
    if (mchip_deblock(c, (uint8_t *)c->d_stage_out, nz, ny, nx, quality)) return -1;
    HIPCHK(hipMemcpyAsync(vol, c->d_stage_out, n, hipMemcpyDeviceToHost, c->stream));
    return mchip_ctx_sync(c);
}

int mchip_ctx_enable_timing(mchip_ctx *c, int on) {
    if (!c) return fail("null context");
    c->timing = on != 0;
    return 0;
}
int mchip_ctx_kernel_time(mchip_ctx *c, const char *kernel, double *total_ms, long *launches) {
    if (!c || !kernel) return fail("null argument");
    if (collect_timing(c)) return -1;
    auto it = c->stats.find(kernel);
    if (total_ms) *total_ms = it == c->stats.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == c->stats.end() ? 0 : it->second.second;
    return 0;
}
int mchip_ctx_reset_timing(mchip_ctx *c) {
    if (!c) return fail("null context");
    if (collect_timing(c)) return -1;
    c->stats.clear();
    return 0;
}

}  // extern "C"
