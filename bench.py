"""Benchmark of the MI355X 16^3-block codec: bit-exact chunk decode.

Workload (default): 64 synthetic 256^3 u8 chunks (seeds 1..64, 41.9% air,
1 GiB decoded) encoded at q=6; one step decodes all 64 chunks from HBM-
resident blobs into HBM, one chunk per launch.  Every decoded chunk is
checked against the reference's decoded xxh64 (tests/golden/
chunk_manifest.json) after the timed region.

Prints ONE JSON line (see the repository's benchmark contract).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
# Per-chunk launches overlap across streams only as far as the process has
# hardware queues (HIP's default is 4): a hardware queue runs one kernel at a
# time, so two streams on one queue take turns.  The runtime's null stream
# holds one queue, so the per-chunk leg uses QUEUES - 1 streams, one per
# remaining queue.  Measured on MI355X (64 chunks of q=6): 73 GB/s with 32
# queues shared by 64 streams, 92 with 16 queues, 111 with 22, and 73 again
# with 24.  The runtime reads the variable when it starts, so it is set
# before the library is loaded.
QUEUES = 22
# The serving leg runs in a child process of its own (--serve-only) with as
# many hardware queues as it has streams: idle queues left by the per-chunk
# leg would otherwise share the hardware's queue slots with the serving
# lanes and delay their dispatches.
SERVE_LANES = 8
# The drop-in leg (--dropin-only) runs at HIP's default queue count, as a
# program that links the library and sets nothing would.
if "--dropin-only" in sys.argv:
    os.environ.pop("GPU_MAX_HW_QUEUES", None)
else:
    os.environ["GPU_MAX_HW_QUEUES"] = str(2 * SERVE_LANES if "--serve-only" in sys.argv else QUEUES)
sys.path.insert(0, os.path.join(ROOT, "matter-compressor_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import mc_amd  # noqa: E402
import workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def xxh64_hex(a):
    import xxhash
    return f"{xxhash.xxh64(memoryview(np.ascontiguousarray(a)).cast('B'), seed=0).intdigest():016x}"


def profile_value(workload, kernel, kind, key):
    """A figure from a committed rocprofv3 summary
    profiles/{kernel}_{workload}_{kind}.txt (a line `key: value`), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_{kind}.txt")
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), or None."""
    v = profile_value(workload, kernel, "traffic", "hbm_bytes_per_launch")
    return None if v is None else int(v)


def cpu_threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except Exception:
        return max(1, min(16, os.cpu_count() or 1))


def cpu_encode_and_baseline(chunks, specs, nthreads):
    """CPU baseline leg of the oracle port: it encodes the given chunks and
    decodes them again, nthreads threads each way."""
    import oracle_bind as ob
    O = ob.oracle()
    n = len(chunks)
    caps = np.full(n, 20 << 20, np.uint64)
    bufs = [np.zeros(20 << 20, np.uint8) for _ in range(n)]
    ptrs = (C.c_void_p * n)(*[b.ctypes.data for b in bufs])
    lens = np.zeros(n, np.uint64)
    q = np.array([s["q"] for s in specs], np.float32)
    tau = np.array([s["tau"] for s in specs], np.int32)
    t0 = time.perf_counter()
    O.orc_encode_chunks_mt(n, ob.p(chunks), ob.p(q), ob.p(tau), ptrs, ob.p(caps), ob.p(lens), nthreads)
    t_enc = time.perf_counter() - t0
    blobs = [bufs[i][: int(lens[i])].tobytes() for i in range(n)]
    # decode sample: enough chunks for ~10 s of single-core work
    ns = min(n, max(nthreads, 8))
    barr = [np.frombuffer(b, np.uint8) for b in blobs[:ns]]
    bp = (C.c_void_p * ns)(*[a.ctypes.data for a in barr])
    bl = np.array([a.size for a in barr], np.uint64)
    out = np.empty((ns, 256, 256, 256), np.uint8)
    t0 = time.perf_counter()
    O.orc_decode_chunks_mt(ns, bp, ob.p(bl), ob.p(out), nthreads)
    t_dec = time.perf_counter() - t0
    enc_gbs = n * mc_amd.CHUNK_BYTES / t_enc / 1e9
    dec_gbs = ns * mc_amd.CHUNK_BYTES / t_dec / 1e9
    return blobs, enc_gbs, dec_gbs, ns


REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def _ref_lib(name):
    vp = C.c_void_p
    R = C.CDLL(os.path.join(REF_DIR, name))
    R.mc_archive_open_dims.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_float]
    R.mc_archive_open_dims.restype = vp
    R.mc_archive_close.argtypes = [vp]
    R.mc_archive_append_chunk_compressed.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_size_t]
    R.mc_archive_append_chunk_par.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_float, C.c_int]
    R.mc_archive_chunk_offset.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    R.mc_archive_chunk_offset.restype = C.c_uint64
    R.mc_archive_decode_chunk.argtypes = [vp, C.c_uint64, vp, C.c_int]
    R.mc_codec_init()
    return R


def _timed(fn, min_s=1.0):
    """Seconds per call of fn(), repeated to at least min_s."""
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= min_s:
            return dt / reps


def cpu_reference_baseline(chunks, blobs, man, nthreads):
    """The reference's own CPU decode and encode, compiled from its sources
    (oracle/_ref: libmcref_fast.so with its shipped fast-math flags,
    libmcref.so with its strict flags), timed on this host's cores over the
    given chunks, or None when those builds are absent.

    Decode goes through mc_archive_decode_chunk on an archive of the blobs
    (written by mc_archive_append_chunk_compressed), both as the API is called
    (chunks in turn, nthreads threads each) and with nthreads workers decoding
    whole chunks at nthreads=1; encode is mc_archive_append_chunk_par with
    nthreads threads at q=6.  The strict build's chunks are checked against
    the decoded hashes; the fast build is not bit-reproducible, and how many
    of its chunks match is reported."""
    import tempfile
    import threading
    if not all(os.path.exists(os.path.join(REF_DIR, f)) for f in ("libmcref.so", "libmcref_fast.so")):
        return None
    ns = len(blobs)
    gb = ns * mc_amd.CHUNK_BYTES / 1e9
    out = np.empty((ns, 256, 256, 256), np.uint8)
    tmp = tempfile.mkdtemp(prefix="mcref_")
    res = {}
    try:
        for tag, so in (("fast", "libmcref_fast.so"), ("strict", "libmcref.so")):
            R = _ref_lib(so)
            path = os.path.join(tmp, f"{tag}.mca")
            a = R.mc_archive_open_dims(path.encode(), 256 * ns, 256, 256, C.c_float(6.0))
            assert a, "reference archive open failed"
            for i, b in enumerate(blobs):
                bb = np.frombuffer(b, np.uint8)
                assert R.mc_archive_append_chunk_compressed(a, 0, 0, 0, i, bb.ctypes.data, bb.size) == 0
            offs = [R.mc_archive_chunk_offset(a, 0, 0, 0, i) for i in range(ns)]

            def api():
                for i in range(ns):
                    R.mc_archive_decode_chunk(a, offs[i], out[i].ctypes.data, nthreads)

            def workers():
                def one(i):
                    R.mc_archive_decode_chunk(a, offs[i], out[i].ctypes.data, 1)
                for s in range(0, ns, nthreads):
                    th = [threading.Thread(target=one, args=(i,)) for i in range(s, min(ns, s + nthreads))]
                    for t in th:
                        t.start()
                    for t in th:
                        t.join()

            res[tag + "_api"] = gb / _timed(api)
            res[tag + "_workers"] = gb / _timed(workers)
            res[tag + "_exact"] = sum(xxh64_hex(out[i]) == man[i]["decoded_xxh64"] for i in range(ns))
            R.mc_archive_close(a)
            epath = os.path.join(tmp, f"{tag}_enc.mca")
            e = R.mc_archive_open_dims(epath.encode(), 256 * ns, 256, 256, C.c_float(6.0))
            assert e, "reference archive open failed"
            t0 = time.perf_counter()
            for i in range(ns):
                R.mc_archive_append_chunk_par(e, 0, 0, 0, i, chunks[i].ctypes.data, C.c_float(6.0), nthreads)
            res[tag + "_enc"] = gb / (time.perf_counter() - t0)
            R.mc_archive_close(e)
    finally:
        import shutil
        shutil.rmtree(tmp, ignore_errors=True)
    return res


def dropin_leg(nthreads, nchunks=64, nappend=16):
    """The drop-in calls as a caller of the reference's API makes them, with
    host buffers (so PCIe transfers are included), in a process at HIP's
    default hardware-queue count: the plain64_q6 chunks are generated and
    GPU-encoded (blobs checked against the reference manifest) into an
    archive; then mc_archive_decode_chunk of every chunk into pageable host
    memory from nthreads caller threads (merged into batched decodes, and,
    for comparison, with merging off: each thread's own decode on its own
    stream) and from one thread, mc_archive_append_chunk_par from one thread,
    and one chunk per mchip_decode_chunks launch rotated over the streams the
    default queues allow.  Every decoded chunk is hashed against the
    manifest.  GB/s of decoded (or encoded) voxels."""
    import tempfile
    import threading
    import shutil
    L = mc_amd.api()
    specs = workloads.WORKLOADS["plain64_q6"]()[:nchunks]
    man = json.load(open(os.path.join(ROOT, "tests", "golden", "chunk_manifest.json")))["plain64_q6"]
    n = len(specs)
    chunks = mc_amd.synth_chunks([s_["seed"] for s_ in specs], [s_["air"] for s_ in specs], nthreads=nthreads)
    ctx = mc_amd.Ctx(0)
    blobs = ctx.encode_chunks_host(chunks, [6.0] * n)
    enc_ok = all(len(b) == m["blob_len"] and xxh64_hex(np.frombuffer(b, np.uint8)) == m["blob_xxh64"]
                 for b, m in zip(blobs, man))
    tmp = tempfile.mkdtemp(prefix="mcamd_")
    res = {"hw_queues": "HIP default (GPU_MAX_HW_QUEUES unset)", "encode_byte_identical": enc_ok}
    bad = []
    try:
        a = L.mc_archive_open_dims(os.path.join(tmp, "d.mca").encode(), 256 * n, 256, 256, C.c_float(6.0))
        assert a, "archive open failed"
        for i, b in enumerate(blobs):
            bb = np.frombuffer(b, np.uint8)
            assert L.mc_archive_append_chunk_compressed(a, 0, 0, 0, i, bb.ctypes.data, bb.size) == 0
        offs = [L.mc_archive_chunk_offset(a, 0, 0, 0, i) for i in range(n)]
        out = np.empty((nthreads, 256, 256, 256), np.uint8)

        def run(t, stride, check):
            for i in range(t, n, stride):
                L.mc_archive_decode_chunk(a, offs[i], out[t].ctypes.data, 0)
                if check and xxh64_hex(out[t]) != man[i]["decoded_xxh64"]:
                    bad.append(i)

        def threaded(rounds=6):
            # a pool of caller threads that live across the rounds: a thread's
            # device context is created by its first call and kept until it
            # exits; the first two rounds warm up (contexts, the runtime's
            # copy staging, first touch of the output pages), the last round
            # hashes every output (outside the timing)
            start, done = threading.Barrier(nthreads + 1), threading.Barrier(nthreads + 1)

            def worker(t):
                for r in range(rounds):
                    start.wait()
                    run(t, nthreads, r == rounds - 1)
                    done.wait()

            pool = [threading.Thread(target=worker, args=(t,)) for t in range(nthreads)]
            for x in pool:
                x.start()
            best = None
            for r in range(rounds):
                t0 = time.perf_counter()
                start.wait()
                done.wait()
                dt = time.perf_counter() - t0
                if 1 < r < rounds - 1:
                    best = dt if best is None else min(best, dt)
            for x in pool:
                x.join()
            return n * mc_amd.CHUNK_BYTES / 1e9 / best

        b0 = C.c_uint64()
        c0 = C.c_uint64()
        res[f"decode_chunk_{nthreads}_threads_gbs"] = round(threaded(), 3)
        mc_amd.check(mc_amd.load().mchip_merge_stats(C.byref(b0), C.byref(c0)))
        res["merged_chunks_per_batch"] = round(c0.value / max(1, b0.value), 2)
        mc_amd.set_tuning("merge", 0)
        res[f"decode_chunk_{nthreads}_threads_unmerged_gbs"] = round(threaded(), 3)
        mc_amd.set_tuning("merge", 1)
        run(0, 1, True)
        t0 = time.perf_counter()
        run(0, 1, False)
        res["decode_chunk_1_thread_gbs"] = round(n * mc_amd.CHUNK_BYTES / 1e9 / (time.perf_counter() - t0), 3)
        L.mc_archive_close(a)
        e = L.mc_archive_open_dims(os.path.join(tmp, "e.mca").encode(), 256 * nappend, 256, 256, C.c_float(6.0))
        assert e, "archive open failed"
        L.mc_archive_append_chunk_par(e, 0, 0, 0, 0, chunks[0].ctypes.data, C.c_float(6.0), 0)   # warm
        t0 = time.perf_counter()
        for i in range(nappend):
            L.mc_archive_append_chunk_par(e, 0, 0, 0, i, chunks[i].ctypes.data, C.c_float(6.0), 0)
        res["append_chunk_par_1_thread_gbs"] = round(nappend * mc_amd.CHUNK_BYTES / 1e9
                                                     / (time.perf_counter() - t0), 3)
        L.mc_archive_close(e)
        # one chunk per launch, device-resident, over the streams the default
        # queues allow (4 queues, one held by the null stream)
        res["per_chunk_resident_default_queues_gbs"] = round(per_chunk_rate(ctx, blobs, man, 3, 3, bad), 3)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
        ctx.close()
    res["bit_exact"] = not bad
    res["note"] = (f"host buffers, PCIe included: decode of {n} blobs from an archive into pageable memory "
                   f"(every output hashed against the reference manifest); append of {nappend} chunks")
    return res


def per_chunk_rate(ctx, blobs, man, nstreams, steps, bad):
    """GB/s of one chunk per mchip_decode_chunks launch over resident blobs,
    rotated over nstreams contexts (ctx first); outputs hashed afterwards."""
    n = len(blobs)
    offs = np.zeros(n, np.uint64)
    tot = 0
    for i, b in enumerate(blobs):
        offs[i] = tot
        tot += (len(b) + 255) // 256 * 256
    host = np.zeros(tot, np.uint8)
    for i, b in enumerate(blobs):
        host[int(offs[i]):int(offs[i]) + len(b)] = np.frombuffer(b, np.uint8)
    d_b = mc_amd.DeviceBuffer(tot)
    d_o = mc_amd.DeviceBuffer(n * mc_amd.CHUNK_BYTES)
    ctx.h2d(d_b, host)
    ctxs = [ctx] + [mc_amd.Ctx(0) for _ in range(nstreams - 1)]
    Lb = mc_amd.load()
    lens = [np.array([len(b)], np.uint64) for b in blobs]
    qv = [np.array([np.frombuffer(b[:4], np.float32)[0]], np.float32) for b in blobs]
    ov = [np.array([offs[i]], np.uint64) for i in range(n)]

    def step(k):
        for i in range(n):
            mc_amd.check(Lb.mchip_decode_chunks(ctxs[k].h, C.c_void_p(d_b.ptr), ov[i].ctypes.data_as(C.c_void_p),
                                                lens[i].ctypes.data_as(C.c_void_p),
                                                qv[i].ctypes.data_as(C.c_void_p), 1,
                                                C.c_void_p(d_o.ptr + i * mc_amd.CHUNK_BYTES), mc_amd.CHUNK_BYTES))
            k = k + 1 if k + 1 < len(ctxs) else 0
        return k

    k = step(0)
    for c_ in ctxs:
        c_.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        k = step(k)
    for c_ in ctxs:
        c_.sync()
    dt = (time.perf_counter() - t0) / steps
    out = np.empty((256, 256, 256), np.uint8)
    for i in range(n):
        ctx.d2h(out, d_o, offset=i * mc_amd.CHUNK_BYTES)
        if xxh64_hex(out) != man[i]["decoded_xxh64"]:
            bad.append(("per_chunk", i))
    for c_ in ctxs[1:]:
        c_.close()
    d_b.free()
    d_o.free()
    return n * mc_amd.CHUNK_BYTES / 1e9 / dt


def host_cpu():
    """The host's CPU model and core counts (as /proc/cpuinfo and the
    scheduler report them)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except Exception:
        usable = None
    return {"model": model, "nproc": os.cpu_count(), "usable": usable}


def cpu_record(ref, port_dec, port_enc, ns, n, nthreads):
    """The cpu_baseline object: the reference's shipped build when it was
    timed (its faster decode policy is the value), else the oracle port."""
    port = {"port_decode_gbs": round(port_dec, 3), "port_encode_gbs": round(port_enc, 3),
            "port_sample": f"oracle decode of {ns} of the {n} chunks, {nthreads} threads over blocks",
            "host": host_cpu()}
    if not ref:
        return {"value": round(port_dec, 3), "unit": "GB/s", "cores": nthreads, "kind": "port",
                "sample": port["port_sample"], "encode_gbs": round(port_enc, 3), "host": port["host"]}
    r3 = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in ref.items()}
    return {"value": round(max(ref["fast_api"], ref["fast_workers"]), 3), "unit": "GB/s", "cores": nthreads,
            "kind": "reference",
            "build": "reference sources with their shipped flags (-O3 -ffast-math -ffp-contract=fast, "
                     "-march=x86-64-v3); strict build -O2 -fno-fast-math -ffp-contract=off",
            "sample": f"mc_archive_decode_chunk over {ns} of the {n} chunks: in turn with nthreads={nthreads} "
                      f"(decode_api_gbs) and {nthreads} workers x nthreads=1 (decode_workers_gbs); "
                      f"mc_archive_append_chunk_par of the same chunks at q=6, nthreads={nthreads}",
            "decode_api_gbs": r3["fast_api"], "decode_workers_gbs": r3["fast_workers"],
            "encode_gbs": r3["fast_enc"],
            "fast_build_bit_exact_chunks": f"{ref['fast_exact']}/{ns}",
            "strict": {"decode_api_gbs": r3["strict_api"], "decode_workers_gbs": r3["strict_workers"],
                       "encode_gbs": r3["strict_enc"], "bit_exact_chunks": f"{ref['strict_exact']}/{ns}"},
            **port}


def cpu_verify_baseline(blobs):
    """The oracle's blob hash (the reference's mc_chunk_compute_hash restated)
    over the blobs on one core, repeated to about half a second: GB/s."""
    import oracle_bind as ob
    O = ob.oracle()
    tot, reps = sum(b.size for b in blobs), 0
    t0 = time.perf_counter()
    while True:
        for b in blobs:
            O.orc_blob_hash(ob.p(b), b.size)
        reps += 1
        dt = time.perf_counter() - t0
        if dt > 0.5:
            return reps * tot / dt / 1e9


def mixed_leg(nthreads):
    """Mixed-quality batch (mixed256: 256 chunks, 4 GiB, presets with max-error
    corrections): batched GPU encode checked byte for byte against the
    reference's blobs, then the batched decode autotuned on the batch, timed
    eagerly and as HIP graph replays, checked against the reference's decoded
    hashes."""
    specs = workloads.WORKLOADS["mixed256"]()
    man = json.load(open(os.path.join(ROOT, "tests", "golden", "chunk_manifest.json")))["mixed256"]
    n = len(specs)
    chunks = mc_amd.synth_chunks([x["seed"] for x in specs], [x["air"] for x in specs], nthreads=nthreads)
    ctx = mc_amd.Ctx(0)
    d_in = mc_amd.DeviceBuffer(n * mc_amd.CHUNK_BYTES)
    ctx.h2d(d_in, chunks)
    del chunks
    cap = max(m["blob_len"] for m in man) + (1 << 20)
    d_enc = mc_amd.DeviceBuffer(n * cap)
    eoffs = np.arange(n, dtype=np.uint64) * cap
    ecaps = np.full(n, cap, np.uint64)
    qv = [x["q"] for x in specs]
    tv = [x["tau"] for x in specs]
    # the encode workspace of a 256-chunk sequence (about 165 MiB per chunk)
    # stays with the context between the timed calls
    mc_amd.set_tuning("enc_keep_mib", 1 << 20)
    ctx.encode_chunks(d_in, qv, tv, d_enc, eoffs, ecaps)          # workspace allocation
    ctx.sync()
    # the encode shapes autotuned on the batch (cached for its signature and
    # used by the timed call)
    enc_tune = ctx.autotune_encode(d_in, qv, tv, d_enc, eoffs, ecaps)
    log(f"[bench] mixed encode autotune: {enc_tune}")
    t0 = time.perf_counter()
    elens, est = ctx.encode_chunks(d_in, qv, tv, d_enc, eoffs, ecaps)
    enc_s = time.perf_counter() - t0

    def check_blobs(lens, st):
        ok = bool((st == 0).all())
        for i in range(n):
            b = np.empty(int(lens[i]), np.uint8)
            ctx.d2h(b, d_enc, offset=int(eoffs[i]))
            ok &= len(b) == man[i]["blob_len"] and xxh64_hex(b) == man[i]["blob_xxh64"]
        return ok

    ok_enc = check_blobs(elens, est)
    log(f"[bench] mixed encode {n * mc_amd.CHUNK_BYTES / enc_s / 1e9:.2f} GB/s, byte-identical={ok_enc}")
    # the same sequence captured as a HIP graph and replayed
    enc_graph_s, ok_graph_enc = None, None
    try:
        eg, egn = ctx.encode_graph(d_in, qv, tv, d_enc, eoffs, ecaps)
        ctx.graph_launch(eg)
        ctx.sync()
        tge = []
        for _ in range(2):
            t0 = time.perf_counter()
            ctx.graph_launch(eg)
            ctx.sync()
            tge.append(time.perf_counter() - t0)
        glens, gst = ctx.encode_graph_results(eg, egn)
        ok_graph_enc = check_blobs(glens, gst)
        enc_graph_s = min(tge)
        ctx.graph_destroy(eg)
        log(f"[bench] mixed encode graph {n * mc_amd.CHUNK_BYTES / enc_graph_s / 1e9:.2f} GB/s, "
            f"byte-identical={ok_graph_enc}")
    except mc_amd.MchipError as e:
        log(f"[bench] mixed encode graph not captured: {e}")
    mc_amd.set_tuning("enc_keep_mib", 2048)
    d_in.free()
    qs = np.array([x["q"] for x in specs], np.float32)
    d_out = mc_amd.DeviceBuffer(n * mc_amd.CHUNK_BYTES)

    def check():
        out = np.empty(mc_amd.CHUNK_BYTES, np.uint8)
        bad = 0
        for i in range(n):
            ctx.d2h(out, d_out, offset=i * mc_amd.CHUNK_BYTES)
            bad += xxh64_hex(out) != man[i]["decoded_xxh64"]
        return bad

    tune = ctx.autotune_decode(d_enc, eoffs, elens, qs, d_out)
    best = next((r for r in tune if r.get("chosen")), min(tune, key=lambda r: r["ms"]))
    log(f"[bench] mixed autotune: {tune}")
    ts = []
    for _ in range(4):
        t0 = time.perf_counter()
        ctx.decode_chunks(d_enc, eoffs, elens, qs, d_out)
        ctx.sync()
        ts.append(time.perf_counter() - t0)
    bad_eager = check()
    g = ctx.decode_graph(d_enc, eoffs, elens, qs, d_out)
    ctx.graph_launch(g)                  # the first replay uploads the graph
    ctx.sync()
    tg = []
    for _ in range(4):
        t0 = time.perf_counter()
        ctx.graph_launch(g)
        ctx.sync()
        tg.append(time.perf_counter() - t0)
    bad_graph = check()
    ctx.graph_destroy(g)
    d_enc.free()
    d_out.free()
    res = {"workload": "mixed256", "chunks": n, "decoded_bytes": n * mc_amd.CHUNK_BYTES,
           "blob_bytes": int(np.sum(elens)),
           "encode_gbs": round(n * mc_amd.CHUNK_BYTES / enc_s / 1e9, 3), "encode_byte_identical": ok_enc,
           "encode_graph_gbs": (round(n * mc_amd.CHUNK_BYTES / enc_graph_s / 1e9, 3) if enc_graph_s else None),
           "encode_graph_byte_identical": ok_graph_enc, "encode_autotune": enc_tune,
           "encode_chosen": next((r for r in enc_tune if r.get("chosen")), min(enc_tune, key=lambda r: r["ms"])),
           "decode_gbs": round(n * mc_amd.CHUNK_BYTES / min(ts) / 1e9, 3),
           "decode_graph_gbs": round(n * mc_amd.CHUNK_BYTES / min(tg) / 1e9, 3),
           "bit_exact": bad_eager == 0 and bad_graph == 0, "autotune": tune, "chosen": best,
           "launch": "all chunks in one launch sequence (persistent entropy grid), and that sequence as a HIP graph"}
    log(f"[bench] mixed decode {res['decode_gbs']} GB/s, graph {res['decode_graph_gbs']} GB/s, "
        f"bit-exact={res['bit_exact']}")
    return res


def build_pyramid(L, ctx, path, nthreads):
    """The 2048^3 volume (volume2048_q6, LOD 0) and its decimations LOD 1-3
    (2x2x2 chunk groups decimated on the device, src/mc_archive.c:61-72), each
    chunk GPU-encoded at q=6 and installed in one archive.  Returns the
    archive and, per level, {(cz, cy, cx): blob}."""
    CH = mc_amd.CHUNK_BYTES
    specs = workloads.WORKLOADS["volume2048_q6"]()
    by = {x["coord"]: x for x in specs}
    a = L.mc_archive_open_dims(path.encode(), 2048, 2048, 2048, C.c_float(6.0))
    if not a:
        raise RuntimeError("mc_archive_open_dims failed")
    blobs = [dict() for _ in range(4)]

    def install(lod, coord, bl):
        if bl is None or len(bl) == 0:
            return
        if L.mc_archive_append_chunk_compressed(a, lod, *coord, bl, len(bl)) != 0:
            raise RuntimeError("install failed")
        blobs[lod][coord] = bl

    cube = [(z, y, x) for z in range(2) for y in range(2) for x in range(2)]
    d_grp = mc_amd.DeviceBuffer(8 * CH)
    levels = [None, mc_amd.DeviceBuffer(64 * CH), mc_amd.DeviceBuffer(8 * CH), mc_amd.DeviceBuffer(CH)]
    # LOD 0: 64 groups of 2x2x2 chunks, synthesized on the host, encoded, and
    # decimated into LOD 1 on the device
    for gz in range(4):
        for gy in range(4):
            for gx in range(4):
                grp = [by[(2 * gz + z, 2 * gy + y, 2 * gx + x)] for z, y, x in cube]
                ch = mc_amd.synth_chunks([x["seed"] for x in grp], [x["air"] for x in grp],
                                         coords=[x["coord"] for x in grp], nthreads=nthreads)
                for x, bl in zip(grp, ctx.encode_chunks_host(ch, [6.0] * 8)):
                    install(0, x["coord"], bl)
                ctx.h2d(d_grp, ch)
                ctx.decimate_chunks([(d_grp, k * CH) for k in range(8)], levels[1],
                                    dst_off=((gz * 4 + gy) * 4 + gx) * CH)
                del ch
    # LOD 1..3: encode the level on the device, decimate it into the next
    d_enc = mc_amd.DeviceBuffer(64 * (20 << 20))
    for lod, n in ((1, 4), (2, 2), (3, 1)):
        cnt = n ** 3
        offs = np.arange(cnt, dtype=np.uint64) * (20 << 20)
        lens, st = ctx.encode_chunks(levels[lod], [6.0] * cnt, [0] * cnt, d_enc, offs, np.full(cnt, 20 << 20))
        assert (st == 0).all(), f"LOD {lod} encode status {st}"
        for i in range(cnt):
            coord = (i // (n * n), (i // n) % n, i % n)
            if lens[i]:
                b = np.empty(int(lens[i]), np.uint8)
                ctx.d2h(b, d_enc, offset=int(offs[i]))
                install(lod, coord, b.tobytes())
        if lod < 3:
            m = n // 2
            for hz in range(m):
                for hy in range(m):
                    for hx in range(m):
                        src = [(levels[lod], (((2 * hz + z) * n + 2 * hy + y) * n + 2 * hx + x) * CH)
                               for z, y, x in cube]
                        ctx.decimate_chunks(src, levels[lod + 1], dst_off=((hz * m + hy) * m + hx) * CH)
    ctx.sync()
    for b in [d_grp, d_enc] + levels[1:]:
        b.free()
    return a, blobs


def volume_manifest():
    with open(os.path.join(ROOT, "tests", "golden", "volume_manifest.json")) as f:
        m = json.load(f)["volume2048_q6"]
    return {int(k): {tuple(r["coord"]): r for r in v} for k, v in m.items()}


def check_pyramid_blobs(blobs, man):
    """Chunks whose GPU-encoded blob differs from the reference's (or that the
    reference has and the pyramid lacks, or the other way round)."""
    bad = 0
    for lod, lb in enumerate(blobs):
        for c, r in man[lod].items():
            b = lb.get(c)
            if r["blob_len"] == 0:
                bad += b is not None and len(b) > 0
            elif b is None or len(b) != r["blob_len"] or xxh64_hex(np.frombuffer(b, np.uint8)) != r["blob_xxh64"]:
                bad += 1
    return bad


def dense_levels(ctx, blobs, man=None):
    """Decoded levels (GPU decode of every installed chunk) as dense rasters,
    and the number of chunks whose decode differs from the manifest's."""
    out = []
    bad = 0
    for lod, lb in enumerate(blobs):
        n = 2048 >> lod
        v = np.zeros((n, n, n), np.uint8)
        items = sorted(lb.items())
        for i in range(0, len(items), 32):
            part = items[i:i + 32]
            dec = ctx.decode_chunks_host([b for _, b in part])
            for (cz, cy, cx), d in zip([c for c, _ in part], dec):
                if man is not None and xxh64_hex(d) != man[lod][(cz, cy, cx)]["decoded_xxh64"]:
                    bad += 1
                v[256 * cz:256 * cz + 256, 256 * cy:256 * cy + 256, 256 * cx:256 * cx + 256] = d
        out.append(v)
    return out, bad


def render_frame(k, w=1024, h=1024):
    """Render request k: a plane through a point of the volume interior with
    a normal uniform on the sphere (seed 5000 + k), as a 129 x 129 control
    grid of 64-voxel spacing; zoom 1, 2, 4, 8 voxels per pixel in turn (the
    step makes mc_render_pick_lod choose LOD 0, 1, 2, 3); the 1024^2 window
    is centred on the point."""
    rng = np.random.default_rng(5000 + k)
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    u = np.cross(n, rng.normal(size=3))
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    o = rng.uniform(512, 1536, 3)
    g = 129
    sp = 64.0
    j = np.arange(g, dtype=np.float64) - (g - 1) / 2
    grid = (o[None, None, :] + sp * (j[:, None, None] * v[None, None, :] + j[None, :, None] * u[None, None, :]))
    grid = grid.astype(np.float32)
    zoom = [1, 2, 4, 8][k % 4]
    step = zoom * 1.001 / sp
    x0 = (g - 1) / 2 - w * step / 2
    return grid, float(x0), float(x0), float(step), w, h, zoom


def render_leg(a, levels, ctx, nthreads, arena_bytes=4 << 30, lanes=8, boxes_per_batch=64, edge=128,
               n_render=40, n_batch=80, lam_render=10.0, lam_batch=20.0):
    """Mixed render + box-batch serving over LODs 0-3 of the 2048^3 volume with
    a 4 GiB decoded-block arena: 1024^2 quad renders (trilinear, 9-step mean
    along the normal, t0 = -4, t1 = 4) and batches of 64 boxes of 128^3
    arrive as independent Poisson processes; renders are checked bit for bit
    against the oracle's renderer fed the same decoded levels."""
    import threading
    import oracle_bind as ob
    L = mc_amd.api()
    sv = mc_amd.Serve(a, 0, arena_bytes=arena_bytes, lanes=lanes, nlods=4)
    fb = [mc_amd.DeviceBuffer(1024 * 1024) for _ in range(2 * lanes)]
    bb = [mc_amd.DeviceBuffer(boxes_per_batch * edge ** 3) for _ in range(2 * lanes)]

    def one(k, buf):
        grid, x0, y0, step, w, h, _ = render_frame(k)
        t = sv.render(grid, x0, y0, step, w, h, buf)
        lod = sv.render_level(t)
        return lod, sv.wait(t)

    # cold (blocks decoded by the request) and resident frames, one at a time
    cold, warm, lods, frames = [], [], [], []
    for k in range(4):
        lod, ms = one(k, fb[k])
        cold.append(round(ms, 3))
        lods.append(lod)
        img = np.empty((1024, 1024), np.uint8)
        ctx.d2h(img, fb[k])
        frames.append(img)
    for k in range(4):
        warm.append(round(one(k, fb[k])[1], 3))
    log(f"[bench] render frames LOD {lods}: cold {cold} ms, resident {warm} ms")
    # parity and the CPU baseline: the oracle renders the same frames from
    # the GPU-decoded levels (one core)
    dense = levels
    exact, cpu_s = [], []
    for k in range(4):
        grid, x0, y0, step, w, h, _ = render_frame(k)
        t0 = time.perf_counter()
        want, wl = ob.render_quad_lod(dense, grid, x0, y0, step, w, h, 1, 2, -4.0, 4.0, 1.0)
        cpu_s.append(time.perf_counter() - t0)
        exact.append(bool(wl == lods[k] and np.array_equal(want, frames[k])))
        lit = float((want > 0).mean())
        log(f"[bench] render frame {k}: bit-exact {exact[-1]}, lit {lit:.2f}, oracle {cpu_s[-1]:.2f}s")
    del dense

    # mixed Poisson phases
    def batch(k):
        arr = (mc_amd.McBox * boxes_per_batch)()
        got = L.mc_archive_sample_boxes(a, 0, C.c_uint64(30000 + k), boxes_per_batch, edge, edge, edge,
                                        C.c_float(0.5), arr)
        return [(arr[i].z0, arr[i].y0, arr[i].x0) for i in range(got)]

    def mixed(n_render, n_batch, lam_render, lam_batch, seed, kbase):
        breqs = [batch(kbase + k) for k in range(n_batch)]
        rng = np.random.default_rng(seed)
        ev = [(t, "r", i) for i, t in enumerate(np.cumsum(rng.exponential(1.0 / lam_render, n_render)))]
        ev += [(t, "b", i) for i, t in enumerate(np.cumsum(rng.exponential(1.0 / lam_batch, n_batch)))]
        ev.sort()
        lat = {"r": [None] * n_render, "b": [None] * n_batch}
        tick = []
        ready = threading.Semaphore(0)

        def waiter():
            for j in range(len(ev)):
                ready.acquire()
                cls, i, t = tick[j]
                lat[cls][i] = sv.wait(t)

        th = threading.Thread(target=waiter)
        th.start()
        s0 = sv.stats()
        start = time.perf_counter()
        used = {"r": 0, "b": 0}
        for tdue, cls, i in ev:
            d = start + tdue - time.perf_counter()
            if d > 0:
                time.sleep(d)
            ring = fb if cls == "r" else bb
            k = used[cls]
            if k >= len(ring):            # output buffer reuse: its earlier request must be done
                while lat[cls][k - len(ring)] is None:
                    time.sleep(0.0002)
            if cls == "r":
                grid, x0, y0, step, w, h, _ = render_frame(kbase + i)
                t = sv.render(grid, x0, y0, step, w, h, ring[k % len(ring)])
            else:
                t = sv.submit(breqs[i], edge, edge, edge, ring[k % len(ring)])
            used[cls] += 1
            tick.append((cls, i, t))
            ready.release()
        th.join()
        wall = time.perf_counter() - start
        s1 = sv.stats()
        lr, lb = np.array(lat["r"]), np.array(lat["b"])
        refs = s1["block_refs"] - s0["block_refs"]
        m = {"renders": n_render, "render_lambda_per_s": lam_render, "batches": n_batch,
             "batch_lambda_per_s": lam_batch, "wall_s": round(wall, 3),
             "render_p50_ms": round(float(np.percentile(lr, 50)), 3),
             "render_p99_ms": round(float(np.percentile(lr, 99)), 3),
             "batch_p50_ms": round(float(np.percentile(lb, 50)), 3),
             "batch_p99_ms": round(float(np.percentile(lb, 99)), 3),
             "frames_per_s": round(n_render / wall, 3),
             "batch_gbs": round(sum(len(r) for r in breqs) * edge ** 3 / wall / 1e9, 3),
             "hit_rate": round((s1["hits"] - s0["hits"]) / max(1, refs), 4),
             "decoded_blocks": s1["misses"] - s0["misses"],
             "evictions": s1["evictions"] - s0["evictions"]}
        log(f"[bench] render+serve lambda {lam_render}/{lam_batch} per s: render p50 {m['render_p50_ms']} "
            f"p99 {m['render_p99_ms']} ms, batch p50 {m['batch_p50_ms']} p99 {m['batch_p99_ms']} ms, "
            f"{m['frames_per_s']} frames/s, {m['batch_gbs']} GB/s, hit rate {m['hit_rate']}")
        return m

    phases = [mixed(n_render, n_batch, lam_render, lam_batch, 77, 100),
              mixed(4 * n_render, 2 * n_batch, 20 * lam_render, 5 * lam_batch, 78, 1000)]
    sv.close()
    for b in fb + bb:
        b.free()
    return {"volume": "2048^3 u8 + LOD 1-3 (device decimation), q=6", "arena_bytes": arena_bytes,
            "lanes": lanes, "frame": "1024^2 quad, trilinear, mean of 9 steps (t0=-4, t1=4, dt=1)",
            "frame_lods": lods, "frame_cold_ms": cold, "frame_resident_ms": warm,
            "bit_exact_vs_oracle_renderer": all(exact),
            "cpu_oracle_frame_s": [round(x, 3) for x in cpu_s],
            "mixed": phases}


def _ref_region_lib(name):
    """The reference build (oracle/_ref) with its region and block entry
    points declared, or None when that build is absent."""
    path = os.path.join(REF_DIR, name)
    if not os.path.exists(path):
        return None
    vp = C.c_void_p
    R = _ref_lib(name)
    R.mc_archive_read_regions.argtypes = [vp, C.c_int, vp, C.c_int, C.c_long, C.c_long, C.c_long, vp, C.c_size_t,
                                          C.c_int]
    R.mc_archive_decode_block.argtypes = [vp, C.c_uint64, C.c_int, C.c_int, C.c_int, vp]
    return R


def _block_latency(L, a, picks, dst, reps=1):
    """Per-call latency (ms) of mc_archive_decode_block over picks."""
    lat = []
    for _ in range(reps):
        for co, bz, by, bx in picks:
            t0 = time.perf_counter()
            L.mc_archive_decode_block(a, co, bz, by, bx, dst.ctypes.data)
            lat.append((time.perf_counter() - t0) * 1e3)
    lat = np.array(lat)
    return {"p50_ms": round(float(np.percentile(lat, 50)), 4), "p99_ms": round(float(np.percentile(lat, 99)), 4),
            "calls": int(lat.size)}


def serve_leg(rates, nthreads, lanes=8, boxes_per_batch=64, edge=128, bounded_bytes=1 << 30, resident_batches=1000,
              with_render=True):
    """Serving of box batches from an HBM-resident 2048^3 archive
    (volume2048_q6): batches of 64 boxes of 128^3 (mc_archive_sample_boxes,
    seed 10000 + k, min material fraction 0.5) arrive as a Poisson process
    (exponential gaps from seed 99) and are served by mc_serve_* into device
    memory; latency is submission to completion.  Each rate runs on a fresh
    server whose arena (half of the free device memory) starts empty, so the
    blocks are decoded by the requests ("cold"); a run through a 1 GiB arena
    (a quarter of the blocks those runs touch) keeps missing ("bounded"); a run after a preload of every block serves
    decoded blocks only ("resident").  The pyramid's GPU-encoded blobs and
    its GPU decode are checked against the reference manifest, and served and
    read_regions boxes against that decode.  mc_archive_read_regions into
    host buffers and per-call mc_archive_decode_block are timed beside the
    reference build's own (same archive file, host cores)."""
    import tempfile
    import threading
    L = mc_amd.api()
    path = os.path.join(tempfile.mkdtemp(prefix="mcserve_"), "vol2048.mca")
    ctx = mc_amd.Ctx(0)
    t0 = time.perf_counter()
    a, lv_blobs = build_pyramid(L, ctx, path, nthreads)
    log(f"[bench] serving volume (LOD 0-3) built in {time.perf_counter() - t0:.1f}s")
    vman = volume_manifest()
    bad_blobs = check_pyramid_blobs(lv_blobs, vman)
    t0 = time.perf_counter()
    levels, bad_dec = dense_levels(ctx, lv_blobs, vman)
    log(f"[bench] pyramid blobs differing from the reference: {bad_blobs}; decoded chunks differing: {bad_dec} "
        f"({time.perf_counter() - t0:.1f}s)")
    vol = levels[0]

    def batch(k):
        arr = (mc_amd.McBox * boxes_per_batch)()
        got = L.mc_archive_sample_boxes(a, 0, C.c_uint64(10000 + k), boxes_per_batch, edge, edge, edge,
                                        C.c_float(0.5), arr)
        return [(arr[i].z0, arr[i].y0, arr[i].x0) for i in range(got)]

    def matches(got, boxes):
        return all(np.array_equal(got[j], vol[z:z + edge, y:y + edge, x:x + edge]) for j, (z, y, x) in enumerate(boxes))

    box_bytes = edge ** 3
    nbuf = 3 * lanes
    bufs = [mc_amd.DeviceBuffer(boxes_per_batch * box_bytes) for _ in range(nbuf)]
    # bit-exactness: two batches served cold and through read_regions (host
    # output), against the manifest-checked decode
    exact = True
    rr_ms = []
    sv = mc_amd.Serve(a, lod=0, arena_bytes=0, lanes=lanes)
    for k in range(2):
        bk = batch(k)
        sv.wait(sv.submit(bk, edge, edge, edge, bufs[0]))
        got = np.empty((len(bk), edge, edge, edge), np.uint8)
        ctx.d2h(got, bufs[0])
        exact &= matches(got, bk)
        arr = (mc_amd.McBox * len(bk))(*[mc_amd.McBox(*o) for o in bk])
        for rep in range(2):
            t0 = time.perf_counter()
            L.mc_archive_read_regions(a, 0, arr, len(bk), edge, edge, edge, got.ctypes.data_as(C.c_void_p),
                                      box_bytes, 0)
            rr_ms.append((time.perf_counter() - t0) * 1e3)
        exact &= matches(got, bk)
        del got
    sv.close()
    rr_batch = batch(0)
    out = []
    kbase = 2
    runs = [("cold", lam, nb, 0) for lam, nb in rates] + [("bounded", 100, rates[-1][1], bounded_bytes),
                                                          ("resident", 100, resident_batches, 0)]
    preload = None
    slots = {}
    for phase, lam, nb, arena in runs:
        if not nb:
            continue
        sv = mc_amd.Serve(a, lod=0, arena_bytes=arena, lanes=lanes)
        slots[phase] = sv.stats()["arena_slots"]
        if phase == "resident":
            tp = time.perf_counter()
            nl = sv.preload()
            preload = {"blocks": nl, "ms": round(1e3 * (time.perf_counter() - tp), 1)}
            log(f"[bench] serving preload: {nl} blocks in {preload['ms']} ms")
        reqs = [batch(kbase + k) for k in range(nb)]
        kbase += nb
        gaps = np.random.default_rng(99).exponential(1.0 / lam, nb)
        s0 = sv.stats()
        lat = [None] * nb
        tickets = [None] * nb
        ready = [threading.Event() for _ in range(nb)]

        def waiter():
            for i in range(nb):
                ready[i].wait()
                lat[i] = sv.wait(tickets[i])

        th = threading.Thread(target=waiter)
        th.start()
        start = time.perf_counter()
        due = start
        for i in range(nb):
            due += gaps[i]
            d = due - time.perf_counter()
            if d > 0:
                time.sleep(d)
            if i >= nbuf:                 # output buffer reuse: its earlier request must be done
                while lat[i - nbuf] is None:
                    time.sleep(0.0002)
            tickets[i] = sv.submit(reqs[i], edge, edge, edge, bufs[i % nbuf])
            ready[i].set()
        th.join()
        wall = time.perf_counter() - start
        s1 = sv.stats()
        # the last batch of the run, checked against the reference decode
        last = np.empty((len(reqs[-1]), edge, edge, edge), np.uint8)
        ctx.d2h(last, bufs[(nb - 1) % nbuf])
        exact &= matches(last, reqs[-1])
        del last
        sv.close()
        lv = np.array(lat)
        vox = sum(len(r) for r in reqs) * box_bytes
        refs = s1["block_refs"] - s0["block_refs"]
        out.append({"phase": phase, "lambda_per_s": lam, "batches": nb, "arena_slots": slots[phase],
                    "p50_ms": round(float(np.percentile(lv, 50)), 3),
                    "p99_ms": round(float(np.percentile(lv, 99)), 3), "mean_ms": round(float(lv.mean()), 3),
                    "gbs": round(vox / wall / 1e9, 3),
                    "hit_rate": round((s1["hits"] - s0["hits"]) / max(1, refs), 4),
                    "decoded_blocks": s1["misses"] - s0["misses"], "evictions": s1["evictions"] - s0["evictions"],
                    "deferred": s1["deferred"] - s0["deferred"]})
        log(f"[bench] serving {phase} lambda={lam}/s: p50 {out[-1]['p50_ms']} ms p99 {out[-1]['p99_ms']} ms "
            f"{out[-1]['gbs']} GB/s hit rate {out[-1]['hit_rate']} decoded {out[-1]['decoded_blocks']} "
            f"evictions {out[-1]['evictions']}")
    for b_ in bufs:
        b_.free()
    # per-call block decode latency (the path a block cache's miss takes)
    rng = np.random.default_rng(5)
    picks = []
    coords = sorted(lv_blobs[0])
    while len(picks) < 2000:
        c = coords[int(rng.integers(0, len(coords)))]
        co = L.mc_archive_chunk_offset(a, 0, *c)
        bz, by, bx = (int(v) for v in rng.integers(0, 16, 3))
        if L.mc_archive_block_present(a, 0, 16 * c[0] + bz, 16 * c[1] + by, 16 * c[2] + bx):
            picks.append((co, bz, by, bx))
    blk = np.empty(4096, np.uint8)
    _block_latency(L, a, picks[:100], blk)
    block_lat = _block_latency(L, a, picks, blk)
    render = render_leg(a, levels, ctx, nthreads, lanes=lanes) if with_render else None
    del levels
    L.mc_archive_close(a)
    # the reference's own region read and block decode on the same archive
    # file (its format is the reference's), on the host cores
    ref = {}
    arr = (mc_amd.McBox * len(rr_batch))(*[mc_amd.McBox(*o) for o in rr_batch])
    got = np.empty((len(rr_batch), edge, edge, edge), np.uint8)
    for tag, so in (("fast", "libmcref_fast.so"), ("strict", "libmcref.so")):
        R = _ref_region_lib(so)
        if R is None:
            continue
        ra = R.mc_archive_open_dims(path.encode(), 2048, 2048, 2048, C.c_float(6.0))
        if not ra:
            continue
        t0 = time.perf_counter()
        R.mc_archive_read_regions(ra, 0, arr, len(rr_batch), edge, edge, edge, got.ctypes.data_as(C.c_void_p),
                                  box_bytes, nthreads)
        ref[tag + "_read_regions_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        if tag == "strict":
            ref["strict_read_regions_equal_ours"] = bool(matches(got, rr_batch))
        ref[tag + "_decode_block"] = _block_latency(R, ra, picks[:1000], blk)
        R.mc_archive_close(ra)
    try:
        os.remove(path)
    except OSError:
        pass
    rr = np.array(rr_ms)
    gb_batch = len(rr_batch) * box_bytes / 1e9
    return {"workload": "volume2048_q6", "volume": "2048^3 u8 (8^3 chunks), q=6, air 0.30-0.95 smooth field",
            "boxes_per_batch": boxes_per_batch, "box": f"{edge}^3", "arrivals": "Poisson, gaps from seed 99",
            "lanes": lanes, "pyramid_blobs_differing_from_reference": bad_blobs,
            "pyramid_decoded_chunks_differing_from_reference": bad_dec,
            "bit_exact_vs_reference_decode": bool(exact) and bad_dec == 0,
            "rates": out, "preload": preload,
            "latency": "submission to completion, output in device memory",
            "read_regions_host_buffers": {"ms_per_batch": round(float(np.median(rr)), 2),
                                          "gbs": round(gb_batch / (float(np.median(rr)) / 1e3), 3),
                                          "note": "mc_archive_read_regions, 64 boxes of 128^3 into pageable memory"},
            "decode_block_per_call": block_lat,
            "reference_cpu": dict(ref, cores=nthreads,
                                  note="the reference build on the same archive file: read_regions of one "
                                       "64-box batch with nthreads=cores; decode_block per call"),
            "render_serving": render}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="plain64_q6")
    ap.add_argument("--no-batched", action="store_true", help="skip the all-chunks-in-one-launch measurement")
    ap.add_argument("--batched", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--streams", type=int, default=QUEUES - 1, help="HIP streams the per-chunk launches rotate over")
    ap.add_argument("--tune", action="append", default=[], help="key=value launch tuning (mchip_set_tuning)")
    ap.add_argument("--timed-streams", type=int, default=1,
                    help="streams whose kernels carry HIP timing events during the timed steps (each "
                         "timed launch records two events; on 5 of 21 streams they cost about 7 %% of the rate)")
    ap.add_argument("--serve-batches", default="1000,1000,1000",
                    help="batches served at 20, 50 and 100 batches/s (serving leg); 0 skips the leg")
    ap.add_argument("--resident-batches", type=int, default=1000, help=argparse.SUPPRESS)
    ap.add_argument("--serve-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--mixed-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--dropin-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-quality batched leg")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the drop-in calls with host buffers (their caller threads stop rocprofv3's tracer)")
    args = ap.parse_args()
    if args.serve_only:
        nbs = [int(x) for x in args.serve_batches.split(",")]
        print(json.dumps(serve_leg(list(zip((20, 50, 100), nbs)), cpu_threads(), lanes=SERVE_LANES,
                                   resident_batches=args.resident_batches)))
        return
    if args.mixed_only:
        print(json.dumps(mixed_leg(cpu_threads())))
        return
    if args.dropin_only:
        print(json.dumps(dropin_leg(cpu_threads())))
        return
    # The legs on other workloads run first, each in a child process, before
    # this process touches the GPU: each then has the device's hardware queue
    # slots to itself.
    import subprocess

    def child(flag, *extra_args):
        r = subprocess.run([sys.executable, "-u", os.path.abspath(__file__), flag, *extra_args],
                           stdout=subprocess.PIPE, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"bench leg {flag} failed with status {r.returncode}")
        return json.loads(r.stdout.strip().splitlines()[-1])

    mixed = serving = dropin = None
    if not args.no_dropin and args.gpus == 1:
        log("[bench] drop-in calls with host buffers at the default queue count (child process)")
        dropin = child("--dropin-only")
        log(f"[bench] drop-in: {dropin}")
    if not args.no_mixed and args.gpus == 1:
        log("[bench] mixed-quality batched leg (child process)")
        mixed = child("--mixed-only")
    nbs = [int(x) for x in args.serve_batches.split(",")]
    if any(nbs) and args.gpus == 1:
        log("[bench] serving leg (child process)")
        serving = child("--serve-only", "--serve-batches", args.serve_batches, "--resident-batches",
                        str(args.resident_batches))
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested); "
              "the codec shards by chunk with no cross-GPU exchange", file=sys.stderr)
        sys.exit(2)

    enc_batch = 0                               # chunks per encode launch sequence (0: the library sizes it)
    for kv in args.tune:
        k, v = kv.split("=")
        mc_amd.set_tuning(k, int(v))
        if k == "enc_batch":
            enc_batch = int(v)
    specs = workloads.WORKLOADS[args.workload]()
    man = json.load(open(os.path.join(ROOT, "tests", "golden", "chunk_manifest.json")))[args.workload]
    n = len(specs)
    nthreads = cpu_threads()
    log(f"[bench] generating {n} chunks ({nthreads} threads)")
    t0 = time.perf_counter()
    chunks = mc_amd.synth_chunks([s["seed"] for s in specs], [s["air"] for s in specs], nthreads=nthreads)
    log(f"[bench] generated in {time.perf_counter() - t0:.1f}s")
    for i in range(min(n, 4)):
        assert xxh64_hex(chunks[i]) == man[i]["input_xxh64"], f"generator drift on chunk {i}"

    ctx = mc_amd.Ctx(0)
    # ---- GPU encode (device-resident rasters -> device blobs), bit-exact check
    d_in = mc_amd.DeviceBuffer(n * mc_amd.CHUNK_BYTES)
    ctx.h2d(d_in, chunks)
    cap = 20 << 20
    d_enc = mc_amd.DeviceBuffer(n * cap)
    eoffs = np.arange(n, dtype=np.uint64) * cap
    ecaps = np.full(n, cap, np.uint64)
    qv = [s_["q"] for s_ in specs]
    tv = [s_["tau"] for s_ in specs]
    log("[bench] GPU encode warmup")
    mc_amd.set_tuning("enc_keep_mib", 1 << 20)     # keep the 64-chunk workspace between the timed calls
    ctx.encode_chunks(d_in, qv, tv, d_enc, eoffs, ecaps)
    ctx.enable_timing(True)
    ctx.reset_timing()
    t0 = time.perf_counter()
    esteps = max(1, min(args.steps, 3))
    for _ in range(esteps):
        elens, est = ctx.encode_chunks(d_in, qv, tv, d_enc, eoffs, ecaps)
    enc_dt = (time.perf_counter() - t0) / esteps
    enc_k = {k: ctx.kernel_time(k) for k in ("enc_transform", "enc_entropy", "enc_assemble", "enc_copy",
                                              "enc_hash")}
    ctx.enable_timing(False)
    mc_amd.set_tuning("enc_keep_mib", 2048)
    assert (est == 0).all(), f"encode status {est}"
    blobs = []
    bad_enc = 0
    for i in range(n):
        b = np.empty(int(elens[i]), np.uint8)
        ctx.d2h(b, d_enc, offset=int(eoffs[i]))
        if len(b) != man[i]["blob_len"] or xxh64_hex(b) != man[i]["blob_xxh64"]:
            bad_enc += 1
        blobs.append(b.tobytes())
    if bad_enc:
        log(f"[bench] ERROR: {bad_enc} encoded blobs differ from the reference")
    enc_gbs = n * mc_amd.CHUNK_BYTES / enc_dt / 1e9
    log(f"[bench] GPU encode {enc_gbs:.2f} GB/s, bit-exact={bad_enc == 0}")
    d_in.free()
    d_enc.free()

    log("[bench] CPU baseline leg")
    _, cpu_enc, cpu_dec, ns = cpu_encode_and_baseline(chunks[: min(n, nthreads)], specs[: min(n, nthreads)],
                                                       nthreads)
    ref_cpu = None
    if all(s_["q"] == 6.0 and s_["tau"] == 0 for s_ in specs):
        ref_cpu = cpu_reference_baseline(chunks[:ns], blobs[:ns], man, nthreads)
    if ref_cpu:
        log("[bench] reference CPU build: " + ", ".join(f"{k} {v:.3f}" for k, v in ref_cpu.items()))

    offs = np.zeros(n, np.uint64)
    tot = 0
    for i, b in enumerate(blobs):
        offs[i] = tot
        tot += (len(b) + 255) // 256 * 256
    host = np.zeros(tot, np.uint8)
    for i, b in enumerate(blobs):
        host[int(offs[i]):int(offs[i]) + len(b)] = np.frombuffer(b, np.uint8)
    lens = np.array([len(b) for b in blobs], np.uint64)
    qs = np.array([np.frombuffer(b[:4], np.float32)[0] for b in blobs], np.float32)
    d_blobs = mc_amd.DeviceBuffer(tot)
    d_out = mc_amd.DeviceBuffer(n * mc_amd.CHUNK_BYTES)
    ctx.h2d(d_blobs, host)
    blob_bytes = int(lens.sum())

    class _View:                     # chunk i of the output buffer
        def __init__(self, buf, i):
            self.ptr = buf.ptr + i * mc_amd.CHUNK_BYTES

    views = [_View(d_out, i) for i in range(n)]
    # copy of known size: its HBM counters calibrate the per-dispatch
    # FETCH_SIZE / WRITE_SIZE readings of the decode kernels
    calib = 256 << 20
    ctx.d2d(d_out, d_out, calib, dst_off=512 << 20, src_off=0)
    ctx.sync()
    extra = {}
    if not args.no_batched:
        # all chunks in one launch sequence, measured before the per-chunk
        # leg creates its streams (idle hardware queues slow single-stream
        # work); on a context of its own, whose second stream (a sliced
        # batch's) is gone again before the per-chunk streams are created
        bctx = mc_amd.Ctx(0)
        bctx.decode_chunks(d_blobs, offs, lens, qs, d_out)
        bctx.sync()
        t0 = time.perf_counter()
        for _ in range(max(2, args.steps // 2)):
            bctx.decode_chunks(d_blobs, offs, lens, qs, d_out)
        bctx.sync()
        dtb = (time.perf_counter() - t0) / max(2, args.steps // 2)
        bctx.close()
        extra["batched_gbs"] = round(n * mc_amd.CHUNK_BYTES / dtb / 1e9, 3)
        # the batched decode's voxels against the reference's decoded hashes
        # (the per-chunk leg rewrites the same output afterwards)
        outb = np.empty((256, 256, 256), np.uint8)
        nbad = 0
        for i in range(n):
            ctx.d2h(outb, d_out, offset=i * mc_amd.CHUNK_BYTES)
            nbad += xxh64_hex(outb) != man[i]["decoded_xxh64"]
        extra["batched_bit_exact"] = nbad == 0
        log(f"[bench] batched decode {extra['batched_gbs']} GB/s, bit-exact={nbad == 0}")
    # ---- verify: blob hashes recomputed and checked against the stored ones,
    # for the resident blobs (one wave per blob) and for host blobs (gathered
    # into pinned slabs and uploaded), against the oracle's hash on one core;
    # measured, like the batched decode, before the per-chunk streams exist
    stored = np.array([int.from_bytes(b[4:12], "little") for b in blobs], np.uint64)
    ctx.verify_blobs(d_blobs, offs, lens)
    vreps = 5
    t0 = time.perf_counter()
    for _ in range(vreps):
        nbad_v, hv, _ = ctx.verify_blobs(d_blobs, offs, lens)
    dtv = (time.perf_counter() - t0) / vreps
    host_blobs = [np.frombuffer(b, np.uint8) for b in blobs]
    ctx.verify_blobs_host(host_blobs)
    t0 = time.perf_counter()
    for _ in range(2):
        nbad_h, hh, _ = ctx.verify_blobs_host(host_blobs)
    dth = (time.perf_counter() - t0) / 2
    verify_ok = nbad_v == 0 and nbad_h == 0 and (hv == stored).all() and (hh == stored).all()
    if not verify_ok:
        log("[bench] ERROR: verify disagrees with the stored blob hashes")
    cpu_v = cpu_verify_baseline(host_blobs)
    verify = {"unit": "GB/s of blob bytes", "blobs": n, "blob_bytes": blob_bytes,
              "device_resident": round(blob_bytes / dtv / 1e9, 3),
              "host_blobs": round(blob_bytes / dth / 1e9, 3), "hashes_equal_stored": bool(verify_ok),
              "cpu_oracle_1_core": round(cpu_v, 3)}
    log(f"[bench] verify resident {verify['device_resident']} GB/s, host {verify['host_blobs']} GB/s, "
        f"oracle 1 core {verify['cpu_oracle_1_core']} GB/s")
    # The per-chunk leg's streams are created after the single-stream legs
    # (idle hardware queues slow single-stream work) and before the drop-in
    # leg, so that each takes a hardware queue of its own: the drop-in leg's
    # caller threads create streams that share queues with them, but those
    # are gone or idle when the per-chunk leg runs.
    ctxs = [ctx] + [mc_amd.Ctx(0) for _ in range(args.streams - 1)]
    # per-chunk call arguments built once: each step issues exactly the C calls
    # mchip_decode_chunks(ctx, blobs, &off[i], &len[i], &q[i], 1, out_i, 16 MiB)
    Lib = mc_amd.load()
    argv = []
    for i in range(n):
        cargs = (np.array([offs[i]], np.uint64), np.array([lens[i]], np.uint64), np.array([qs[i]], np.float32))
        argv.append((C.c_void_p(d_blobs.ptr), *[a.ctypes.data_as(C.c_void_p) for a in cargs],
                     C.c_void_p(views[i].ptr), cargs))
    handles = [c_.h for c_ in ctxs]
    # the chunks go to the streams in one rotation that continues from step to
    # step, so every stream (and its hardware queue) gets the same number of
    # chunks over the run rather than a fixed share of each step
    turn = [0]

    def step():
        k = turn[0]
        for b_, o_, l_, q_, out_, _keep in argv:
            if Lib.mchip_decode_chunks(handles[k], b_, o_, l_, q_, 1, out_, mc_amd.CHUNK_BYTES) != 0:
                raise mc_amd.MchipError(mc_amd.last_error())
            k = k + 1 if k + 1 < len(handles) else 0
        turn[0] = k

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    for c_ in ctxs:
        c_.sync()
    # kernel durations from HIP events on a spread subset of the streams (every
    # stream runs the same kernels; the two event records around each timed
    # launch slow the launching, so by default one stream carries them, about
    # 30 launches of each kernel over 10 steps)
    every = max(1, len(ctxs) // max(1, args.timed_streams))
    timed_ctxs = ctxs[::every]
    for c_ in timed_ctxs:
        c_.enable_timing(True)
        c_.reset_timing()
    log(f"[bench] timing {args.steps} steps ({len(timed_ctxs)} of {len(ctxs)} streams carry kernel events)")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    for c_ in ctxs:
        c_.sync()
    dt = time.perf_counter() - t0
    ktimes = {}
    for k in ("index", "entropy", "recon"):
        tot, cnt = 0.0, 0
        for c_ in timed_ctxs:
            a_, b_ = c_.kernel_time(k)
            tot += a_
            cnt += b_
        ktimes[k] = (tot, cnt)
    for c_ in timed_ctxs:
        c_.enable_timing(False)
    ms_per_step = dt * 1000.0 / args.steps
    value = n * mc_amd.CHUNK_BYTES / (dt / args.steps) / 1e9

    # bit-exactness check of the last step's output, and the air-voxel leak:
    # source air voxels (value 0) that decode nonzero, as the reference's
    # round-trip test reports it (tests/mc_roundtrip.c:50-60); air is exactly
    # 0 only in blocks the bitmap marks absent
    out = np.empty((256, 256, 256), np.uint8)
    bad = 0
    leak = dict(air=0, air_nonzero=0, air_sq=0, material=0, mat_sq=0, max_err=0)
    for i in range(n):
        ctx.d2h(out, d_out, offset=i * mc_amd.CHUNK_BYTES)
        if xxh64_hex(out) != man[i]["decoded_xxh64"]:
            bad += 1
        src = chunks[i].reshape(-1)
        dec = out.reshape(-1)
        air = src == 0
        da = dec[air].astype(np.int64)
        e = np.abs(dec[~air].astype(np.int64) - src[~air].astype(np.int64))
        leak["air"] += int(air.sum())
        leak["air_nonzero"] += int(np.count_nonzero(da))
        leak["air_sq"] += int((da * da).sum())
        leak["material"] += int(e.size)
        leak["mat_sq"] += int((e * e).sum())
        leak["max_err"] = max(leak["max_err"], int(e.max()) if e.size else 0)
    del chunks
    if bad:
        log(f"[bench] ERROR: {bad} decoded chunks differ from the reference")
    mat_rmse = (leak["mat_sq"] / max(1, leak["material"])) ** 0.5
    air_leak = {"air_voxels": leak["air"], "air_voxels_nonzero": leak["air_nonzero"],
                "air_nonzero_frac": round(leak["air_nonzero"] / max(1, leak["air"]), 5),
                "air_rmse": round((leak["air_sq"] / max(1, leak["air"])) ** 0.5, 4),
                "material_rmse": round(mat_rmse, 4), "material_max_err": leak["max_err"],
                "psnr_db": round(10 * np.log10(255.0 ** 2 / mat_rmse ** 2), 2) if mat_rmse > 0 else 99.0,
                "note": "decoded vs generated voxels over all chunks (reference tests/mc_roundtrip.c:50-60)"}
    log(f"[bench] air leak: {air_leak}")


    # roofline of the dominant kernel (entropy decode, one launch per chunk):
    # algorithmic bytes per chunk = 16 MiB of voxels written + blob bytes read
    # (the decode path's bytes, charged to its dominant kernel), divided by
    # that kernel's mean duration from HIP events on its own stream.
    alg_bytes = (mc_amd.CHUNK_BYTES + blob_bytes / n)
    kt = {k: (v[0] / max(1, v[1])) for k, v in ktimes.items()}
    dom = max(kt, key=kt.get)
    achieved = alg_bytes / (kt[dom] / 1e3) / 1e9
    path_gbs = alg_bytes / (sum(kt.values()) / 1e3) / 1e9
    traffic = measured_traffic(args.workload, dom)
    # the decode path's HBM traffic per chunk (every decode kernel's
    # calibrated FETCH_SIZE + WRITE_SIZE) against its algorithmic bytes
    path_traffic = profile_value(args.workload, dom, "traffic", "decode_path_hbm_bytes_per_chunk")
    # issue roofline of the dominant kernel: VALU wave-instructions per launch
    # (SQ_INSTS_VALU summed over its waves) over its mean duration, against
    # one wave64 VALU instruction per SIMD per 4 cycles on 256 CUs x 4 SIMDs
    valu_per_launch = profile_value(args.workload, dom, "valu", "valu_insts_per_launch")
    valu_peak = 256 * 4 * 2.4e9 / 4.0               # 2.4 GHz: MI355X peak engine clock
    valu_roof = None
    if valu_per_launch:
        ach = valu_per_launch / (kt[dom] / 1e3)
        valu_roof = {"bound": "valu_issue", "achieved": round(ach / 1e9, 2), "peak": round(valu_peak / 1e9, 1),
                     "unit": "G wave64 VALU instructions/s", "frac": round(ach / valu_peak, 4),
                     "valu_insts_per_launch": int(valu_per_launch),
                     "note": "per launch alone; up to 21 chunks decode at once in the timed steps"}
    rec = {
        "metric": "decode GB/s of decoded voxels on 1xMI355X, bit-exact (64 x 256^3 u8 chunks, q=6, one chunk per launch)",
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "bit_exact": bad == 0,
        "config": {"workload": args.workload, "chunks": n, "chunk": "256^3 u8", "q": 6.0, "air": 0.419,
                   "decoded_bytes_per_step": n * mc_amd.CHUNK_BYTES, "blob_bytes": blob_bytes,
                   "launch": "one chunk per launch", "streams": args.streams, "hw_queues": QUEUES, "tuning": args.tune,
                   "kernel_events_on_streams": len(timed_ctxs)},
        "kernels_ms_avg": {k: round(v, 4) for k, v in kt.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "alg_bytes_per_launch": round(alg_bytes), "decode_path_gbs": round(path_gbs, 2),
                     "decode_path_traffic_ratio": (round(path_traffic / alg_bytes, 3) if path_traffic else None),
                     "valu": valu_roof},
        "cpu_baseline": cpu_record(ref_cpu, cpu_dec, cpu_enc, ns, n, nthreads),
        "encode": {"value": round(enc_gbs, 3), "unit": "GB/s", "byte_identical": bad_enc == 0,
                   "launch": (f"{min(n, enc_batch)} chunks per launch" if enc_batch
                              else f"{n} chunks per launch (automatic sequence size, 16-256 chunks)"),
                   "kernels_ms_per_chunk": {k: round(v[0] / (esteps * n), 4) for k, v in enc_k.items()}},
        "verify": verify,
        "air_leak": air_leak,
        "dropin_host_buffers": dropin,
    }
    rec.update(extra)
    d_blobs.free()
    d_out.free()
    for c_ in ctxs:
        c_.close()
    if mixed is not None:
        rec["mixed"] = mixed
    if serving is not None:
        rec["render_serving"] = serving.pop("render_serving", None)
        rec["serving"] = serving
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
