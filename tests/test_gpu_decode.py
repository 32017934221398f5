"""Decode parity of the HIP path against the oracle and the reference's
golden fixtures (bit-exact)."""
import json
import os

import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def kat_payload(kat, i):
    return kat["payload"][kat["offs"][i]:kat["offs"][i + 1]].tobytes()


def test_block_decode_matches_reference_kat(gpu_ctx, kat):
    idx = [i for i in range(len(kat["q"])) if kat["offs"][i + 1] > kat["offs"][i]]
    print(f"decoding {len(idx)} KAT payloads", flush=True)
    out = gpu_ctx.decode_blocks_host([kat_payload(kat, i) for i in idx], kat["q"][idx])
    for j, i in enumerate(idx):
        assert np.array_equal(out[j], kat["dec"][i]), f"case {i} {kat['names'][i]} q={kat['q'][i]}"


def test_block_decode_corrupt_payloads_match_oracle(gpu_ctx, kat):
    rng = np.random.default_rng(9)
    pays, qs = [], []
    for trial in range(300):
        i = int(rng.integers(0, len(kat["q"])))
        p = bytearray(kat_payload(kat, i) or b"\x00")
        for _ in range(int(rng.integers(1, 4))):
            p[int(rng.integers(0, len(p)))] ^= int(rng.integers(1, 256))
        if trial % 4 == 0:
            p = p[: int(rng.integers(0, len(p) + 1))]
        pays.append(bytes(p))
        qs.append(float(kat["q"][i]))
    print("decoding 300 corrupted payloads", flush=True)
    out = gpu_ctx.decode_blocks_host(pays, np.array(qs, np.float32))
    for k in range(len(pays)):
        assert np.array_equal(out[k], ob.dec_block(pays[k], qs[k])), f"trial {k}"


def test_block_decode_absent_empty_and_present_in_one_call(gpu_ctx, kat):
    """One call with absent blocks (no payload: zeros) between present ones,
    an empty payload and a truncated one: every present block equals the
    oracle's decode of its payload, every absent one is zeros."""
    idx = [i for i in range(len(kat["q"])) if kat["offs"][i + 1] > kat["offs"][i]][:12]
    pays, qs = [], []
    for k, i in enumerate(idx):
        if k % 3 == 0:
            pays.append(None)
            qs.append(6.0)
        p = kat_payload(kat, i)
        pays.append(p if k % 4 else p[: len(p) // 2])
        qs.append(float(kat["q"][i]))
    pays.insert(5, b"")
    qs.insert(5, 6.0)
    pays.append(None)
    qs.append(1.0)
    print(f"decoding {len(pays)} blocks, {sum(p is None for p in pays)} absent", flush=True)
    out = gpu_ctx.decode_blocks_host(pays, np.array(qs, np.float32))
    for k, p in enumerate(pays):
        want = np.zeros(4096, np.uint8) if p is None else ob.dec_block(p, qs[k])
        assert np.array_equal(out[k], want), f"block {k} ({'absent' if p is None else len(p)})"


def _chunks(specs):
    import mc_amd
    return mc_amd.synth_chunks([s["seed"] for s in specs], [s["air"] for s in specs])


def test_chunk_decode_matches_oracle_and_manifest(gpu_ctx):
    man = json.load(open(os.path.join(ROOT, "tests", "golden", "chunk_manifest.json")))
    specs = [man["plain64_q6"][0], man["mixed256"][0], man["mixed256"][1], man["mixed256"][5]]
    chunks = _chunks(specs)
    blobs = []
    for s, ch in zip(specs, chunks):
        assert f"{ob.xxh64(ch):016x}" == s["input_xxh64"], "generator output drifted"
        b = ob.encode_chunk(ch, s["q"], s["tau"])
        assert f"{ob.xxh64(b):016x}" == s["blob_xxh64"]
        blobs.append(b)
    print("decoding 4 chunks on the GPU", flush=True)
    out = gpu_ctx.decode_chunks_host(blobs)
    for k, s in enumerate(specs):
        assert f"{ob.xxh64(out[k]):016x}" == s["decoded_xxh64"], f"chunk {k} seed {s['seed']}"


def test_chunk_decode_truncated_and_corrupt_blobs_match_oracle(gpu_ctx):
    import mc_amd
    ch = mc_amd.synth_chunks([42], [0.5])[0]
    blob = bytearray(ob.encode_chunk(ch, 3.0, 2))
    rng = np.random.default_rng(4)
    variants = [bytes(blob[: len(blob) - 1000]), bytes(blob[: len(blob) // 2])]
    b2 = bytearray(blob)
    for _ in range(50):
        b2[int(rng.integers(700, len(b2)))] ^= 0x5A
    variants.append(bytes(b2))
    want = [ob.decode_chunk(v) for v in variants]
    out = gpu_ctx.decode_chunks_host(variants)
    for k, v in enumerate(variants):
        assert np.array_equal(out[k], want[k]), f"variant {k}"
    # the persistent decoder with every payload window size (a 16-word window
    # slides every 64 payload bytes, also across the corrupt and cut-off ends)
    try:
        mc_amd.set_tuning("persist_min", 1)
        for win in (16, 64, 32):
            mc_amd.set_tuning("persist_win", win)
            print(f"persist_win={win}", flush=True)
            out = gpu_ctx.decode_chunks_host(variants)
            for k in range(len(variants)):
                assert np.array_equal(out[k], want[k]), f"persist_win={win} variant {k}"
    finally:
        mc_amd.set_tuning("persist_win", 32)
        mc_amd.set_tuning("persist_min", 2)


def test_all_air_and_sparse_chunks(gpu_ctx):
    ch = np.zeros((2, 256, 256, 256), np.uint8)
    ch[1, 100:105, 7, 200] = 77
    ch[1, 255, 255, 255] = 3
    blobs = [ob.encode_chunk(ch[1], 6.0, 0)]
    assert ob.encode_chunk(ch[0], 6.0, 0) == b""
    out = gpu_ctx.decode_chunks_host(blobs)
    assert np.array_equal(out[0], ob.decode_chunk(blobs[0]))


def test_device_resident_batch_equals_per_chunk(gpu_ctx):
    import mc_amd
    specs = [dict(seed=s, air=a) for s, a in [(7, 0.3), (8, 0.6), (9, 0.9)]]
    chunks = _chunks(specs)
    blobs = [ob.encode_chunk(c, q, 0) for c, q in zip(chunks, [6.0, 1.0, 32.0])]
    offs = np.zeros(len(blobs), np.uint64)
    tot = 0
    for k, b in enumerate(blobs):
        offs[k] = tot
        tot += (len(b) + 255) // 256 * 256
    buf = np.zeros(tot, np.uint8)
    for k, b in enumerate(blobs):
        buf[int(offs[k]):int(offs[k]) + len(b)] = np.frombuffer(b, np.uint8)
    d_in = mc_amd.DeviceBuffer(tot)
    d_out = mc_amd.DeviceBuffer(len(blobs) * mc_amd.CHUNK_BYTES)
    gpu_ctx.h2d(d_in, buf)
    qs = [np.frombuffer(b[:4], np.float32)[0] for b in blobs]
    gpu_ctx.decode_chunks(d_in, offs, [len(b) for b in blobs], qs, d_out)
    gpu_ctx.sync()
    got = np.empty((len(blobs), 256, 256, 256), np.uint8)
    gpu_ctx.d2h(got, d_out)
    for k in range(len(blobs)):
        assert np.array_equal(got[k], ob.decode_chunk(blobs[k]))
    # one chunk per launch gives the same bytes
    for k in range(len(blobs)):
        gpu_ctx.decode_chunks(d_in, offs[k:k + 1], [len(blobs[k])], qs[k:k + 1], d_out)
    gpu_ctx.sync()
    one = np.empty((256, 256, 256), np.uint8)
    gpu_ctx.d2h(one, d_out)
    assert np.array_equal(one, got[-1])
    d_in.free()
    d_out.free()


def _resident(gpu_ctx, blobs):
    import mc_amd
    offs = np.zeros(len(blobs), np.uint64)
    tot = 0
    for k, b in enumerate(blobs):
        offs[k] = tot
        tot += (len(b) + 255) // 256 * 256
    buf = np.zeros(tot, np.uint8)
    for k, b in enumerate(blobs):
        buf[int(offs[k]):int(offs[k]) + len(b)] = np.frombuffer(b, np.uint8)
    d_in = mc_amd.DeviceBuffer(tot)
    gpu_ctx.h2d(d_in, buf)
    qs = [np.frombuffer(b[:4], np.float32)[0] for b in blobs]
    return d_in, offs, [len(b) for b in blobs], qs


def test_persistent_batch_and_graph_replays_equal_eager(gpu_ctx):
    """A mixed-quality batch through the persistent launch and through HIP
    graph replays (with eager work between replays) equals the oracle."""
    import mc_amd
    specs = [dict(seed=s, air=a) for s, a in [(21, 0.3), (22, 0.7), (23, 0.5), (24, 0.9)]]
    chunks = _chunks(specs)
    blobs = [ob.encode_chunk(c, q, t) for c, (q, t) in zip(chunks, [(0.5, 1), (6.0, 16), (2.5, 8), (64.0, 128)])]
    want = [ob.decode_chunk(b) for b in blobs]
    d_in, offs, lens, qs = _resident(gpu_ctx, blobs)
    d_out = mc_amd.DeviceBuffer(len(blobs) * mc_amd.CHUNK_BYTES)
    d_tmp = mc_amd.DeviceBuffer(mc_amd.CHUNK_BYTES)
    got = np.empty((len(blobs), 256, 256, 256), np.uint8)
    for persistent in (1, 0):
        mc_amd.set_tuning("persistent", persistent)
        print(f"batch persistent={persistent}", flush=True)
        gpu_ctx.decode_chunks(d_in, offs, lens, qs, d_out)
        gpu_ctx.sync()
        gpu_ctx.d2h(got, d_out)
        for k in range(len(blobs)):
            assert np.array_equal(got[k], want[k]), f"persistent={persistent} chunk {k}"
    mc_amd.set_tuning("persistent", 1)
    print("graph", flush=True)
    g = gpu_ctx.decode_graph(d_in, offs, lens, qs, d_out)
    try:
        for rep in range(3):
            # eager work between replays, then a replay over a scrambled output
            gpu_ctx.decode_chunks(d_in, offs[1:2], lens[1:2], qs[1:2], d_tmp)
            gpu_ctx.h2d(d_out, np.full(len(blobs) * mc_amd.CHUNK_BYTES, 7, np.uint8))
            gpu_ctx.graph_launch(g)
            gpu_ctx.sync()
            gpu_ctx.d2h(got, d_out)
            for k in range(len(blobs)):
                assert np.array_equal(got[k], want[k]), f"replay {rep} chunk {k}"
    finally:
        gpu_ctx.graph_destroy(g)
    d_in.free()
    d_out.free()
    d_tmp.free()


def test_sliced_batch_on_two_streams_equals_oracle(gpu_ctx):
    """A batched decode split into chunk slices (each slice's reconstruction
    on one of two streams, overlapping the next slice's entropy pass) equals
    the oracle for every slice count, eagerly and as HIP graph replays with
    eager work and a scrambled output between the replays."""
    import mc_amd
    specs = [dict(seed=40 + k, air=a) for k, a in enumerate([0.3, 0.5, 0.7, 0.9, 0.4, 0.6, 0.8])]
    chunks = _chunks(specs)
    qt = [(6.0, 0), (0.5, 1), (16.0, 32), (2.5, 8), (6.0, 16), (64.0, 128), (1.0, 4)]
    blobs = [ob.encode_chunk(c, q, t) for c, (q, t) in zip(chunks, qt)]
    want = [ob.decode_chunk(b) for b in blobs]
    d_in, offs, lens, qs = _resident(gpu_ctx, blobs)
    d_out = mc_amd.DeviceBuffer(len(blobs) * mc_amd.CHUNK_BYTES)
    d_tmp = mc_amd.DeviceBuffer(mc_amd.CHUNK_BYTES)
    got = np.empty((len(blobs), 256, 256, 256), np.uint8)

    def check(what):
        gpu_ctx.d2h(got, d_out)
        for k in range(len(blobs)):
            assert np.array_equal(got[k], want[k]), f"{what}: chunk {k}"

    try:
        for ns in (0, 1, 2, 3, 7):
            mc_amd.set_tuning("slices", ns)
            print(f"slices={ns}", flush=True)
            gpu_ctx.h2d(d_out, np.full(len(blobs) * mc_amd.CHUNK_BYTES, 3, np.uint8))
            gpu_ctx.decode_chunks(d_in, offs, lens, qs, d_out)
            gpu_ctx.sync()
            check(f"slices={ns}")
        mc_amd.set_tuning("slices", 3)
        print("graph, 3 slices", flush=True)
        g = gpu_ctx.decode_graph(d_in, offs, lens, qs, d_out)
        try:
            for rep in range(3):
                gpu_ctx.decode_chunks(d_in, offs[2:3], lens[2:3], qs[2:3], d_tmp)
                gpu_ctx.h2d(d_out, np.full(len(blobs) * mc_amd.CHUNK_BYTES, 7, np.uint8))
                gpu_ctx.graph_launch(g)
                gpu_ctx.sync()
                check(f"replay {rep}")
        finally:
            gpu_ctx.graph_destroy(g)
    finally:
        mc_amd.set_tuning("slices", 0)
        d_in.free()
        d_out.free()
        d_tmp.free()


def test_every_launch_shape_and_the_autotuner_keep_results_exact(gpu_ctx):
    """Every value of every decode launch-shape knob the autotuner searches
    decodes a mixed-quality batch exactly as the oracle; so do the autotuned
    shapes cached for the batch's signature."""
    import mc_amd
    specs = [dict(seed=s, air=a) for s, a in [(31, 0.4), (32, 0.6), (33, 0.8)]]
    chunks = _chunks(specs)
    blobs = [ob.encode_chunk(c, q, t) for c, (q, t) in zip(chunks, [(1.0, 4), (16.0, 0), (6.0, 16)])]
    want = [ob.decode_chunk(b) for b in blobs]
    d_in, offs, lens, qs = _resident(gpu_ctx, blobs)
    d_out = mc_amd.DeviceBuffer(len(blobs) * mc_amd.CHUNK_BYTES)
    got = np.empty((len(blobs), 256, 256, 256), np.uint8)
    defaults = [("decoder", 3), ("persistent", 1), ("persist_min", 2), ("persist_wg", 0), ("persist_win", 32),
                ("recon_bpw", 1), ("len_shift", 3), ("group_shift", 6)]

    def check(what):
        gpu_ctx.decode_chunks(d_in, offs, lens, qs, d_out)
        gpu_ctx.sync()
        gpu_ctx.d2h(got, d_out)
        for k in range(len(blobs)):
            assert np.array_equal(got[k], want[k]), f"{what}: chunk {k}"

    try:
        knobs = [("decoder", [0, 2, 3]), ("persistent", [0, 1]), ("persist_wg", [0, 64, 300]),
                 ("persist_win", [16, 64, 32]), ("recon_bpw", [1, 2, 4]), ("len_shift", [2, 3, 4, 5]),
                 ("group_shift", [4, 5, 6, 7, 8])]
        for key, vals in knobs:
            for v in vals:
                mc_amd.set_tuning(key, v)
                print(f"{key}={v}", flush=True)
                check(f"{key}={v}")
            mc_amd.set_tuning(key, dict(defaults)[key])
        print("autotune", flush=True)
        rep = gpu_ctx.autotune_decode(d_in, offs, lens, qs, d_out)
        assert len(rep) >= 6 and all(r["ms"] > 0 for r in rep)
        check("autotuned")
        mc_amd.set_tuning("tuned_clear", 1)
        check("after clearing the autotuned shapes")
    finally:
        for k, v in defaults:
            mc_amd.set_tuning(k, v)
        mc_amd.set_tuning("tuned_clear", 1)
        d_in.free()
        d_out.free()


def _replace_payloads(blob, repl):
    """The chunk blob with the payloads of some present blocks replaced
    (repl: block index -> payload bytes); header, fraction map and bitmap
    kept, lengths rewritten."""
    b = np.frombuffer(blob, np.uint8)
    fml = int(b[12]) | (int(b[13]) << 8)
    bm = 14 + fml
    bits = np.unpackbits(b[bm:bm + 512], bitorder="little")
    present = np.nonzero(bits)[0]
    lens = b[bm + 512:bm + 512 + 2 * len(present)].view("<u2").astype(np.int64)
    offs = bm + 512 + 2 * len(present) + np.concatenate([[0], np.cumsum(lens)[:-1]])
    pays = [repl.get(int(bi), b[offs[k]:offs[k] + lens[k]].tobytes()) for k, bi in enumerate(present)]
    new_lens = np.array([len(p) for p in pays], "<u2")
    return bytes(b[:bm + 512]) + new_lens.tobytes() + b"".join(pays)


def test_chunk_decode_parallel_and_ordered_corrections(gpu_ctx):
    """k_recon applies a block's corrections in parallel when their positions
    strictly increase and in list order otherwise: blocks with 300 spread
    corrections, with repeated positions whose clamps depend on the order, and
    with positions that wrap past 2^32 decode exactly as in the oracle."""
    import ctypes as C
    from test_device_logic_cpu import harness, qtab
    import mc_amd
    H = harness()
    ch = mc_amd.synth_chunks([77], [0.3])[0]
    q = 6.0
    blob = ob.encode_chunk(ch, q, 0)
    t = qtab(q)
    rng = np.random.default_rng(21)

    def craft(gaps, negs, mags, dc):
        out = np.zeros(1 << 16, np.uint8)
        n = H.dl_encode_corrections(ob.p(t), dc, len(gaps), ob.p(np.array(gaps, np.uint32)),
                                    ob.p(np.array(negs, np.int32)), ob.p(np.array(mags, np.uint32)), ob.p(out),
                                    out.size)
        return out[:n].tobytes()

    bits = np.unpackbits(np.frombuffer(blob, np.uint8)[14 + (blob[12] | (blob[13] << 8)):][:512], bitorder="little")
    present = [int(i) for i in np.nonzero(bits)[0]]
    spread = sorted(rng.choice(4096, 300, replace=False))
    repl = {
        present[0]: craft([int(spread[0])] + [int(b - a) for a, b in zip(spread, spread[1:])],
                          list(rng.integers(0, 2, 300)), list(rng.integers(1, 200, 300)), 90),
        present[1]: craft([5, 0, 0, 3, 0], [0, 0, 1, 1, 0], [200, 100, 150, 90, 7], 128),
        present[2]: craft([4000, 0xFFFFFFFF - 3999, 10, 0], [0, 1, 0, 1], [50, 60, 70, 80], 200),
    }
    crafted = _replace_payloads(blob, repl)
    print("decoding a chunk with crafted correction lists", flush=True)
    out = gpu_ctx.decode_chunks_host([crafted])
    want = ob.decode_chunk(crafted)
    assert np.array_equal(out[0], want.reshape(out[0].shape))
