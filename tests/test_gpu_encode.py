"""Encode parity of the HIP path: blobs byte-identical to the reference
(golden manifest) and to the oracle."""
import json
import os

import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_encode_matches_reference_manifest(gpu_ctx):
    import mc_amd
    man = json.load(open(os.path.join(ROOT, "tests", "golden", "chunk_manifest.json")))
    specs = [man["plain64_q6"][0], man["plain64_q6"][1]] + [man["mixed256"][k] for k in range(8)]
    chunks = mc_amd.synth_chunks([s["seed"] for s in specs], [s["air"] for s in specs])
    print(f"encoding {len(specs)} chunks on the GPU", flush=True)
    blobs = gpu_ctx.encode_chunks_host(chunks, [s["q"] for s in specs], [s["tau"] for s in specs])
    for k, s in enumerate(specs):
        assert len(blobs[k]) == s["blob_len"], f"chunk {k} seed {s['seed']} q={s['q']} tau={s['tau']}"
        assert f"{ob.xxh64(blobs[k]):016x}" == s["blob_xxh64"], f"chunk {k} seed {s['seed']}"


def test_encode_edge_chunks_match_oracle(gpu_ctx):
    rng = np.random.default_rng(12)
    ch = np.zeros((5, 256, 256, 256), np.uint8)
    ch[1, 10:12, 3, 250] = 9                                  # sparse
    ch[1, 255, 255, 255] = 200
    ch[2] = rng.integers(0, 256, (256, 256, 256), dtype=np.uint8)   # noise (large payloads)
    ch[3] = 128                                               # constant
    ch[4, ::2] = 77                                           # stripes, half air
    qs = [6.0, 6.0, 0.5, 6.0, 1.0]
    taus = [0, 3, 1, 0, 4]
    print("encoding 5 edge-case chunks", flush=True)
    blobs = gpu_ctx.encode_chunks_host(ch, qs, taus)
    assert blobs[0] == b""
    for k in range(1, 5):
        want = ob.encode_chunk(ch[k], qs[k], taus[k])
        assert blobs[k] == want, f"chunk {k}"


def test_encode_decode_roundtrip_tau_bound(gpu_ctx):
    import mc_amd
    chunks = mc_amd.synth_chunks([77, 78], [0.6, 0.3])
    blobs = gpu_ctx.encode_chunks_host(chunks, [2.5, 16.0], [8, 32])
    out = gpu_ctx.decode_chunks_host(blobs)
    for k, tau in enumerate([8, 32]):
        err = np.abs(out[k].astype(np.int16) - chunks[k].astype(np.int16))
        assert err.max() <= tau


def test_encode_graph_replays_equal_reference(gpu_ctx):
    """One encode sequence captured as a HIP graph: every replay (with eager
    encodes and a scrambled output buffer between replays) writes the
    reference's blobs (manifest), and the graph's results give their
    lengths."""
    import mc_amd
    man = json.load(open(os.path.join(ROOT, "tests", "golden", "chunk_manifest.json")))
    specs = [man["plain64_q6"][2]] + [man["mixed256"][k] for k in (3, 9, 17)]
    chunks = mc_amd.synth_chunks([s["seed"] for s in specs], [s["air"] for s in specs])
    n = len(specs)
    cap = 20 << 20
    d_in = mc_amd.DeviceBuffer(n * mc_amd.CHUNK_BYTES)
    d_out = mc_amd.DeviceBuffer(n * cap)
    d_tmp = mc_amd.DeviceBuffer(cap)
    gpu_ctx.h2d(d_in, chunks)
    offs = np.arange(n, dtype=np.uint64) * cap
    caps = np.full(n, cap, np.uint64)
    qs = [s["q"] for s in specs]
    taus = [s["tau"] for s in specs]
    print("encode graph", flush=True)
    g, gn = gpu_ctx.encode_graph(d_in, qs, taus, d_out, offs, caps)
    try:
        for rep in range(3):
            gpu_ctx.encode_chunks(d_in, qs[rep:rep + 1], taus[rep:rep + 1], d_tmp, [0], [cap])
            gpu_ctx.h2d(d_out, np.full(n * cap, 0x33, np.uint8))
            gpu_ctx.graph_launch(g)
            gpu_ctx.sync()
            lens, st = gpu_ctx.encode_graph_results(g, gn)
            assert (st == 0).all()
            for k, s in enumerate(specs):
                assert int(lens[k]) == s["blob_len"], f"replay {rep} chunk {k} length"
                b = np.empty(int(lens[k]), np.uint8)
                gpu_ctx.d2h(b, d_out, offset=int(offs[k]))
                assert f"{ob.xxh64(b):016x}" == s["blob_xxh64"], f"replay {rep} chunk {k}"
    finally:
        gpu_ctx.graph_destroy(g)
        d_in.free()
        d_out.free()
        d_tmp.free()


def test_sliced_encode_on_two_streams_equals_reference(gpu_ctx):
    """An encode sequence split into chunk slices (alternate slices on two
    streams, the entropy pass heaviest chunk first within each slice) writes
    the reference's blobs (manifest) for every slice count, eagerly and as HIP
    graph replays (blocks costliest first) with eager work and a scrambled
    output between them."""
    import mc_amd
    man = json.load(open(os.path.join(ROOT, "tests", "golden", "chunk_manifest.json")))
    specs = [man["plain64_q6"][3]] + [man["mixed256"][k] for k in (1, 4, 6, 11, 20)]
    chunks = mc_amd.synth_chunks([s["seed"] for s in specs], [s["air"] for s in specs])
    n = len(specs)
    cap = 20 << 20
    d_in = mc_amd.DeviceBuffer(n * mc_amd.CHUNK_BYTES)
    d_out = mc_amd.DeviceBuffer(n * cap)
    d_tmp = mc_amd.DeviceBuffer(cap)
    gpu_ctx.h2d(d_in, chunks)
    offs = np.arange(n, dtype=np.uint64) * cap
    caps = np.full(n, cap, np.uint64)
    qs = [s["q"] for s in specs]
    taus = [s["tau"] for s in specs]

    def check(what, lens):
        for k, s in enumerate(specs):
            assert int(lens[k]) == s["blob_len"], f"{what}: chunk {k} length"
            b = np.empty(int(lens[k]), np.uint8)
            gpu_ctx.d2h(b, d_out, offset=int(offs[k]))
            assert f"{ob.xxh64(b):016x}" == s["blob_xxh64"], f"{what}: chunk {k}"

    try:
        for ns in (1, 2, 3, 6):
            mc_amd.set_tuning("enc_slices", ns)
            print(f"enc_slices={ns}", flush=True)
            gpu_ctx.h2d(d_out, np.full(n * cap, 0x5A, np.uint8))
            lens, st = gpu_ctx.encode_chunks(d_in, qs, taus, d_out, offs, caps)
            assert (st == 0).all()
            check(f"enc_slices={ns}", lens)
        mc_amd.set_tuning("enc_slices", 3)
        mc_amd.set_tuning("enc_lanes", 1)
        print("encode graph, 3 slices, blocks costliest first", flush=True)
        g, gn = gpu_ctx.encode_graph(d_in, qs, taus, d_out, offs, caps)
        try:
            for rep in range(3):
                gpu_ctx.encode_chunks(d_in, qs[rep:rep + 1], taus[rep:rep + 1], d_tmp, [0], [cap])
                gpu_ctx.h2d(d_out, np.full(n * cap, 0x33, np.uint8))
                gpu_ctx.graph_launch(g)
                gpu_ctx.sync()
                lens, st = gpu_ctx.encode_graph_results(g, gn)
                assert (st == 0).all()
                check(f"replay {rep}", lens)
        finally:
            gpu_ctx.graph_destroy(g)
    finally:
        mc_amd.set_tuning("enc_slices", 0)
        mc_amd.set_tuning("enc_lanes", 0)
        d_in.free()
        d_out.free()
        d_tmp.free()


def test_encode_tiny_blocks_and_unaligned_outputs_match_oracle(gpu_ctx):
    """Blobs written at odd output offsets (every alignment of the payload
    copy), from chunks whose present blocks have payloads of a few bytes (a
    64-byte piece of the payload stream then spans many blocks), from a
    noise chunk (near-full slots) and from a mixed one, equal the oracle's
    byte for byte, with the entropy pass's blocks in raster order and
    costliest first (most blocks of the sparse chunks are empty)."""
    import mc_amd
    rng = np.random.default_rng(31)
    ch = np.zeros((5, 256, 256, 256), np.uint8)
    ch[0, ::16, ::16, ::16] = 200                                   # one voxel per block
    ch[1][rng.random((256, 256, 256)) < 0.002] = 90                 # scattered voxels
    ch[2] = rng.integers(0, 256, (256, 256, 256), dtype=np.uint8)   # noise
    ch[3] = mc_amd.synth_chunks([5], [0.5])[0]
    ch[4, 3::16, 5::16, 7::16] = 17                                 # one voxel per block, elsewhere
    qs = [6.0, 2.0, 0.5, 6.0, 16.0]
    taus = [0, 2, 0, 8, 0]
    n = len(qs)
    cap = 20 << 20
    d_in = mc_amd.DeviceBuffer(n * mc_amd.CHUNK_BYTES)
    d_out = mc_amd.DeviceBuffer(n * cap + 64)
    gpu_ctx.h2d(d_in, ch)
    offs = np.array([k * cap + (k % 4) for k in range(n)], np.uint64)
    caps = np.full(n, cap - 8, np.uint64)
    wants = [ob.encode_chunk(ch[k], qs[k], taus[k]) for k in range(n)]
    try:
        for lanes in (0, 1):
            mc_amd.set_tuning("enc_lanes", lanes)
            print(f"encoding 5 chunks at odd output offsets, enc_lanes={lanes}", flush=True)
            gpu_ctx.h2d(d_out, np.full(n * cap + 64, 0x5A, np.uint8))
            lens, st = gpu_ctx.encode_chunks(d_in, qs, taus, d_out, offs, caps)
            assert (st == 0).all(), st
            for k in range(n):
                b = np.empty(int(lens[k]), np.uint8)
                gpu_ctx.d2h(b, d_out, offset=int(offs[k]))
                assert b.tobytes() == wants[k], f"enc_lanes={lanes} chunk {k}: {len(b)} vs {len(wants[k])} bytes"
    finally:
        mc_amd.set_tuning("enc_lanes", 0)
        d_in.free()
        d_out.free()


def test_encode_autotuner_and_every_encode_shape_keep_blobs_identical(gpu_ctx):
    """Every encode shape the autotuner searches (entropy-pass chunk order and
    block order, chunks per launch sequence, chunk slices) writes the reference's blobs
    (manifest); so do the autotuned shapes cached for the chunks' signature."""
    import mc_amd
    man = json.load(open(os.path.join(ROOT, "tests", "golden", "chunk_manifest.json")))
    specs = [man["mixed256"][k] for k in (2, 5, 7, 12, 30)]
    chunks = mc_amd.synth_chunks([s["seed"] for s in specs], [s["air"] for s in specs])
    n = len(specs)
    cap = 20 << 20
    d_in = mc_amd.DeviceBuffer(n * mc_amd.CHUNK_BYTES)
    d_out = mc_amd.DeviceBuffer(n * cap)
    gpu_ctx.h2d(d_in, chunks)
    offs = np.arange(n, dtype=np.uint64) * cap
    caps = np.full(n, cap, np.uint64)
    qs = [s["q"] for s in specs]
    taus = [s["tau"] for s in specs]

    def check(what):
        gpu_ctx.h2d(d_out, np.full(n * cap, 0x77, np.uint8))
        lens, st = gpu_ctx.encode_chunks(d_in, qs, taus, d_out, offs, caps)
        assert (st == 0).all(), what
        for k, s in enumerate(specs):
            assert int(lens[k]) == s["blob_len"], f"{what}: chunk {k} length"
            b = np.empty(int(lens[k]), np.uint8)
            gpu_ctx.d2h(b, d_out, offset=int(offs[k]))
            assert f"{ob.xxh64(b):016x}" == s["blob_xxh64"], f"{what}: chunk {k}"

    try:
        for key, vals, default in [("enc_order", [0, 1], 1), ("enc_lanes", [1, 0], 0), ("enc_batch", [2, 3, 0], 0),
                                   ("enc_slices", [2, 0], 0)]:
            for v in vals:
                mc_amd.set_tuning(key, v)
                print(f"{key}={v}", flush=True)
                check(f"{key}={v}")
            mc_amd.set_tuning(key, default)
        mc_amd.set_tuning("enc_lanes", 1)
        mc_amd.set_tuning("enc_slices", 2)
        print("enc_lanes=1 enc_slices=2", flush=True)
        check("enc_lanes=1 enc_slices=2")
        mc_amd.set_tuning("enc_lanes", 0)
        mc_amd.set_tuning("enc_slices", 0)
        print("autotune encode", flush=True)
        rep = gpu_ctx.autotune_encode(d_in, qs, taus, d_out, offs, caps)
        assert len(rep) >= 2 and all(r["ms"] > 0 for r in rep)
        check("autotuned")
    finally:
        mc_amd.set_tuning("tuned_clear", 1)
        mc_amd.set_tuning("enc_order", 1)
        mc_amd.set_tuning("enc_lanes", 0)
        mc_amd.set_tuning("enc_batch", 0)
        mc_amd.set_tuning("enc_slices", 0)
        d_in.free()
        d_out.free()
