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
