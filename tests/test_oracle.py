"""Pins the CPU oracle: reference known answers, published xxHash values,
closed forms of the transform and range-coder identities."""
import json
import os

import numpy as np
import pytest

import oracle_bind as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def kat_payload(kat, i):
    return kat["payload"][kat["offs"][i]:kat["offs"][i + 1]].tobytes()


def test_oracle_block_encode_matches_reference_kat(kat):
    for i in range(len(kat["q"])):
        pay = ob.enc_block(kat["vox"][i], float(kat["q"][i]), int(kat["tau"][i]))
        assert pay == kat_payload(kat, i), f"case {i} {kat['names'][i]} q={kat['q'][i]} tau={kat['tau'][i]}"


def test_oracle_block_decode_matches_reference_kat(kat):
    for i in range(len(kat["q"])):
        pay = kat_payload(kat, i)
        if not pay:
            assert not kat["vox"][i].any()
            continue
        dec = ob.dec_block(pay, float(kat["q"][i]))
        assert np.array_equal(dec, kat["dec"][i]), f"case {i}"


def test_kat_tau_bound_and_all_air(kat):
    # same-build tau bound (reference tests/mc_api_test.c:45-66) on the fixtures themselves
    for i in range(len(kat["q"])):
        tau = int(kat["tau"][i])
        if tau > 0 and kat["vox"][i].any():
            err = np.abs(kat["dec"][i].astype(int) - kat["vox"][i].astype(int)).max()
            assert err <= tau
    assert ob.enc_block(np.zeros(4096, np.uint8), 6.0, 3) == b""


def test_xxh64_published_known_answers():
    xxhash = pytest.importorskip("xxhash")
    assert ob.xxh64(b"", 0) == 0xEF46DB3751D8E999
    rng = np.random.default_rng(3)
    for n in [1, 3, 4, 7, 8, 31, 32, 33, 63, 64, 100, 1000, 4099]:
        data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        for seed in [0, 1, 0x6D636368756E6B, (1 << 64) - 1]:
            assert ob.xxh64(data, seed) == xxhash.xxh64(data, seed=seed).intdigest()


def test_dct_matrix_orthonormal_and_closed_form():
    cm = np.zeros(256, np.float32)
    ob.oracle().orc_dct_matrix(ob.p(cm))
    cm = cm.reshape(16, 16).astype(np.float64)
    k = np.arange(16)[:, None]
    n = np.arange(16)[None, :]
    ref = np.where(k == 0, np.sqrt(1 / 16), np.sqrt(2 / 16)) * np.cos(np.pi * (2 * n + 1) * k / 32)
    assert np.abs(cm - ref).max() < 1e-7
    assert np.abs(cm @ cm.T - np.eye(16)).max() < 1e-6


def test_dct3_against_float64_separable_product():
    cm = np.zeros(256, np.float32)
    ob.oracle().orc_dct_matrix(ob.p(cm))
    M = cm.reshape(16, 16).astype(np.float64)
    rng = np.random.default_rng(5)
    x = rng.integers(-256, 256, (16, 16, 16)).astype(np.float32)
    out = np.zeros_like(x)
    ob.oracle().orc_dct3_fwd(ob.p(x), ob.p(out))
    want = np.einsum("ka,lb,mc,abc->klm", M, M, M, x.astype(np.float64))
    assert np.abs(out - want).max() < 2e-3
    back = np.zeros_like(x)
    ob.oracle().orc_dct3_inv(ob.p(out), ob.p(back))
    assert np.abs(back - x).max() < 1e-2           # src/mc_codec_float.h:264-273


def test_step_table_closed_form():
    st = np.zeros(4096, np.float32)
    rs = np.zeros(4096, np.float32)
    ob.oracle().orc_step_table(6.0, ob.p(st), ob.p(rs))
    i = np.arange(4096)
    l1 = (i >> 8) + ((i >> 4) & 15) + (i & 15)
    want = 6.0 * (1.0 + l1) ** 0.65
    assert np.allclose(st, want, rtol=1e-6)
    assert np.allclose(rs * st, 1.0, rtol=1e-6)


def test_scan_order_is_a_permutation_sorted_by_l1():
    sc = np.zeros(4096, np.uint16)
    ob.oracle().orc_scan_order(ob.p(sc))
    assert sorted(sc.tolist()) == list(range(4096))
    s = sc.astype(int)
    l1 = (s >> 8) + ((s >> 4) & 15) + (s & 15)
    assert (np.diff(l1) >= 0).all()
    assert s[0] == 0 and s[-1] == 4095


def test_range_coder_roundtrip_random_ops():
    O = ob.oracle()
    rng = np.random.default_rng(11)
    for trial in range(40):
        n = int(rng.integers(1, 3000))
        kinds = rng.integers(0, 3, n).astype(np.uint8)
        ctx = rng.integers(0, 64, n).astype(np.uint8)
        nb = rng.integers(0, 33, n).astype(np.uint8)
        vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        skew = rng.random(n) < 0.8
        vals = np.where((kinds == 0) & skew, 0, vals).astype(np.uint32)
        vals = np.where(kinds < 2, vals & 1, vals & ((np.uint64(1) << nb.astype(np.uint64)) - 1).astype(np.uint32))
        out = np.zeros(n * 8 + 64, np.uint8)
        ln = O.orc_rc_encode_ops(n, ob.p(kinds), ob.p(ctx), ob.p(nb), ob.p(vals), ob.p(out), out.size)
        got = np.zeros(n, np.uint32)
        O.orc_rc_decode_ops(n, ob.p(kinds), ob.p(ctx), ob.p(nb), ob.p(out), ln, ob.p(got))
        assert np.array_equal(got, vals), f"trial {trial}"


def test_fracmap_roundtrip():
    O = ob.oracle()
    rng = np.random.default_rng(2)
    for pattern in ["smooth", "random", "zeros", "full"]:
        if pattern == "smooth":
            fr = (np.clip(np.sin(np.arange(4096) / 200.0) * 10 + 7, 0, 15)).astype(np.uint8)
        elif pattern == "random":
            fr = rng.integers(0, 16, 4096).astype(np.uint8)
        elif pattern == "zeros":
            fr = np.zeros(4096, np.uint8)
        else:
            fr = np.full(4096, 15, np.uint8)
        buf = np.zeros(16384, np.uint8)
        n = O.orc_enc_fracmap(ob.p(fr), ob.p(buf), buf.size)
        back = np.zeros(4096, np.uint8)
        O.orc_dec_fracmap(ob.p(buf), n, ob.p(back))
        assert np.array_equal(back, fr), pattern


def test_oracle_matches_reference_build_on_random_blocks():
    if not ob.ref_available():
        pytest.skip("reference build not present (only in the build container)")
    import ctypes as C
    R = ob.ref()
    ctx = R.mc_codec_ctx_new()
    rng = np.random.default_rng(77)
    try:
        for trial in range(150):
            vox = rng.integers(0, 256, 4096).astype(np.uint8)
            if trial % 3 == 1:
                vox = np.where(rng.random(4096) < 0.4, 0, 128 + (vox >> 3)).astype(np.uint8)
            q = float(rng.choice([0.3, 0.5, 1, 2.5, 6, 12, 33, 64, 500, 2000]))
            tau = int(rng.choice([0, 1, 3, 16, 200]))
            R.mc_codec_ctx_set_quality(ctx, q)
            R.mc_codec_ctx_set_max_error(ctx, tau)
            b = ob._McBuf()
            n = C.c_uint32()
            R.mc_enc_block(ctx, ob.p(vox), C.byref(b), C.byref(n))
            want = C.string_at(b.p, n.value) if n.value else b""
            assert ob.enc_block(vox, q, tau) == want
            if want:
                corrupt = bytearray(want)
                corrupt[rng.integers(0, len(corrupt))] ^= 0x5A
                corrupt = bytes(corrupt[: max(1, len(corrupt) - int(rng.integers(0, 4)))])
                d_ref = np.zeros(4096, np.uint8)
                R.mc_codec_ctx_set_quality(ctx, q)
                R.mc_dec_block(ctx, corrupt, len(corrupt), ob.p(d_ref))
                assert np.array_equal(ob.dec_block(corrupt, q), d_ref)
    finally:
        R.mc_codec_ctx_free(ctx)


def test_oracle_chunk_blobs_match_manifest():
    import mc_amd
    path = os.path.join(ROOT, "tests", "golden", "chunk_manifest.json")
    if not os.path.exists(path):
        pytest.skip("chunk manifest not generated")
    man = json.load(open(path))
    picks = [("plain64_q6", 0), ("mixed256", 3), ("mixed256", 10)]
    for name, k in picks:
        e = man[name][k]
        chunk = mc_amd.synth_chunks([e["seed"]], [e["air"]])[0]
        assert f"{ob.xxh64(chunk):016x}" == e["input_xxh64"]
        blob = ob.encode_chunk(chunk, e["q"], e["tau"])
        assert len(blob) == e["blob_len"]
        assert f"{ob.xxh64(blob):016x}" == e["blob_xxh64"]
        assert ob.oracle().orc_blob_hash(ob.p(np.frombuffer(blob, np.uint8)), len(blob)) == \
            int.from_bytes(blob[4:12], "little")
        dec = ob.decode_chunk(blob)
        assert f"{ob.xxh64(dec):016x}" == e["decoded_xxh64"]
