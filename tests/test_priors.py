"""Per-volume prior overrides (mc_archive_set_priors, src/mc_archive.c:650-672)
and the lifetime of quality tables.

The reference fixture (tests/golden/priors_golden.json) holds what the
reference writes after mc_archive_set_priors: the endpoint tables, every
chunk blob's length and xxh64, the decoded chunk's xxh64 and the whole
archive file's xxh64.  The CPU test pins the oracle's override path to it;
the GPU tests check that this library writes the same archive bytes and
decodes the same voxels under the override, that a serving store keeps the
tables it was opened with when the process priors change afterwards, and
that more than 4096 distinct quality values never disturb tables already
handed out."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_bind as ob
from test_dropin_archive import _api, _blobs, _build_ours, _tmp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "priors_golden.json")


def _gold():
    with open(GOLD) as f:
        g = json.load(f)
    lo = np.frombuffer(bytes.fromhex(g["plo_u16le"]), "<u2").reshape(8, 32).astype(np.uint16)
    hi = np.frombuffer(bytes.fromhex(g["phi_u16le"]), "<u2").reshape(8, 32).astype(np.uint16)
    return g, lo, hi


def _inputs(g):
    import mc_amd
    return mc_amd.synth_chunks([c["seed"] for c in g["chunks"]], [c["air"] for c in g["chunks"]])


def test_fixture_tables_are_the_seeded_override():
    g, lo, hi = _gold()
    want_lo, want_hi = ob.custom_priors()
    assert np.array_equal(lo, want_lo) and np.array_equal(hi, want_hi)
    dlo, dhi = ob.default_priors()
    assert not np.array_equal(lo, dlo) and not np.array_equal(hi, dhi)


def test_oracle_under_override_matches_reference_fixture():
    g, lo, hi = _gold()
    chunks = _inputs(g)
    ob.set_priors(lo, hi)
    try:
        for k, c in enumerate(g["chunks"]):
            assert f"{ob.xxh64(chunks[k]):016x}" == c["input_xxh64"], f"input {k}"
            blob = ob.encode_chunk(chunks[k], c["q"], c["tau"])
            assert len(blob) == c["blob_len"] and f"{ob.xxh64(blob):016x}" == c["blob_xxh64"], f"blob {k}"
            assert f"{ob.xxh64(ob.decode_chunk(blob)):016x}" == c["decoded_xxh64"], f"decode {k}"
    finally:
        ob.set_priors()
    # the default tables give different bytes (the override is really in force above)
    blob = ob.encode_chunk(chunks[0], g["chunks"][0]["q"], 0)
    assert f"{ob.xxh64(blob):016x}" != g["chunks"][0]["blob_xxh64"]


def _restore_default_priors():
    _api().mc_codec_set_priors(None, None)


@pytest.mark.gpu
def test_gpu_archive_under_override_matches_reference():
    """append_chunk_par / append_chunk_ctx after mc_archive_set_priors write the
    reference's archive bytes; decode_chunk and a serving batch give the
    reference's voxels."""
    import mc_amd
    g, lo, hi = _gold()
    chunks = _inputs(g)
    L = _api()
    path = _tmp("priors.mca")
    a = L.mc_archive_open_dims(path.encode(), *g["dims"], C.c_float(g["quality"]))
    assert a
    print("set priors + appends", flush=True)
    try:
        assert L.mc_archive_set_priors(a, ob.p(lo), ob.p(hi)) == 0
        ctx = L.mc_codec_ctx_new()
        for k, c in enumerate(g["chunks"]):
            v = np.ascontiguousarray(chunks[k])
            if c["path"] == "par":
                assert L.mc_archive_append_chunk_par(a, 0, *c["coord"], ob.p(v), C.c_float(c["q"]), 0) == 0
            else:
                L.mc_apply_preset(ctx, c["preset"])
                assert L.mc_archive_append_chunk_ctx(a, ctx, 0, *c["coord"], ob.p(v)) == 0
        L.mc_codec_ctx_free(ctx)
        print("decode_chunk", flush=True)
        dec = np.empty((256, 256, 256), np.uint8)
        for k, c in enumerate(g["chunks"]):
            off = L.mc_archive_chunk_offset(a, 0, *c["coord"])
            assert off == c["chunk_off"], f"chunk {k} offset"
            L.mc_archive_decode_chunk(a, off, ob.p(dec), 0)
            assert f"{ob.xxh64(dec):016x}" == c["decoded_xxh64"], f"decode_chunk {k}"
        print("serving under the override", flush=True)
        ob.set_priors(lo, hi)
        want = [ob.decode_chunk(ob.encode_chunk(chunks[k], c["q"], c["tau"])) for k, c in enumerate(g["chunks"])]
        ob.set_priors()
        sv = mc_amd.Serve(a, lod=0, arena_bytes=64 << 20, lanes=2)
        boxes = [(c["coord"][0] * 256 + 17, c["coord"][1] * 256 + 33, c["coord"][2] * 256 + 9)
                 for c in g["chunks"]]
        d_out = mc_amd.DeviceBuffer(len(boxes) * 40 * 48 * 64)
        try:
            sv.wait(sv.submit(boxes, 40, 48, 64, d_out))
            got = np.empty((len(boxes), 40, 48, 64), np.uint8)
            mc_amd.Ctx(0).d2h(got, d_out)
        finally:
            sv.close()
            d_out.free()
        for k, (z, y, x) in enumerate(boxes):
            assert np.array_equal(got[k], want[k][17:57, 33:81, 9:73]), f"served box {k}"
    finally:
        L.mc_archive_close(a)
        _restore_default_priors()
    data = np.fromfile(path, np.uint8)
    assert data.size == g["archive_len"] and f"{ob.xxh64(data):016x}" == g["archive_xxh64"]
    for c in g["chunks"]:
        n = int(ob.oracle().orc_blob_len(ob.p(data[c["chunk_off"]:])))
        blob = data[c["chunk_off"]:c["chunk_off"] + n].tobytes()
        assert n == c["blob_len"] and f"{ob.xxh64(blob):016x}" == c["blob_xxh64"]
    # a reopen installs the archive's override as the process priors again
    a = L.mc_archive_open_dims(path.encode(), *g["dims"], C.c_float(g["quality"]))
    try:
        c = g["chunks"][1]
        L.mc_archive_decode_chunk(a, c["chunk_off"], ob.p(dec), 0)
        assert f"{ob.xxh64(dec):016x}" == c["decoded_xxh64"]
    finally:
        L.mc_archive_close(a)
        _restore_default_priors()


def _serve_check(sv, big_of, boxes, d):
    import mc_amd
    d_out = mc_amd.DeviceBuffer(len(boxes) * d ** 3)
    try:
        sv.wait(sv.submit(boxes, d, d, d, d_out))
        got = np.empty((len(boxes), d, d, d), np.uint8)
        mc_amd.Ctx(0).d2h(got, d_out)
    finally:
        d_out.free()
    for k, o in enumerate(boxes):
        assert np.array_equal(got[k], big_of(o, d)), f"box {k} {o}"


@pytest.mark.gpu
def test_store_keeps_its_tables_across_prior_changes_and_4097_qualities():
    """A server opened before the process priors change, and before more than
    4096 distinct qualities get tables, keeps decoding with its own tables."""
    import mc_amd
    blobs, _ = _blobs()
    L = _api()
    a = _build_ours(_tmp("keep.mca"), blobs)
    dec = [ob.decode_chunk(b) for b in blobs]

    def crop(o, d):          # the three chunks sit at (0,0,0), (0,0,1) and (0,1,1) of LOD 0 (_build_ours)
        out = np.zeros((d, d, d), np.uint8)
        for (cz, cy, cx), ch in zip([(0, 0, 0), (0, 0, 1)], dec[:2]):
            z0, y0, x0 = o[0] - cz * 256, o[1] - cy * 256, o[2] - cx * 256
            zs, ys, xs = max(0, -z0), max(0, -y0), max(0, -x0)
            ze, ye, xe = min(d, 256 - z0), min(d, 256 - y0), min(d, 256 - x0)
            if ze > zs and ye > ys and xe > xs:
                out[zs:ze, ys:ye, xs:xe] = ch[z0 + zs:z0 + ze, y0 + ys:y0 + ye, x0 + xs:x0 + xe]
        return out

    boxes = [(10, 20, 30), (100, 200, 250), (0, 128, 200), (200, 5, 300)]
    print("serve open (default priors)", flush=True)
    sv = mc_amd.Serve(a, lod=0, arena_bytes=2 << 20, lanes=2)   # 512 slots: later batches decode again
    try:
        _serve_check(sv, crop, boxes, 48)
        g, lo, hi = _gold()
        print("process priors changed", flush=True)
        L.mc_codec_set_priors(ob.p(lo), ob.p(hi))
        ctx = mc_amd.Ctx(0)
        pay = ob.enc_block(np.ascontiguousarray(_inputs(g)[0][64:80, 64:80, 64:80]).ravel(), 6.0)
        ob.set_priors(lo, hi)
        try:
            got = ctx.decode_blocks_host([pay, pay], [6.0, 7.25])
            assert np.array_equal(got[0], ob.dec_block(pay, 6.0))
            assert np.array_equal(got[1], ob.dec_block(pay, 7.25))
        finally:
            ob.set_priors()
        _serve_check(sv, crop, [(b[0] + 16, b[1], b[2]) for b in boxes], 48)
        _restore_default_priors()
        print("4097 distinct qualities", flush=True)
        qs = (1.0 + np.arange(4097, dtype=np.float32) / 64.0).astype(np.float32)
        got = ctx.decode_blocks_host([pay] * len(qs), qs)
        for k in range(0, len(qs), 97):
            assert np.array_equal(got[k], ob.dec_block(pay, float(qs[k]))), f"q {qs[k]}"
        assert np.array_equal(got[-1], ob.dec_block(pay, float(qs[-1])))
        _serve_check(sv, crop, [(b[0], b[1] + 16, b[2]) for b in boxes], 48)
        # a quality that had a table before the 4097 new ones still decodes right
        again = ctx.decode_blocks_host([pay], [6.0])
        assert np.array_equal(again[0], ob.dec_block(pay, 6.0))
        ctx.close()
    finally:
        sv.close()
        L.mc_archive_close(a)
        _restore_default_priors()
