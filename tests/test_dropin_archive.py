"""The drop-in mc_archive API: file format, index, install, metadata, priors,
blob hash, fraction map and sampler against the reference (CPU parts), and the
GPU-backed decode / append / region entry points against the oracle."""
import ctypes as C
import json
import os
import tempfile

import numpy as np
import pytest

import oracle_bind as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _api():
    import mc_amd
    return mc_amd.api()


def _chunks(seeds, air):
    import mc_amd
    return mc_amd.synth_chunks(seeds, air)


def _tmp(name):
    d = tempfile.mkdtemp(prefix="mcamd_")
    return os.path.join(d, name)


def _blobs():
    """Small set of oracle blobs (the oracle is pinned to the reference)."""
    ch = _chunks([5, 6, 7], [0.5, 0.8, 0.3])
    return [ob.encode_chunk(ch[0], 6.0), ob.encode_chunk(ch[1], 2.5, 8), ob.encode_chunk(ch[2], 16.0)], ch


PLACES = [(0, 0, 0, 0), (0, 0, 0, 1), (0, 1, 1, 0)]


def _build_ours(path, blobs, meta=b"scroll"):
    L = _api()
    a = L.mc_archive_open_dims(path.encode(), 512, 512, 256, C.c_float(6.0))
    assert a
    assert L.mc_archive_set_metadata(a, meta, len(meta)) == 0
    for (lod, cz, cy, cx), b in zip(PLACES, blobs):
        assert L.mc_archive_append_chunk_compressed(a, lod, cz, cy, cx, b, len(b)) == 0
    L.mc_archive_reserve_index(a, 1, 0, 0, 0)
    return a


def test_archive_file_identical_to_reference():
    if not ob.ref_available():
        pytest.skip("reference build not present")
    R = ob.ref()
    R.mc_archive_append_chunk_compressed.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                                     C.c_size_t]
    R.mc_archive_set_metadata.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    R.mc_archive_reserve_index.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
    blobs, _ = _blobs()
    p_ref, p_ours = _tmp("ref.mca"), _tmp("ours.mca")
    a = R.mc_archive_open_dims(p_ref.encode(), 512, 512, 256, C.c_float(6.0))
    R.mc_archive_set_metadata(a, b"scroll", 6)
    for (lod, cz, cy, cx), b in zip(PLACES, blobs):
        assert R.mc_archive_append_chunk_compressed(a, lod, cz, cy, cx, b, len(b)) == 0
    R.mc_archive_reserve_index(a, 1, 0, 0, 0)
    R.mc_archive_close(a)
    L = _api()
    a2 = _build_ours(p_ours, blobs)
    L.mc_archive_close(a2)
    assert open(p_ref, "rb").read() == open(p_ours, "rb").read()


def test_archive_reopen_index_metadata_hash():
    L = _api()
    blobs, _ = _blobs()
    path = _tmp("a.mca")
    a = _build_ours(path, blobs)
    offs = [L.mc_archive_chunk_offset(a, *pl) for pl in PLACES]
    assert all(o > 1 for o in offs)
    assert L.mc_archive_chunk_offset(a, 0, 1, 0, 0) == 0
    assert L.mc_archive_chunk_coverage(a, 0, 0, 0, 1) == 1
    assert L.mc_archive_chunk_coverage(a, 0, 1, 1, 1) == 0
    n = C.c_size_t()
    m = L.mc_archive_metadata(a, C.byref(n))
    assert C.string_at(m, n.value) == b"scroll"
    L.mc_archive_close(a)
    # reopen with the same dims works, other dims refused
    a = L.mc_archive_open_dims(path.encode(), 512, 512, 256, C.c_float(6.0))
    assert a and [L.mc_archive_chunk_offset(a, *pl) for pl in PLACES] == offs
    assert not L.mc_archive_open_dims(path.encode(), 256, 512, 256, C.c_float(6.0))
    L.mc_archive_close(a)
    # mc_verify_archive hashes on the device: tests/test_gpu_hash.py
    for b in blobs:
        arr = np.frombuffer(b, np.uint8)
        assert L.mc_chunk_compute_hash(ob.p(arr), arr.size) == int.from_bytes(b[4:12], "little")


def test_fracmap_and_sampler_match_reference():
    if not ob.ref_available():
        pytest.skip("reference build not present")
    L = _api()
    blobs, _ = _blobs()
    path = _tmp("s.mca")
    a = _build_ours(path, blobs)
    R = ob.ref()
    R.mc_archive_block_fraction.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
    R.mc_archive_block_fraction.restype = C.c_float
    L.mc_archive_close(a)
    ra = R.mc_archive_open_dims(path.encode(), 512, 512, 256, C.c_float(6.0))
    a = L.mc_archive_open_dims(path.encode(), 512, 512, 256, C.c_float(6.0))
    for bz, by, bx in [(0, 0, 0), (3, 7, 20), (15, 15, 31), (5, 20, 3), (1, 1, 1)]:
        assert L.mc_archive_block_fraction(a, 0, bz, by, bx) == R.mc_archive_block_fraction(ra, 0, bz, by, bx)
    for seed in [0, 1, 10000, 99]:
        b1 = (ob.p.__self__ if False else None)
        o1 = (C.c_long * (3 * 32))()
        o2 = (C.c_long * (3 * 32))()
        n1 = L.mc_archive_sample_boxes(a, 0, seed, 32, 64, 64, 64, C.c_float(0.5), o1)
        n2 = R.mc_archive_sample_boxes(ra, 0, seed, 32, 64, 64, 64, C.c_float(0.5), o2)
        assert n1 == n2 and list(o1)[:3 * n1] == list(o2)[:3 * n2]
    L.mc_archive_close(a)
    R.mc_archive_close(ra)


def test_presets_and_ctx():
    L = _api()
    cx = L.mc_codec_ctx_new()
    prev = 0.0
    for lv in range(8):
        assert L.mc_preset_tau(lv) == 1 << lv
        q = L.mc_apply_preset(cx, lv)
        assert q >= prev and L.mc_codec_ctx_get_quality(cx) == q and L.mc_codec_ctx_get_max_error(cx) == 1 << lv
        prev = q
    L.mc_codec_ctx_set_quality(cx, C.c_float(-1.0))
    assert L.mc_codec_ctx_get_quality(cx) == 8.0
    L.mc_codec_ctx_set_quality(cx, C.c_float(2000.0))
    assert L.mc_codec_ctx_get_quality(cx) == 8.0
    L.mc_codec_ctx_free(cx)


def test_host_fracmap_matches_oracle():
    L = _api()
    rng = np.random.default_rng(8)
    for fr in [rng.integers(0, 16, 4096).astype(np.uint8), np.clip(np.arange(4096) // 300, 0, 15).astype(np.uint8)]:
        b1 = np.zeros(8192, np.uint8)
        b2 = np.zeros(8192, np.uint8)
        n1 = L.mc_enc_fracmap(ob.p(fr), ob.p(b1), b1.size)
        n2 = ob.oracle().orc_enc_fracmap(ob.p(fr), ob.p(b2), b2.size)
        assert n1 == n2 and np.array_equal(b1[:n1], b2[:n2])
        back = np.zeros(4096, np.uint8)
        L.mc_dec_fracmap(ob.p(b1), n1, ob.p(back))
        assert np.array_equal(back, fr)


# ---------------------------------------------------------------- GPU paths
@pytest.mark.gpu
def test_gpu_append_par_and_decode_chunk():
    L = _api()
    man = json.load(open(os.path.join(ROOT, "tests", "golden", "chunk_manifest.json")))["plain64_q6"]
    ch = _chunks([man[0]["seed"], man[1]["seed"]], [man[0]["air"], man[1]["air"]])
    path = _tmp("par.mca")
    a = L.mc_archive_open_dims(path.encode(), 512, 256, 256, C.c_float(6.0))
    print("append_chunk_par x2 + all-air", flush=True)
    for k in range(2):
        assert L.mc_archive_append_chunk_par(a, 0, 0, 0, k, ob.p(ch[k]), C.c_float(6.0), 0) == 0
    air = np.zeros((256, 256, 256), np.uint8)
    assert L.mc_archive_append_chunk_par(a, 0, 1, 0, 0, ob.p(air), C.c_float(6.0), 0) == 0
    assert L.mc_archive_chunk_coverage(a, 0, 1, 0, 0) == 0          # all-air par append stays absent
    out = np.zeros((256, 256, 256), np.uint8)
    for k in range(2):
        co = L.mc_archive_chunk_offset(a, 0, 0, 0, k)
        L.mc_archive_decode_chunk(a, co, ob.p(out), 0)
        assert f"{ob.xxh64(out):016x}" == man[k]["decoded_xxh64"]
    L.mc_archive_decode_chunk(a, 0, ob.p(out), 0)
    assert not out.any()
    L.mc_archive_close(a)
    # blobs in the file are the reference's bytes
    data = open(path, "rb").read()
    arr = np.frombuffer(data, np.uint8)
    assert L.mc_verify_archive(ob.p(arr), arr.size, 0) == 0


@pytest.mark.gpu
def test_gpu_append_paths_byte_identical_to_compressed_install():
    L = _api()
    blobs, ch = _blobs()
    p1, p2 = _tmp("x.mca"), _tmp("y.mca")
    a = L.mc_archive_open_dims(p1.encode(), 512, 512, 256, C.c_float(6.0))
    b = L.mc_archive_open_dims(p2.encode(), 512, 512, 256, C.c_float(6.0))
    ctx = L.mc_codec_ctx_new()
    L.mc_codec_ctx_set_quality(ctx, C.c_float(2.5))
    L.mc_codec_ctx_set_max_error(ctx, 8)
    print("append raw_q / ctx / par", flush=True)
    assert L.mc_archive_append_chunk_raw_q(a, 0, 0, 0, 0, ob.p(ch[0]), C.c_float(6.0)) == 0
    assert L.mc_archive_append_chunk_ctx(a, ctx, 0, 0, 0, 1, ob.p(ch[1])) == 0
    assert L.mc_archive_append_chunk_par(a, 0, 0, 1, 1, ob.p(ch[2]), C.c_float(16.0), 0) == 0
    co_par = L.mc_archive_chunk_offset(a, 0, 0, 1, 1)
    for (lod, cz, cy, cx), bl in zip(PLACES[:2], blobs[:2]):
        assert L.mc_archive_append_chunk_compressed(b, lod, cz, cy, cx, bl, len(bl)) == 0
    L.mc_codec_ctx_free(ctx)
    co_a = [L.mc_archive_chunk_offset(a, *pl) for pl in PLACES[:2]]
    co_b = [L.mc_archive_chunk_offset(b, *pl) for pl in PLACES[:2]]
    assert co_a == co_b
    L.mc_archive_close(a)
    L.mc_archive_close(b)
    da, db = open(p1, "rb").read(), open(p2, "rb").read()
    for ca, bl in zip(co_a, blobs[:2]):
        assert da[ca:ca + len(bl)] == bl
    assert da[co_par:co_par + len(blobs[2])] == blobs[2]


@pytest.mark.gpu
def test_gpu_decode_block_and_single_block_codec(kat):
    L = _api()
    cx = L.mc_codec_ctx_new()
    print("mc_enc_block / mc_dec_block on KAT cases", flush=True)
    for i in range(0, len(kat["q"]), 7):
        L.mc_codec_ctx_set_quality(cx, C.c_float(float(kat["q"][i])))
        L.mc_codec_ctx_set_max_error(cx, int(kat["tau"][i]))
        buf = __import__("mc_amd").McBuf()
        n = C.c_uint32()
        vox = np.ascontiguousarray(kat["vox"][i])
        coded = L.mc_enc_block(cx, ob.p(vox), C.byref(buf), C.byref(n))
        want = kat["payload"][kat["offs"][i]:kat["offs"][i + 1]].tobytes()
        got = C.string_at(buf.p, n.value) if n.value else b""
        assert coded == (1 if want else 0) and got == want, f"case {i}"
        if want:
            dst = np.zeros(4096, np.uint8)
            L.mc_dec_block(cx, want, len(want), ob.p(dst))
            assert np.array_equal(dst, kat["dec"][i])
    L.mc_codec_ctx_free(cx)


@pytest.mark.gpu
def test_gpu_read_regions_match_oracle_crops():
    L = _api()
    import mc_amd
    blobs, ch = _blobs()
    path = _tmp("r.mca")
    a = _build_ours(path, blobs)
    vol = np.zeros((512, 512, 512), np.uint8)         # z, y, x (dims 256 x 512 x 512 padded)
    for (lod, cz, cy, cx), bl in zip(PLACES, blobs):
        vol[cz * 256:(cz + 1) * 256, cy * 256:(cy + 1) * 256, cx * 256:(cx + 1) * 256] = ob.decode_chunk(bl)
    boxes = [(0, 0, 0), (10, 20, 30), (60, 100, 3), (64, 64, 300), (200, 250, 250), (250, 500, 490),
             (-5, -7, 100), (240, 240, 240)]
    dz, dy, dx = 40, 48, 64
    arr = (mc_amd.McBox * len(boxes))(*[mc_amd.McBox(*b) for b in boxes])
    out = np.zeros((len(boxes), dz, dy, dx), np.uint8)
    print("read_regions", flush=True)
    L.mc_archive_read_regions(a, 0, arr, len(boxes), dz, dy, dx, ob.p(out), out[0].nbytes, 0)
    big = np.zeros((1024, 1024, 1024), np.uint8)
    big[256:768, 256:768, 256:768] = vol
    for k, (z0, y0, x0) in enumerate(boxes):
        want = big[256 + z0:256 + z0 + dz, 256 + y0:256 + y0 + dy, 256 + x0:256 + x0 + dx]
        assert np.array_equal(out[k], want), f"box {k} {boxes[k]}"
    # single region with strides
    reg = np.zeros((20, 30, 50), np.uint8)
    L.mc_archive_read_region(a, 0, 100, 200, 250, 20, 30, 50, ob.p(reg), 30 * 50, 50, 0)
    assert np.array_equal(reg, vol[100:120, 200:230, 250:300])
    blk = np.zeros(4096, np.uint8)
    co = L.mc_archive_chunk_offset(a, 0, 0, 0, 1)
    L.mc_archive_decode_block(a, co, 3, 4, 5, ob.p(blk))
    assert np.array_equal(blk.reshape(16, 16, 16), vol[48:64, 64:80, 256 + 80:256 + 96])
    L.mc_archive_close(a)


def test_concurrent_appends_share_index_nodes():
    """tests/mc_archive_concurrent_test.c: 16 threads append disjoint chunks
    whose index paths share root, inner and shard nodes, released together by
    a barrier; every appended chunk must be reachable afterwards, on each of a
    number of fresh archives."""
    import threading
    L = _api()
    coords = [0, 16, 48, 80, 272, 288, 528, 800]          # index nibbles vary at every level
    dim = (800 + 1) * 256
    blob = np.full(64, 0xAB, np.uint8)                    # the index does not parse the blob
    places = [(z, y, x) for z in coords for y in coords for x in coords]
    nthread = 16
    for trial in range(12):
        a = L.mc_archive_open_dims(_tmp(f"c{trial}.mca").encode(), dim, dim, dim, C.c_float(6.0))
        assert a
        start = threading.Barrier(nthread)
        fails = [0] * nthread

        def worker(t):
            start.wait()
            for k in range(t, len(places), nthread):
                if L.mc_archive_append_chunk_compressed(a, 0, *places[k], ob.p(blob), blob.size) != 0:
                    fails[t] += 1

        th = [threading.Thread(target=worker, args=(t,)) for t in range(nthread)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        missing = sum(L.mc_archive_chunk_offset(a, 0, *p) <= 1 for p in places)
        L.mc_archive_close(a)
        assert sum(fails) == 0 and missing == 0, f"trial {trial}: {sum(fails)} failed, {missing} missing"
