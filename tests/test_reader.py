"""The archive readers (mc_open / mc_open_streaming, whole-blob and partial
fetch) and rate-controlled append against the reference compiled from its
sources (oracle/_ref/libmcref.so, which holds src/mc_stream.c and
src/mc_archive.c).

CPU: chunk resolution (present, air and absent chunks, failing reads), blob
lengths and bytes, header fields, priors, through flat and streaming readers;
the archive and reader error paths of tests/mc_archive_errpaths_test.c.
GPU: block decodes through every reader mode equal the reference reader's
(including a corrupted archive), the batched decode equals per-block decodes,
partial fetch moves far fewer bytes for scattered blocks
(tests/mc_stream_partial.c), and mc_archive_append_chunk_target picks the
reference's q and installs its bytes."""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob
from test_dropin_archive import PLACES, _api, _blobs, _build_ours, _tmp

needs_ref = pytest.mark.skipif(not ob.ref_available(), reason="reference build not present")

READ_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p)


class Source:
    """A byte-range source over an in-memory archive that counts what it serves."""

    def __init__(self, data, fail_after=None):
        self.data = np.ascontiguousarray(data)
        self.bytes = 0
        self.calls = 0
        self.fail_after = fail_after
        self.fn = READ_FN(self._read)

    def _read(self, ud, off, n, dst):
        if self.fail_after is not None and self.calls >= self.fail_after:
            return -1
        if off + n > self.data.size:
            return -1
        C.memmove(dst, self.data.ctypes.data + off, n)
        self.bytes += n
        self.calls += 1
        return 0


def _declare(L):
    vp = C.c_void_p
    L.mc_open.restype = vp
    L.mc_open.argtypes = [vp, C.c_size_t]
    L.mc_open_streaming.restype = vp
    L.mc_open_streaming.argtypes = [READ_FN, vp, C.c_uint64]
    L.mc_close.argtypes = [vp]
    L.mc_reader_set_partial_fetch.argtypes = [vp, C.c_int]
    L.mc_reader_set_quality.argtypes = [vp, C.c_float]
    L.mc_chunk_offset.restype = C.c_uint64
    L.mc_chunk_offset.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.mc_chunk_offset_chk.restype = C.c_uint64
    L.mc_chunk_offset_chk.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]
    L.mc_decode_block.argtypes = [vp, C.c_uint64, C.c_int, C.c_int, C.c_int, vp]
    L.mc_reader_chunk_blob_len.restype = C.c_uint64
    L.mc_reader_chunk_blob_len.argtypes = [vp, C.c_uint64]
    L.mc_reader_read_blob.argtypes = [vp, C.c_uint64, C.c_size_t, vp]
    L.mc_reader_priors.argtypes = [vp, C.POINTER(vp), C.POINTER(vp)]
    L.mc_reader_dims.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.mc_reader_quality.restype = C.c_float
    L.mc_reader_quality.argtypes = [vp]
    L.mc_reader_nlods.argtypes = [vp]
    return L


def _ours():
    L = _declare(_api())
    L.mc_reader_decode_blocks.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def _ref():
    return _declare(ob.ref())


@pytest.fixture(scope="module")
def archive():
    """Three present chunks and a chunk marked air (slot 1), written by the
    reference's own writer (CPU only)."""
    if not ob.ref_available():
        pytest.skip("reference build not present")
    R = ob.ref()
    R.mc_archive_append_chunk_compressed.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                                     C.c_size_t]
    R.mc_archive_append_chunk_raw.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    blobs, _ = _blobs()
    path = _tmp("r.mca")
    a = R.mc_archive_open_dims(path.encode(), 512, 512, 256, C.c_float(6.0))
    for (lod, cz, cy, cx), b in zip(PLACES, blobs):
        assert R.mc_archive_append_chunk_compressed(a, lod, cz, cy, cx, b, len(b)) == 0
    air = np.zeros((256, 256, 256), np.uint8)
    assert R.mc_archive_append_chunk_raw(a, 0, 1, 1, 1, ob.p(air)) == 0
    R.mc_archive_close(a)
    return np.fromfile(path, np.uint8)


def _readers(L, data, sources):
    flat = L.mc_open(ob.p(data), data.size)
    s_full, s_part = Source(data), Source(data)
    sources += [s_full, s_part]
    full = L.mc_open_streaming(s_full.fn, None, data.size)
    part = L.mc_open_streaming(s_part.fn, None, data.size)
    L.mc_reader_set_partial_fetch(part, 1)
    assert flat and full and part
    return [flat, full, part]


PROBES = [(l, z, y, x) for l in (0, 1) for z in range(3) for y in range(3) for x in range(3)]


@needs_ref
def test_reader_index_and_accessors_match_reference(archive):
    ours, ref = _ours(), _ref()
    keep = []
    ro, rr = _readers(ours, archive, keep), _readers(ref, archive, keep)
    try:
        for a, b in zip(ro, rr):
            for pr in PROBES:
                assert ours.mc_chunk_offset(a, *pr) == ref.mc_chunk_offset(b, *pr), pr
            dims = [[C.c_int() for _ in range(3)] for _ in range(2)]
            ours.mc_reader_dims(a, *[C.byref(v) for v in dims[0]])
            ref.mc_reader_dims(b, *[C.byref(v) for v in dims[1]])
            assert [v.value for v in dims[0]] == [v.value for v in dims[1]] == [512, 512, 256]
            assert ours.mc_reader_quality(a) == ref.mc_reader_quality(b)
            assert ours.mc_reader_nlods(a) == ref.mc_reader_nlods(b) == 1
            assert ours.mc_reader_priors(a, None, None) == ref.mc_reader_priors(b, None, None)
            for pl in PLACES:
                co = ours.mc_chunk_offset(a, *pl)
                n = ours.mc_reader_chunk_blob_len(a, co)
                assert n == ref.mc_reader_chunk_blob_len(b, co) and n > 0
                x, y = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
                assert ours.mc_reader_read_blob(a, co, n, ob.p(x)) == 0
                assert ref.mc_reader_read_blob(b, co, n, ob.p(y)) == 0
                assert np.array_equal(x, y) and np.array_equal(x, archive[co:co + n])
        # a flat read past the archive's end is refused (the reference would
        # read out of bounds here, so it is not called)
        z = np.zeros(64, np.uint8)
        assert ours.mc_reader_read_blob(ro[0], archive.size - 32, 64, ob.p(z)) == -1
        assert ours.mc_reader_read_blob(ro[0], archive.size - 64, 64, ob.p(z)) == 0
        assert np.array_equal(z, archive[-64:])
    finally:
        for a in ro:
            ours.mc_close(a)
        for b in rr:
            ref.mc_close(b)
    # a streaming read that fails: the resolve reports it instead of "absent"
    for L in (ours, ref):
        src = Source(archive)
        r = L.mc_open_streaming(src.fn, None, archive.size)
        src.fail_after = src.calls
        err = C.c_int(0)
        assert L.mc_chunk_offset_chk(r, 0, 0, 0, 0, C.byref(err)) == 0 and err.value == 1
        L.mc_close(r)
    # garbage is refused
    junk = np.zeros(300, np.uint8)
    assert not ours.mc_open(ob.p(junk), junk.size) and not ref.mc_open(ob.p(junk), junk.size)
    assert not ours.mc_open(ob.p(junk), 100)


def test_archive_and_reader_error_paths():
    """The error paths of the reference's tests/mc_archive_errpaths_test.c on
    this library: a reopen with other dims is refused and leaves the file
    usable; a file or buffer without the archive magic, or shorter than the
    header, is refused by the writer and the flat reader; priors stored with
    mc_archive_set_priors come back through a flat reader."""
    L = _ours()
    blobs, _ = _blobs()
    path = _tmp("e.mca")
    a = _build_ours(path, blobs)
    offs = [L.mc_archive_chunk_offset(a, *pl) for pl in PLACES]
    L.mc_archive_close(a)
    assert not L.mc_archive_open_dims(path.encode(), 128, 128, 128, C.c_float(6.0))
    a = L.mc_archive_open_dims(path.encode(), 512, 512, 256, C.c_float(6.0))
    assert a and [L.mc_archive_chunk_offset(a, *pl) for pl in PLACES] == offs
    plo = np.full((8, 32), 900, np.uint16)
    phi = np.full((8, 32), 3100, np.uint16)
    try:
        assert L.mc_archive_set_priors(a, ob.p(plo), ob.p(phi)) == 0
        L.mc_archive_close(a)
        data = np.fromfile(path, np.uint8)
        r = L.mc_open(ob.p(data), data.size)
        lo, hi = C.c_void_p(), C.c_void_p()
        assert r and L.mc_reader_priors(r, C.byref(lo), C.byref(hi)) == 1
        got_lo = np.ctypeslib.as_array(C.cast(lo, C.POINTER(C.c_uint16)), (8 * 32,))
        got_hi = np.ctypeslib.as_array(C.cast(hi, C.POINTER(C.c_uint16)), (8 * 32,))
        assert np.array_equal(got_lo, plo.ravel()) and np.array_equal(got_hi, phi.ravel())
        L.mc_close(r)
    finally:
        L.mc_codec_set_priors(None, None)          # priors are process-wide: back to the defaults
    garbage = _tmp("g.mca")
    with open(garbage, "wb") as f:
        f.write(b"\xab" * 300)
    assert not L.mc_archive_open_dims(garbage.encode(), 256, 256, 256, C.c_float(6.0))
    junk = np.full(300, 0xAB, np.uint8)
    assert not L.mc_open(ob.p(junk), junk.size)
    assert not L.mc_open(ob.p(data), 255) and not L.mc_open(None, 4096)


BLOCKS = [(z, y, x) for z in (0, 3, 8, 15) for y in (0, 7, 15) for x in (0, 5, 12, 15)]


@needs_ref
@pytest.mark.gpu
def test_reader_block_decode_matches_reference_every_mode(archive):
    ours, ref = _ours(), _ref()
    keep = []
    ro, rr = _readers(ours, archive, keep), _readers(ref, archive, keep)
    offs = [ours.mc_chunk_offset(ro[0], *pl) for pl in PLACES] + [0, 1, ours.mc_chunk_offset(ro[0], 0, 1, 1, 1)]
    try:
        for mode, (a, b) in enumerate(zip(ro, rr)):
            print(f"reader mode {mode}", flush=True)
            got, want = np.zeros(4096, np.uint8), np.zeros(4096, np.uint8)
            for co in offs:
                for bz, by, bx in BLOCKS:
                    ours.mc_decode_block(a, co, bz, by, bx, ob.p(got))
                    ref.mc_decode_block(b, co, bz, by, bx, ob.p(want))
                    assert np.array_equal(got, want), f"mode {mode} chunk {co} block {bz},{by},{bx}"
            # batched decode equals the per-block decodes
            cos = np.array([co for co in offs for _ in BLOCKS], np.uint64)
            zyx = np.array([b3 for _ in offs for b3 in BLOCKS], np.int32).reshape(-1)
            out = np.zeros((cos.size, 4096), np.uint8)
            assert ours.mc_reader_decode_blocks(a, cos.size, ob.p(cos), ob.p(zyx), ob.p(out)) == 0
            for k in range(cos.size):
                ref.mc_decode_block(b, int(cos[k]), *[int(v) for v in zyx[3 * k:3 * k + 3]], ob.p(want))
                assert np.array_equal(out[k], want), f"batch mode {mode} entry {k}"
    finally:
        for a in ro:
            ours.mc_close(a)
        for b in rr:
            ref.mc_close(b)


@needs_ref
@pytest.mark.gpu
def test_partial_fetch_moves_fewer_bytes_and_agrees():
    """tests/mc_stream_partial.c: 5 scattered blocks through a whole-blob and a
    partial-fetch streaming reader give the same voxels, and partial fetch
    moves at least 4x fewer bytes; then a full sweep of the chunk agrees."""
    ours = _ours()
    blobs, _ = _blobs()
    path = _tmp("p.mca")
    a = _build_ours(path, blobs)
    _api().mc_archive_close(a)
    data = np.fromfile(path, np.uint8)
    sf, sp = Source(data), Source(data)
    rf = ours.mc_open_streaming(sf.fn, None, data.size)
    rp = ours.mc_open_streaming(sp.fn, None, data.size)
    ours.mc_reader_set_partial_fetch(rp, 1)
    co = ours.mc_chunk_offset(rf, 0, 0, 0, 0)
    assert co and co == ours.mc_chunk_offset(rp, 0, 0, 0, 0)
    b0f, b0p = sf.bytes, sp.bytes
    x, y = np.zeros(4096, np.uint8), np.zeros(4096, np.uint8)
    print("scattered blocks", flush=True)
    for pick in [(8, 8, 8), (0, 0, 0), (15, 15, 15), (3, 12, 7), (12, 3, 9)]:
        ours.mc_decode_block(rf, co, *pick, ob.p(x))
        ours.mc_decode_block(rp, co, *pick, ob.p(y))
        assert np.array_equal(x, y), pick
    full, part = sf.bytes - b0f, sp.bytes - b0p
    assert part * 4 <= full, (full, part)
    print("sweep", flush=True)
    zyx = np.array([(z, yy, xx) for z in range(16) for yy in range(16) for xx in range(16)], np.int32).reshape(-1)
    cos = np.full(4096, co, np.uint64)
    of, op = np.zeros((4096, 4096), np.uint8), np.zeros((4096, 4096), np.uint8)
    assert ours.mc_reader_decode_blocks(rf, 4096, ob.p(cos), ob.p(zyx), ob.p(of)) == 0
    assert ours.mc_reader_decode_blocks(rp, 4096, ob.p(cos), ob.p(zyx), ob.p(op)) == 0
    assert np.array_equal(of, op)
    want = ob.decode_chunk(blobs[0]).reshape(16, 16, 16, 16, 16, 16).transpose(0, 2, 4, 1, 3, 5).reshape(4096, 4096)
    assert np.array_equal(of, want)
    ours.mc_close(rf)
    ours.mc_close(rp)


@needs_ref
@pytest.mark.gpu
def test_reader_on_corrupted_archive_matches_reference(archive):
    ours, ref = _ours(), _ref()
    rng = np.random.default_rng(9)
    got, want = np.zeros(4096, np.uint8), np.zeros(4096, np.uint8)
    base = _ours().mc_open(ob.p(archive), archive.size)
    offs = [ours.mc_chunk_offset(base, *pl) for pl in PLACES]
    ours.mc_close(base)
    for trial in range(6):
        data = archive.copy()
        for _ in range(8):
            co = offs[int(rng.integers(0, 3))]
            data[co + int(rng.integers(0, 2000))] ^= np.uint8(1 << int(rng.integers(0, 8)))
        keep = []
        ro, rr = _readers(ours, data, keep), _readers(ref, data, keep)
        print(f"corrupted archive {trial}", flush=True)
        for mode, (a, b) in enumerate(zip(ro, rr)):
            for co in offs:
                for bz, by, bx in BLOCKS[::3]:
                    ours.mc_decode_block(a, co, bz, by, bx, ob.p(got))
                    ref.mc_decode_block(b, co, bz, by, bx, ob.p(want))
                    assert np.array_equal(got, want), f"trial {trial} mode {mode} chunk {co} {bz},{by},{bx}"
        for a in ro:
            ours.mc_close(a)
        for b in rr:
            ref.mc_close(b)


@needs_ref
@pytest.mark.gpu
def test_append_chunk_target_matches_reference():
    L = _api()
    L.mc_archive_append_chunk_target.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                                 C.c_float, C.POINTER(C.c_float)]
    R = ob.ref()
    R.mc_archive_append_chunk_target.argtypes = L.mc_archive_append_chunk_target.argtypes
    import mc_amd
    chunks = mc_amd.synth_chunks([41, 42, 43], [0.2, 0.6, 0.95])
    for k, ratio in enumerate([4.0, 12.0, 40.0]):
        p1, p2 = _tmp("t1.mca"), _tmp("t2.mca")
        a = L.mc_archive_open_dims(p1.encode(), 256, 256, 256, C.c_float(6.0))
        b = R.mc_archive_open_dims(p2.encode(), 256, 256, 256, C.c_float(6.0))
        qa, qb = C.c_float(), C.c_float()
        print(f"rate control chunk {k} ratio {ratio}", flush=True)
        assert L.mc_archive_append_chunk_target(a, 0, 0, 0, 0, ob.p(chunks[k]), ratio, C.byref(qa)) == 0
        assert R.mc_archive_append_chunk_target(b, 0, 0, 0, 0, ob.p(chunks[k]), ratio, C.byref(qb)) == 0
        assert qa.value == qb.value, (k, qa.value, qb.value)
        L.mc_archive_close(a)
        R.mc_archive_close(b)
        assert open(p1, "rb").read() == open(p2, "rb").read(), f"chunk {k}"
    a = L.mc_archive_open_dims(_tmp("t3.mca").encode(), 256, 256, 256, C.c_float(6.0))
    assert L.mc_archive_append_chunk_target(a, 0, 0, 0, 0, ob.p(chunks[0]), 1.0, None) == -1
    L.mc_archive_close(a)
