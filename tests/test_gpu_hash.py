"""Device XXH64 and chunk-blob verification (mchip_xxh64_ranges,
mchip_verify_blobs[_host], mc_verify_archive): plain XXH64 of ranges of every
length class and alignment against the published algorithm (the xxhash
package, or the oracle where it is absent; the oracle is pinned to xxhash's
known answers), blob hashes against the hash stored by the reference's
encoder, tamper and truncation detection, and mc_verify_archive's count and
messages against the reference build on clean and corrupted archives."""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob
from test_dropin_archive import PLACES, _api, _blobs, _build_ours, _tmp

pytestmark = pytest.mark.gpu


def _xxh64(data, seed):
    try:
        import xxhash
        return xxhash.xxh64(bytes(data), seed=seed).intdigest()
    except ImportError:
        return ob.xxh64(np.frombuffer(bytes(data), np.uint8), seed)


@pytest.fixture(scope="module")
def ctx():
    import mc_amd
    return mc_amd.Ctx(0)


# every path of the kernel: empty, tail only (< 32), one stripe, partial and
# full 256-stripe tiles, several tiles, and large ranges
LENGTHS = [0, 1, 3, 4, 7, 8, 9, 31, 32, 33, 63, 64, 255, 256, 257, 1000, 8191, 8192, 8193, 8192 + 31,
           8192 * 2 + 32 * 7 + 5, 65536, 100003, 1 << 20, 3 * (1 << 20) + 17]


def test_xxh64_ranges_all_lengths_and_alignments(ctx):
    import mc_amd
    rng = np.random.default_rng(21)
    pad = 64
    total = sum(LENGTHS) * 4 + pad * 4 * len(LENGTHS) + 64
    data = rng.integers(0, 256, total, dtype=np.uint8)
    buf = mc_amd.DeviceBuffer(total)
    try:
        ctx.h2d(buf, data)
        offs, lens, seeds = [], [], []
        pos = 0
        for n in LENGTHS:
            for al in range(4):
                pos = (pos + pad) // 4 * 4 + al
                offs.append(pos)
                lens.append(n)
                seeds.append(int(rng.integers(0, 1 << 63)) * (1 + (al & 1)))
                pos += n
        seeds[0] = 0
        print(f"xxh64 ranges n={len(offs)} bytes={pos}", flush=True)
        got = ctx.xxh64_ranges(buf, offs, lens, seeds)
    finally:
        buf.free()
    for o, n, sd, g in zip(offs, lens, seeds, got):
        want = _xxh64(data[o:o + n].tobytes(), sd)
        assert int(g) == want, f"length {n} at offset {o} (alignment {o & 3}) seed {sd:#x}"
    assert int(got[0]) == 0xEF46DB3751D8E999                      # XXH64("", 0)


def _pack(blobs, shift=0):
    offs, pos = [], 0
    for b in blobs:
        pos = (pos + 15) // 16 * 16 + shift
        offs.append(pos)
        pos += len(b)
    host = np.zeros(pos + 16, np.uint8)
    for o, b in zip(offs, blobs):
        host[o:o + len(b)] = np.frombuffer(b, np.uint8)
    return host, offs


def test_verify_blobs_device_and_host(ctx):
    import mc_amd
    blobs, _ = _blobs()
    stored = [int.from_bytes(b[4:12], "little") for b in blobs]
    # tampered copies: a payload byte, a length-table byte, a fraction-map byte
    t1 = bytearray(blobs[0]); t1[len(t1) - 100] ^= 0x01
    fml = int.from_bytes(blobs[1][12:14], "little")
    t2 = bytearray(blobs[1]); t2[14 + fml + 512 + 3] ^= 0x80
    t3 = bytearray(blobs[2]); t3[20] ^= 0x10
    trunc = blobs[0][:14 + 100]                                  # structure does not fit
    cases = blobs + [bytes(t1), bytes(t2), bytes(t3), trunc]
    want_bad = [0, 0, 0, 1, 1, 1, 1]
    for shift in (0, 1, 3):
        host, offs = _pack(cases, shift)
        buf = mc_amd.DeviceBuffer(host.size)
        try:
            ctx.h2d(buf, host)
            print(f"verify_blobs shift={shift}", flush=True)
            n, h, bad = ctx.verify_blobs(buf, offs, [len(b) for b in cases])
        finally:
            buf.free()
        assert n == 4 and list(bad) == want_bad, f"shift {shift}"
        assert [int(x) for x in h[:3]] == stored
        assert int(h[6]) == 0
    print("verify_blobs_host", flush=True)
    n, h, bad = ctx.verify_blobs_host([np.frombuffer(b, np.uint8) for b in cases])
    assert n == 4 and list(bad) == want_bad
    assert [int(x) for x in h[:3]] == stored
    # the host helper agrees on every structurally valid blob
    L = _api()
    for b, x in zip(cases[:5], h[:5]):
        a = np.frombuffer(b, np.uint8)
        assert L.mc_chunk_compute_hash(ob.p(a), a.size) == int(x)


def test_verify_blobs_host_many_blobs_fill_both_slabs(ctx):
    """About 600 MiB of blobs: three slabs, so each pinned slab is refilled
    while the other one's hashes are still running."""
    blobs, _ = _blobs()
    big = np.frombuffer(blobs[0], np.uint8)
    reps = (600 << 20) // big.size + 1
    arrs = [big] * reps
    bad_at = {7, reps // 2, reps - 1}
    tam = big.copy()
    tam[-1] ^= 0xFF
    for k in bad_at:
        arrs[k] = tam
    print(f"verify_blobs_host reps={reps}", flush=True)
    n, h, bad = ctx.verify_blobs_host(arrs)
    assert n == len(bad_at)
    assert set(np.nonzero(bad)[0].tolist()) == bad_at
    good = int.from_bytes(blobs[0][4:12], "little")
    assert all(int(h[k]) == good for k in range(reps) if k not in bad_at)


def _archive_bytes():
    blobs, _ = _blobs()
    path = _tmp("v.mca")
    a = _build_ours(path, blobs)
    offs = [_api().mc_archive_chunk_offset(a, *pl) for pl in PLACES]
    _api().mc_archive_close(a)
    return np.fromfile(path, np.uint8), offs


def _ref_verify(data, verbose=0):
    R = ob.ref()
    R.mc_verify_archive.argtypes = [C.c_void_p, C.c_size_t, C.c_int]
    R.mc_verify_archive.restype = C.c_long
    return R.mc_verify_archive(ob.p(data), data.size, verbose)


def test_verify_archive_clean_and_tampered():
    L = _api()
    data, offs = _archive_bytes()
    print("verify clean", flush=True)
    assert L.mc_verify_archive(ob.p(data), data.size, 0) == 0
    data[offs[0] + 2000] ^= 0x55
    print("verify tampered", flush=True)
    assert L.mc_verify_archive(ob.p(data), data.size, 0) == 1
    data[offs[2] + 14] ^= 0x01                                 # fraction map of another chunk
    assert L.mc_verify_archive(ob.p(data), data.size, 0) == 2
    assert L.mc_verify_archive(ob.p(data), 10, 0) == -1


def test_verify_archive_matches_reference_on_corruptions(capfd):
    """Random corruptions of the index nodes, blob headers, length tables and
    payloads: the count and the stderr report equal the reference build's."""
    if not ob.ref_available():
        pytest.skip("reference build not present")
    L = _api()
    clean, offs = _archive_bytes()
    rng = np.random.default_rng(5)
    hdr_end = 4096
    for trial in range(24):
        data = clean.copy()
        for _ in range(int(rng.integers(1, 4))):
            kind = trial % 4
            if kind == 0:                                      # anywhere in the file
                at = int(rng.integers(0, data.size))
            elif kind == 1:                                    # blob headers and tables
                at = offs[int(rng.integers(0, 3))] + int(rng.integers(0, 1200))
            elif kind == 2:                                    # index region before the first blob
                at = int(rng.integers(hdr_end, min(offs)))
            else:                                              # payloads
                o = offs[int(rng.integers(0, 3))]
                at = o + int(rng.integers(600, 20000))
            data[min(at, data.size - 1)] ^= np.uint8(1 << int(rng.integers(0, 8)))
        print(f"corruption trial {trial}", flush=True)
        capfd.readouterr()
        ours = L.mc_verify_archive(ob.p(data), data.size, 1)
        err_ours = capfd.readouterr().err
        ref = _ref_verify(data, 1)
        err_ref = capfd.readouterr().err
        assert ours == ref, f"trial {trial}"
        assert err_ours == err_ref, f"trial {trial}"
