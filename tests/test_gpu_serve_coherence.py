"""Serving coherence and arena memory (mc_serve_refresh / _invalidate_chunk /
_resize, mchip_server_resize / _update_blocks): chunks appended or appended
again after mc_serve_open are served with their new voxels, as the
reference's cache sees them (src/mc_cache.c:690-697, :722-736); the arena
shrinks and grows while the server runs, grows by itself as requests need
slots, and a request that cannot get memory fails cleanly while later ones
succeed.  Voxels are checked against the oracle's decode of the same blobs
and against mc_archive_read_regions."""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob
from test_dropin_archive import _api, _tmp

E = 48


def _oracle_vol(blobs_at, nchunks_x):
    """Oracle decode of LOD-0 chunks (0, 0, x) -> blob, as one 256 x 256 x 256*n raster."""
    vol = np.zeros((256, 256, 256 * nchunks_x), np.uint8)
    for x, b in blobs_at.items():
        if b:
            vol[:, :, 256 * x:256 * (x + 1)] = ob.decode_chunk(b).reshape(256, 256, 256)
    return vol


def _serve(sv, ctx, boxes):
    import mc_amd
    d_out = mc_amd.DeviceBuffer(len(boxes) * E ** 3)
    try:
        sv.wait(sv.submit(boxes, E, E, E, d_out))
        got = np.empty((len(boxes), E, E, E), np.uint8)
        ctx.d2h(got, d_out)
    finally:
        d_out.free()
    return got


def _check(got, vol, boxes, what):
    for j, (z, y, x) in enumerate(boxes):
        want = np.zeros((E, E, E), np.uint8)
        zs, ys, xs = max(0, -z), max(0, -y), max(0, -x)
        ze, ye, xe = min(E, vol.shape[0] - z), min(E, vol.shape[1] - y), min(E, vol.shape[2] - x)
        if ze > zs and ye > ys and xe > xs:
            want[zs:ze, ys:ye, xs:xe] = vol[z + zs:z + ze, y + ys:y + ye, x + xs:x + xe]
        assert np.array_equal(got[j], want), f"{what}: box {j} {(z, y, x)}"


BOXES = [(10, 20, 30), (100, 200, 250), (0, 128, 500), (200, 5, 700), (30, 40, 520), (120, 120, 240)]


@pytest.mark.gpu
def test_serve_sees_appends_and_invalidations():
    import mc_amd
    L = _api()
    ch = mc_amd.synth_chunks([51, 52, 53, 54], [0.4, 0.5, 0.6, 0.35])
    b0, b2 = ob.encode_chunk(ch[0], 6.0), ob.encode_chunk(ch[1], 2.5, 8)
    b0_new, b1 = ob.encode_chunk(ch[2], 16.0), ob.encode_chunk(ch[3], 6.0)
    path = _tmp("coh.mca")
    a = L.mc_archive_open_dims(path.encode(), 768, 256, 256, C.c_float(6.0))
    assert a
    ctx = mc_amd.Ctx(0)
    try:
        for x, b in ((0, b0), (2, b2)):
            assert L.mc_archive_append_chunk_compressed(a, 0, 0, 0, x, b, len(b)) == 0
        print("serve before appends", flush=True)
        sv = mc_amd.Serve(a, lod=0, arena_bytes=64 << 20, lanes=2)
        try:
            _check(_serve(sv, ctx, BOXES), _oracle_vol({0: b0, 2: b2}, 3), BOXES, "before")
            # chunk (0,0,1) appended inside the grid, chunk (0,0,0) appended again
            assert L.mc_archive_append_chunk_compressed(a, 0, 0, 0, 1, b1, len(b1)) == 0
            assert L.mc_archive_append_chunk_compressed(a, 0, 0, 0, 0, b0_new, len(b0_new)) == 0
            print("serve after appends", flush=True)
            vol = _oracle_vol({0: b0_new, 1: b1, 2: b2}, 3)
            _check(_serve(sv, ctx, BOXES), vol, BOXES, "after appends")
            st = sv.stats()
            assert st["updated_blocks"] == 2 * 4096
            # a chunk outside the grid the level had at open is counted, not served
            assert L.mc_archive_append_chunk_compressed(a, 0, 1, 0, 0, b1, len(b1)) == 0
            assert sv.refresh() == 1
            # an explicit invalidation re-reads the chunk
            assert sv.invalidate_chunk(0, 0, 0, 2) == 0
            assert sv.invalidate_chunk(0, 1, 0, 0) == 1
            _check(_serve(sv, ctx, BOXES), vol, BOXES, "after invalidation")
        finally:
            sv.close()
        arr = (mc_amd.McBox * len(BOXES))(*[mc_amd.McBox(*b) for b in BOXES])
        rr = np.zeros((len(BOXES), E, E, E), np.uint8)
        L.mc_archive_read_regions(a, 0, arr, len(BOXES), E, E, E, ob.p(rr), rr[0].nbytes, 0)
        _check(rr, vol, BOXES, "read_regions after appends")
    finally:
        ctx.close()
        L.mc_archive_close(a)


@pytest.mark.gpu
def test_arena_resize_growth_and_memory_shortage():
    import mc_amd
    L = _api()
    n = 20                                          # 20 chunks in a row: up to 81920 blocks
    ch = mc_amd.synth_chunks(list(range(61, 61 + n)), [0.1] * n)
    path = _tmp("grow.mca")
    a = L.mc_archive_open_dims(path.encode(), 256 * n, 256, 256, C.c_float(6.0))
    assert a
    ctx = mc_amd.Ctx(0)
    print("appending 20 chunks on the GPU", flush=True)
    for x in range(n):
        assert L.mc_archive_append_chunk_par(a, 0, 0, 0, x, ob.p(np.ascontiguousarray(ch[x])), C.c_float(6.0),
                                             0) == 0
    vol = np.zeros((256, 256, 256 * n), np.uint8)
    dec = np.empty((256, 256, 256), np.uint8)
    for x in range(n):
        L.mc_archive_decode_chunk(a, L.mc_archive_chunk_offset(a, 0, 0, 0, x), ob.p(dec), 0)
        vol[:, :, 256 * x:256 * (x + 1)] = dec
    rng = np.random.default_rng(3)
    boxes = [(int(rng.integers(0, 208)), int(rng.integers(0, 208)), int(rng.integers(0, 256 * n - E)))
             for _ in range(8)]
    try:
        print("shrink and grow while serving", flush=True)
        sv = mc_amd.Serve(a, lod=0, arena_bytes=64 << 20, lanes=2)
        try:
            _check(_serve(sv, ctx, boxes), vol, boxes, "64 MiB arena")
            sv.resize(2 << 20)                      # 512 slots
            assert sv.stats()["arena_slots"] == 512
            _check(_serve(sv, ctx, boxes[:2]), vol, boxes[:2], "2 MiB arena")
            sv.resize(1 << 30)
            _check(_serve(sv, ctx, boxes), vol, boxes, "grown back")
        finally:
            sv.close()
        print("growth by segments on preload", flush=True)
        sv = mc_amd.Serve(a, lod=0, arena_bytes=1 << 30, lanes=2)
        try:
            nl = sv.preload()
            st = sv.stats()
            assert nl > 65536 and st["arena_grows"] >= 1 and st["arena_slots"] >= nl
            assert st["arena_mapped_bytes"] >= nl * 4096
            _check(_serve(sv, ctx, boxes), vol, boxes, "after growth")
            sv.resize(200 << 20)                    # gives the second segment back
            st = sv.stats()
            assert st["arena_slots"] == (200 << 20) // 4096 and st["arena_mapped_bytes"] == 256 << 20
            _check(_serve(sv, ctx, boxes), vol, boxes, "after shrinking")
        finally:
            sv.close()
        print("a request the device memory cannot hold", flush=True)
        sv = mc_amd.Serve(a, lod=0, arena_bytes=1 << 30, lanes=2)
        hog = None
        try:
            free, total = mc_amd.mem_info()
            hog = mc_amd.DeviceBuffer(free - (1 << 30))             # leaves less than the arena's headroom
            big = [(0, 0, 256 * x) for x in range(n)]                # every block: more than one segment
            d_out = mc_amd.DeviceBuffer(len(big) * 256 ** 3)
            t = sv.submit(big, 256, 256, 256, d_out)
            with pytest.raises(mc_amd.MchipError, match="device memory left free"):
                sv.wait(t)
            hog.free()
            hog = None
            t = sv.submit(big, 256, 256, 256, d_out)
            sv.wait(t)
            got = np.empty((256, 256, 256), np.uint8)
            for x in (0, n // 2, n - 1):
                ctx.d2h(got, d_out, offset=x * 256 ** 3)
                assert np.array_equal(got, vol[:, :, 256 * x:256 * (x + 1)]), f"chunk {x}"
            d_out.free()
        finally:
            if hog is not None:
                hog.free()
            sv.close()
    finally:
        ctx.close()
        L.mc_archive_close(a)
