"""HBM-resident serving (mc_serve_* / mchip_server_*): box batches through the
device decoded-block arena give the voxels of mc_archive_read_regions and of
the oracle, cold and warm, alone and in a mix of concurrent requests that
forces evictions, and a request larger than the arena fails cleanly."""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob
from test_dropin_archive import PLACES, _api, _blobs, _build_ours, _tmp

DZ, DY, DX = 40, 48, 64


def _volume(blobs):
    """Oracle decode of the archive's LOD-0 chunks, padded by 256 on every side."""
    big = np.zeros((1024, 1024, 1024), np.uint8)
    for (lod, cz, cy, cx), bl in zip(PLACES, blobs):
        big[256 + cz * 256:256 + (cz + 1) * 256, 256 + cy * 256:256 + (cy + 1) * 256,
            256 + cx * 256:256 + (cx + 1) * 256] = ob.decode_chunk(bl)
    return big


def _crop(big, o):
    z0, y0, x0 = o
    return big[256 + z0:256 + z0 + DZ, 256 + y0:256 + y0 + DY, 256 + x0:256 + x0 + DX]


@pytest.fixture(scope="module")
def served():
    import mc_amd
    blobs, _ = _blobs()
    a = _build_ours(_tmp("s.mca"), blobs)
    big = _volume(blobs)
    ctx = mc_amd.Ctx(0)
    yield a, big, ctx
    _api().mc_archive_close(a)


def _run(sv, ctx, origins):
    import mc_amd
    n = len(origins)
    d_out = mc_amd.DeviceBuffer(n * DZ * DY * DX)
    t = sv.submit(origins, DZ, DY, DX, d_out)
    lat = sv.wait(t)
    out = np.empty((n, DZ, DY, DX), np.uint8)
    ctx.d2h(out, d_out)
    d_out.free()
    return out, lat


BOXES = [(0, 0, 0), (10, 20, 30), (60, 100, 3), (64, 64, 300), (200, 250, 250), (250, 500, 490),
         (-5, -7, 100), (240, 240, 240), (300, 300, 10), (500, 250, 400)]


@pytest.mark.gpu
def test_serve_matches_read_regions_and_oracle_cold_and_warm(served):
    import mc_amd
    a, big, ctx = served
    print("serve open", flush=True)
    sv = mc_amd.Serve(a, lod=0, arena_bytes=64 << 20, lanes=4)
    try:
        print("cold batch", flush=True)
        cold, _ = _run(sv, ctx, BOXES)
        s1 = sv.stats()
        print("warm batch", flush=True)
        warm, _ = _run(sv, ctx, BOXES)
        s2 = sv.stats()
    finally:
        sv.close()
    arr = (mc_amd.McBox * len(BOXES))(*[mc_amd.McBox(*b) for b in BOXES])
    rr = np.zeros((len(BOXES), DZ, DY, DX), np.uint8)
    _api().mc_archive_read_regions(a, 0, arr, len(BOXES), DZ, DY, DX, ob.p(rr), rr[0].nbytes, 0)
    for k, o in enumerate(BOXES):
        want = _crop(big, o)
        assert np.array_equal(cold[k], want), f"cold box {k} {o}"
        assert np.array_equal(warm[k], want), f"warm box {k} {o}"
        assert np.array_equal(rr[k], want), f"read_regions box {k} {o}"
    assert s1["misses"] > 0 and s1["hits"] == 0
    refs2 = s2["block_refs"] - s1["block_refs"]
    assert refs2 == s1["block_refs"] and s2["hits"] - s1["hits"] == refs2 and s2["misses"] == s1["misses"]


@pytest.mark.gpu
def test_serve_mixed_requests_equal_alone_under_evictions(served):
    import mc_amd
    a, big, ctx = served
    rng = np.random.default_rng(11)
    reqs = [[(int(rng.integers(-20, 500)), int(rng.integers(-20, 500)), int(rng.integers(-20, 500)))
             for _ in range(4)] for _ in range(12)]
    # alone: each request on an idle server with a large arena
    print("alone", flush=True)
    sv = mc_amd.Serve(a, lod=0, arena_bytes=64 << 20, lanes=2)
    try:
        alone = [_run(sv, ctx, r)[0] for r in reqs]
    finally:
        sv.close()
    # mixed: all requests submitted at once, 4 lanes, an arena of 600 slots
    # (one request touches at most 320 blocks), so admissions wait for slots
    # and resident blocks are evicted and decoded again
    print("mixed", flush=True)
    sv = mc_amd.Serve(a, lod=0, arena_bytes=600 * 4096, lanes=4)
    outs = [mc_amd.DeviceBuffer(4 * DZ * DY * DX) for _ in reqs]
    try:
        tickets = [sv.submit(r, DZ, DY, DX, o) for r, o in zip(reqs, outs)]
        lats = [sv.wait(t) for t in tickets]
        st = sv.stats()
    finally:
        sv.close()
    assert st["completed"] == len(reqs) and st["evictions"] > 0
    assert all(l >= 0 for l in lats)
    for k, (r, o) in enumerate(zip(reqs, outs)):
        got = np.empty((4, DZ, DY, DX), np.uint8)
        ctx.d2h(got, o)
        o.free()
        assert np.array_equal(got, alone[k]), f"request {k}: mixed differs from alone"
        for j, org in enumerate(r):
            assert np.array_equal(got[j], _crop(big, org)), f"request {k} box {j} {org}"


@pytest.mark.gpu
def test_serve_request_larger_than_arena_fails_cleanly(served):
    import mc_amd
    a, big, ctx = served
    sv = mc_amd.Serve(a, lod=0, arena_bytes=16 * 4096, lanes=2)
    try:
        d_out = mc_amd.DeviceBuffer(DZ * DY * DX)
        print("oversized request", flush=True)
        t = sv.submit([(0, 0, 0)], DZ, DY, DX, d_out)
        with pytest.raises(mc_amd.MchipError):
            sv.wait(t)
        # a request that fits still works afterwards
        print("small request", flush=True)
        t = sv.submit([(3, 5, 7)], 8, 8, 8, d_out)
        sv.wait(t)
        got = np.empty((8, 8, 8), np.uint8)
        ctx.d2h(got, d_out)
        assert np.array_equal(got, big[256 + 3:256 + 11, 256 + 5:256 + 13, 256 + 7:256 + 15])
        d_out.free()
    finally:
        sv.close()


@pytest.mark.gpu
def test_serve_preload_then_hits_only(served):
    import mc_amd
    a, big, ctx = served
    sv = mc_amd.Serve(a, lod=0, arena_bytes=64 << 20, lanes=2)
    try:
        print("preload", flush=True)
        n = sv.preload()
        s0 = sv.stats()
        out, _ = _run(sv, ctx, BOXES)
        s1 = sv.stats()
    finally:
        sv.close()
    assert n > 0 and s0["misses"] == n
    assert s1["misses"] == s0["misses"] and s1["hits"] - s0["hits"] == s1["block_refs"] - s0["block_refs"]
    for k, o in enumerate(BOXES):
        assert np.array_equal(out[k], _crop(big, o)), f"box {k} {o}"
    # an arena smaller than the store: preload fills it, requests stay exact
    sv = mc_amd.Serve(a, lod=0, arena_bytes=700 * 4096, lanes=2)
    try:
        assert sv.preload() == 700
        out, _ = _run(sv, ctx, BOXES[:4])
    finally:
        sv.close()
    for k, o in enumerate(BOXES[:4]):
        assert np.array_equal(out[k], _crop(big, o)), f"small arena box {k} {o}"
