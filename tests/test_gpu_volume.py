"""Serving parity at the dataloader workload's shape, on the 2x2x2-chunk
corner of the 2048^3 serving volume (volume2048_q6).

The corner's chunks are generated, appended with mc_archive_append_chunk_par
and decoded with mc_archive_decode_chunk; blobs and voxels are checked
against the reference's manifest (tests/golden/volume_manifest.json).  Batches
of 64 boxes of 128^3 drawn by mc_archive_sample_boxes (minimum material
fraction 0.5, src/mc_archive.c:922-948) are then served cold, warm and
through a small arena that forces evictions, and read with
mc_archive_read_regions into host buffers; every box equals the crop of the
manifest-checked voxels."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_bind as ob
from test_dropin_archive import _api, _tmp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CORNER = [(z, y, x) for z in range(2) for y in range(2) for x in range(2)]
EDGE = 128


def _manifest():
    with open(os.path.join(ROOT, "tests", "golden", "volume_manifest.json")) as f:
        m = json.load(f)["volume2048_q6"]
    return {int(k): {tuple(r["coord"]): r for r in v} for k, v in m.items()}


def _corner_chunks():
    import mc_amd
    import workloads
    specs = {s["coord"]: s for s in workloads.WORKLOADS["volume2048_q6"]()}
    sel = [specs[c] for c in CORNER]
    return mc_amd.synth_chunks([s["seed"] for s in sel], [s["air"] for s in sel], coords=CORNER)


def test_volume_manifest_inputs_are_the_generator_and_its_decimation():
    """The manifest's LOD-0 inputs are the workload generator's chunks and its
    LOD-1 chunk (0,0,0) is their 2x2x2 decimation (the oracle's decimation,
    pinned to the reference's mc_build in test_oracle_render.py)."""
    man = _manifest()
    ch = _corner_chunks()
    big = np.zeros((512, 512, 512), np.uint8)
    for k, (z, y, x) in enumerate(CORNER):
        assert f"{ob.xxh64(ch[k]):016x}" == man[0][(z, y, x)]["input_xxh64"], f"LOD 0 {(z, y, x)}"
        big[z * 256:(z + 1) * 256, y * 256:(y + 1) * 256, x * 256:(x + 1) * 256] = ch[k]
    assert f"{ob.xxh64(ob.decimate(big)):016x}" == man[1][(0, 0, 0)]["input_xxh64"]


def _sample(L, a, seed, count=64):
    import mc_amd
    arr = (mc_amd.McBox * count)()
    n = L.mc_archive_sample_boxes(a, 0, seed, count, EDGE, EDGE, EDGE, C.c_float(0.5), arr)
    return [(arr[i].z0, arr[i].y0, arr[i].x0) for i in range(n)]


def _crop(vol, o):
    z0, y0, x0 = o
    return vol[z0:z0 + EDGE, y0:y0 + EDGE, x0:x0 + EDGE]


@pytest.mark.gpu
def test_sampled_128_cube_batches_served_and_read_bit_exact():
    import mc_amd
    L = _api()
    man = _manifest()
    ch = _corner_chunks()
    path = _tmp("corner.mca")
    a = L.mc_archive_open_dims(path.encode(), 512, 512, 512, C.c_float(6.0))
    assert a
    print("appending the 8 corner chunks", flush=True)
    for k, c in enumerate(CORNER):
        assert L.mc_archive_append_chunk_par(a, 0, *c, ob.p(np.ascontiguousarray(ch[k])), C.c_float(6.0), 0) == 0
    vol = np.zeros((512, 512, 512), np.uint8)
    dec = np.empty((256, 256, 256), np.uint8)
    offs = {}
    for c in CORNER:
        off = L.mc_archive_chunk_offset(a, 0, *c)
        offs[c] = off
        L.mc_archive_decode_chunk(a, off, ob.p(dec), 0)
        assert f"{ob.xxh64(dec):016x}" == man[0][c]["decoded_xxh64"], f"decode {c}"
        z, y, x = c
        vol[z * 256:(z + 1) * 256, y * 256:(y + 1) * 256, x * 256:(x + 1) * 256] = dec
    L.mc_archive_close(a)
    data = np.fromfile(path, np.uint8)
    for c in CORNER:
        bl = data[offs[c]:offs[c] + man[0][c]["blob_len"]]
        assert f"{ob.xxh64(bl):016x}" == man[0][c]["blob_xxh64"], f"blob {c}"
    a = L.mc_archive_open_dims(path.encode(), 512, 512, 512, C.c_float(6.0))
    assert a
    ctx = mc_amd.Ctx(0)
    try:
        batches = [_sample(L, a, 10000 + k) for k in range(4)]
        assert all(len(b) >= 16 for b in batches), [len(b) for b in batches]
        nb = max(len(b) for b in batches)
        bytes_box = EDGE ** 3
        outs = [mc_amd.DeviceBuffer(nb * bytes_box) for _ in range(4)]
        got = np.empty((nb, EDGE, EDGE, EDGE), np.uint8)

        def check(k, out, what):
            n = len(batches[k])
            ctx.d2h(got[:n], out)
            for j, o in enumerate(batches[k]):
                assert np.array_equal(got[j], _crop(vol, o)), f"{what}: batch {k} box {j} {o}"

        print("serve cold and warm", flush=True)
        sv = mc_amd.Serve(a, lod=0, arena_bytes=256 << 20, lanes=4)
        try:
            sv.wait(sv.submit(batches[0], EDGE, EDGE, EDGE, outs[0]))
            s1 = sv.stats()
            check(0, outs[0], "cold")
            sv.wait(sv.submit(batches[0], EDGE, EDGE, EDGE, outs[1]))
            s2 = sv.stats()
            check(0, outs[1], "warm")
        finally:
            sv.close()
        assert s1["misses"] > 0 and s2["misses"] == s1["misses"]
        assert s2["hits"] - s1["hits"] == s2["block_refs"] - s1["block_refs"]
        # distinct present blocks of each batch (a block with any material
        # voxel has a payload); the arena holds the largest batch but not all
        present = np.zeros((32, 32, 32), bool)
        for k, (z, y, x) in enumerate(CORNER):
            present[z * 16:(z + 1) * 16, y * 16:(y + 1) * 16, x * 16:(x + 1) * 16] = \
                ch[k].reshape(16, 16, 16, 16, 16, 16).any(axis=(1, 3, 5))
        sets = []
        for bt in batches:
            m = np.zeros_like(present)
            for z0, y0, x0 in bt:
                m[z0 >> 4:(z0 + EDGE + 15) >> 4, y0 >> 4:(y0 + EDGE + 15) >> 4, x0 >> 4:(x0 + EDGE + 15) >> 4] = True
            sets.append(m & present)
        need = max(int(m.sum()) for m in sets)
        union = int(np.logical_or.reduce(sets).sum())
        slots = need + 64
        print(f"serve 4 batches at once: {slots}-slot arena, largest batch {need} blocks, union {union}", flush=True)
        sv = mc_amd.Serve(a, lod=0, arena_bytes=slots * 4096, lanes=4)
        try:
            tickets = [sv.submit(b, EDGE, EDGE, EDGE, o) for b, o in zip(batches, outs)]
            for t in tickets:
                sv.wait(t)
            st = sv.stats()
        finally:
            sv.close()
        assert st["completed"] == 4 and (st["evictions"] > 0 or slots >= union)
        for k in range(4):
            check(k, outs[k], "bounded arena")
        for o in outs:
            o.free()
        print("read_regions into host buffers", flush=True)
        b = batches[1]
        arr = (mc_amd.McBox * len(b))(*[mc_amd.McBox(*o) for o in b])
        rr = np.zeros((len(b), EDGE, EDGE, EDGE), np.uint8)
        L.mc_archive_read_regions(a, 0, arr, len(b), EDGE, EDGE, EDGE, ob.p(rr), rr[0].nbytes, 0)
        for j, o in enumerate(b):
            assert np.array_equal(rr[j], _crop(vol, o)), f"read_regions box {j} {o}"
    finally:
        ctx.close()
        L.mc_archive_close(a)
