"""Concurrent callers of the drop-in entry points.

The reference's decode paths are lock-free and safe to call from many
threads, and appends to distinct chunks run concurrently with decodes
(src/mc_archive.c:751-762, :783-820, :458-485).  Here 16 host threads call
mc_archive_decode_chunk, mc_archive_read_regions and
mc_archive_append_chunk_par at once, on shared and on their own archives,
while another thread keeps changing the decode tuning and decoding a
resident batch; every output is checked against the reference's manifest
hashes (tests/golden/chunk_manifest.json) or against voxels checked that
way."""
import ctypes as C
import json
import os
import threading

import numpy as np
import pytest

import oracle_bind as ob
from test_dropin_archive import _api, _tmp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NTHREADS = 16
COORDS = [(z, y, x) for z in range(2) for y in range(2) for x in range(2)]


def _manifest():
    with open(os.path.join(ROOT, "tests", "golden", "chunk_manifest.json")) as f:
        return json.load(f)["plain64_q6"][:8]


@pytest.mark.gpu
def test_sixteen_threads_decode_read_and_append_bit_exact():
    import mc_amd
    L = _api()
    man = _manifest()
    chunks = mc_amd.synth_chunks([m["seed"] for m in man], [m["air"] for m in man])
    for k, m in enumerate(man):
        assert f"{ob.xxh64(chunks[k]):016x}" == m["input_xxh64"]
    path = _tmp("shared.mca")
    a = L.mc_archive_open_dims(path.encode(), 512, 512, 512, C.c_float(6.0))
    assert a
    print("building the shared archive", flush=True)
    vol = np.zeros((512, 512, 512), np.uint8)
    offs = []
    dec = np.empty((256, 256, 256), np.uint8)
    for k, (z, y, x) in enumerate(COORDS):
        assert L.mc_archive_append_chunk_par(a, 0, z, y, x, ob.p(np.ascontiguousarray(chunks[k])),
                                             C.c_float(6.0), 0) == 0
        off = L.mc_archive_chunk_offset(a, 0, z, y, x)
        L.mc_archive_decode_chunk(a, off, ob.p(dec), 0)
        assert f"{ob.xxh64(dec):016x}" == man[k]["decoded_xxh64"]
        vol[z * 256:(z + 1) * 256, y * 256:(y + 1) * 256, x * 256:(x + 1) * 256] = dec
        offs.append(off)
    # the same blobs, read back from the file, resident on the device for the
    # tuning-churn thread
    L.mc_archive_close(a)
    data = np.fromfile(path, np.uint8)
    a = L.mc_archive_open_dims(path.encode(), 512, 512, 512, C.c_float(6.0))
    assert a
    blobs = [data[o:o + m["blob_len"]] for o, m in zip(offs, man)]
    for b, m in zip(blobs, man):
        assert f"{ob.xxh64(b):016x}" == m["blob_xxh64"]
    ctx = mc_amd.Ctx(0)
    pos = np.zeros(len(blobs), np.uint64)
    tot = 0
    for k, b in enumerate(blobs):
        pos[k] = tot
        tot += (b.size + 255) // 256 * 256
    flat = np.zeros(tot, np.uint8)
    for k, b in enumerate(blobs):
        flat[int(pos[k]):int(pos[k]) + b.size] = b
    d_blobs = mc_amd.DeviceBuffer(tot)
    ctx.h2d(d_blobs, flat)
    d_out = mc_amd.DeviceBuffer(len(blobs) * mc_amd.CHUNK_BYTES)

    errors = []
    stop = threading.Event()

    def worker(t):
        try:
            rng = np.random.default_rng(100 + t)
            out = np.empty((256, 256, 256), np.uint8)
            own = _tmp(f"own{t}.mca")
            b = L.mc_archive_open_dims(own.encode(), 256, 256, 256, C.c_float(6.0))
            assert b
            for it in range(6):
                k = int(rng.integers(0, len(COORDS))) if it % 2 else t % len(COORDS)
                L.mc_archive_decode_chunk(a, offs[k], ob.p(out), 0)
                if f"{ob.xxh64(out):016x}" != man[k]["decoded_xxh64"]:
                    errors.append(f"thread {t} decode_chunk {k} iteration {it}")
                org = rng.integers(-32, 480, (4, 3))
                boxes = (mc_amd.McBox * 4)(*[mc_amd.McBox(*map(int, o)) for o in org])
                got = np.zeros((4, 64, 64, 64), np.uint8)
                L.mc_archive_read_regions(a, 0, boxes, 4, 64, 64, 64, ob.p(got), got[0].nbytes, 0)
                for j, (z0, y0, x0) in enumerate(org):
                    want = np.zeros((64, 64, 64), np.uint8)
                    zs, ys, xs = max(0, -z0), max(0, -y0), max(0, -x0)
                    ze, ye, xe = min(64, 512 - z0), min(64, 512 - y0), min(64, 512 - x0)
                    want[zs:ze, ys:ye, xs:xe] = vol[z0 + zs:z0 + ze, y0 + ys:y0 + ye, x0 + xs:x0 + xe]
                    if not np.array_equal(got[j], want):
                        errors.append(f"thread {t} read_regions box {tuple(org[j])} iteration {it}")
                if it < 2:
                    k = (t + it) % len(COORDS)
                    if L.mc_archive_append_chunk_par(b, 0, 0, 0, 0, ob.p(np.ascontiguousarray(chunks[k])),
                                                     C.c_float(6.0), 0) != 0:
                        errors.append(f"thread {t} append failed")
                        continue
                    co = L.mc_archive_chunk_offset(b, 0, 0, 0, 0)
                    with open(own, "rb") as f:
                        f.seek(co)
                        blob = f.read(man[k]["blob_len"])
                    if f"{ob.xxh64(blob):016x}" != man[k]["blob_xxh64"]:
                        errors.append(f"thread {t} append_chunk_par blob {k}")
            L.mc_archive_close(b)
        except Exception as e:   # noqa: BLE001 - reported below
            errors.append(f"thread {t}: {e!r}")

    def churn():
        try:
            tctx = mc_amd.Ctx(0)
            lens = [b.size for b in blobs]
            qs = [6.0] * len(blobs)
            k = 0
            while not stop.is_set():
                mc_amd.set_tuning("persistent", k % 2)
                mc_amd.set_tuning("persist_wg", (0, 64, 200)[k % 3])
                mc_amd.set_tuning("persist_min", 1 + k % 3)
                tctx.decode_chunks(d_blobs, pos, lens, qs, d_out)
                tctx.sync()
                h = tctx.xxh64_ranges(d_out, np.arange(len(blobs), dtype=np.uint64) * mc_amd.CHUNK_BYTES,
                                      np.full(len(blobs), mc_amd.CHUNK_BYTES, np.uint64),
                                      np.zeros(len(blobs), np.uint64))
                for j, m in enumerate(man):
                    if f"{int(h[j]):016x}" != m["decoded_xxh64"]:
                        errors.append(f"batched decode under tuning change {k}: chunk {j}")
                k += 1
            tctx.close()
        except Exception as e:   # noqa: BLE001
            errors.append(f"churn: {e!r}")

    print(f"{NTHREADS} caller threads + tuning churn", flush=True)
    ths = [threading.Thread(target=worker, args=(t,)) for t in range(NTHREADS)]
    ch = threading.Thread(target=churn)
    ch.start()
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    stop.set()
    ch.join()
    for k, v in (("persistent", 1), ("persist_wg", 0), ("persist_min", 2)):
        mc_amd.set_tuning(k, v)
    d_blobs.free()
    d_out.free()
    ctx.close()
    L.mc_archive_close(a)
    assert not errors, errors[:10]
