"""Generate the committed golden fixtures from the reference codec.

Runs ONLY in the container that holds the reference sources: it loads
oracle/_ref/libmcref.so (the reference's src/mc_codec.c, src/mc_archive.c and
src/mc_stream.c compiled in place by oracle/Makefile with the reference's
deterministic flags) and records what the reference produces:

  block_kat.npz       per-block known answers: input voxels, (q, tau),
                      payload bytes from mc_enc_block, voxels from
                      mc_dec_block, for the fuzz-seed block kinds
                      (tests/fuzz/mc_fuzz_block_seed.c:16-24), crops of the
                      synthetic workload, random and sparse blocks, over the
                      preset ladder and q in {1, 3, 6, 12};
  chunk_manifest.json per-chunk input xxh64, blob length, blob xxh64 and
                      decoded xxh64 for the plain q=6 workload (64 chunks,
                      mc_archive_append_chunk_par + mc_archive_decode_chunk)
                      and the mixed-preset workload (256 chunks,
                      mc_archive_append_chunk_ctx with mc_apply_preset);
  volume_manifest.json the same per chunk of the 2048^3 serving volume
                      (volume2048_q6) at LOD 0 and of its decimations LOD 1-3
                      (each 2x2x2 chunk group decimated as src/mc_archive.c:
                      61-72 does, by the oracle's decimation, which the tests
                      pin to the reference's mc_build), every chunk encoded by
                      mc_archive_append_chunk_par at q=6;
  priors_golden.json  a per-volume prior override: the endpoint tables, the
                      blobs mc_archive_append_chunk_par (and, for one chunk,
                      mc_archive_append_chunk_ctx at tau 16) writes after
                      mc_archive_set_priors, their decoded xxh64 and the xxh64
                      of the whole archive file.

Usage: python tests/golden/make_golden.py [--blocks-only | --only NAME ...]
"""
import ctypes as C
import json
import os
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "matter-compressor_amd"))
import oracle_bind as ob  # noqa: E402
import mc_amd  # noqa: E402
import workloads  # noqa: E402

QS_EXTRA = [1.0, 3.0, 6.0, 12.0]


def ref_enc(R, ctx, vox, q, tau):
    R.mc_codec_ctx_set_quality(ctx, q)
    R.mc_codec_ctx_set_max_error(ctx, tau)
    b = ob._McBuf()
    n = C.c_uint32()
    v = np.ascontiguousarray(vox, np.uint8)
    R.mc_enc_block(ctx, ob.p(v), C.byref(b), C.byref(n))
    data = C.string_at(b.p, n.value) if n.value else b""
    return data


def ref_dec(R, ctx, payload, q):
    R.mc_codec_ctx_set_quality(ctx, q)
    out = np.zeros(4096, np.uint8)
    a = np.frombuffer(payload, np.uint8) if payload else np.zeros(1, np.uint8)
    R.mc_dec_block(ctx, ob.p(a), len(payload), ob.p(out))
    return out


def block_inputs():
    rng = np.random.default_rng(20240611)
    z, y, x = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    i = (z * 16 + y) * 16 + x
    kinds = {
        "textured": (40 + ((i * 7) % 160)).astype(np.uint8),
        "constant": np.full(4096, 128, np.uint8).reshape(16, 16, 16),
        "ball": np.where(np.sqrt((x - 8) ** 2 + (y - 8) ** 2 + (z - 8) ** 2) > 6, 0,
                         (200 - np.sqrt((x - 8) ** 2 + (y - 8) ** 2 + (z - 8) ** 2) * 10)).astype(np.uint8),
    }
    blocks = [(k, v.ravel()) for k, v in kinds.items()]
    chunk = mc_amd.synth_chunks([1], 0.419)[0]
    for j in range(12):
        bz, by, bx = rng.integers(0, 16, 3)
        crop = chunk[bz * 16:bz * 16 + 16, by * 16:by * 16 + 16, bx * 16:bx * 16 + 16]
        blocks.append((f"synth{j}", np.ascontiguousarray(crop).ravel()))
    blocks.append(("random", rng.integers(0, 256, 4096).astype(np.uint8)))
    sp = np.zeros(4096, np.uint8)
    sp[rng.integers(0, 4096, 9)] = rng.integers(1, 256, 9)
    blocks.append(("sparse", sp))
    return blocks


def make_blocks(R):
    ctx = R.mc_codec_ctx_new()
    cases = []
    for name, vox in block_inputs():
        settings = list(workloads.PRESETS) + [(q, 0) for q in QS_EXTRA]
        for q, tau in settings:
            pay = ref_enc(R, ctx, vox, q, tau)
            dec = ref_dec(R, ctx, pay, q) if pay else np.zeros(4096, np.uint8)
            cases.append((name, vox, q, tau, pay, dec))
    R.mc_codec_ctx_free(ctx)
    offs = np.zeros(len(cases) + 1, np.int64)
    for k, c in enumerate(cases):
        offs[k + 1] = offs[k] + len(c[4])
    payload = np.frombuffer(b"".join(c[4] for c in cases), np.uint8)
    np.savez_compressed(os.path.join(HERE, "block_kat.npz"),
                        names=np.array([c[0] for c in cases]),
                        vox=np.stack([c[1] for c in cases]).astype(np.uint8),
                        q=np.array([c[2] for c in cases], np.float32),
                        tau=np.array([c[3] for c in cases], np.int32),
                        offs=offs, payload=payload,
                        dec=np.stack([c[5] for c in cases]).astype(np.uint8))
    print(f"block_kat.npz: {len(cases)} cases, {payload.size} payload bytes")


def blob_from_file(path, off):
    data = np.fromfile(path, np.uint8)
    n = int(ob.oracle().orc_blob_len(ob.p(data[off:])))
    return data[off:off + n].tobytes()


def ref_chunk(spec, chunk, mode):
    """Encode one chunk with the reference; returns (blob, decoded)."""
    R = ob.ref()
    fd, path = tempfile.mkstemp(suffix=".mca", dir="/tmp")
    os.close(fd)
    os.remove(path)
    a = R.mc_archive_open_dims(path.encode(), 256, 256, 256, C.c_float(spec["q"]))
    vox = np.ascontiguousarray(chunk)
    if mode == "par":
        rc = R.mc_archive_append_chunk_par(a, 0, 0, 0, 0, ob.p(vox), C.c_float(spec["q"]), 1)
    else:
        ctx = R.mc_codec_ctx_new()
        R.mc_apply_preset(ctx, spec["preset"])
        rc = R.mc_archive_append_chunk_ctx(a, ctx, 0, 0, 0, 0, ob.p(vox))
        R.mc_codec_ctx_free(ctx)
    assert rc == 0
    off = R.mc_archive_chunk_offset(a, 0, 0, 0, 0)
    dec = np.zeros((256, 256, 256), np.uint8)
    R.mc_archive_decode_chunk(a, off, ob.p(dec), 1)
    R.mc_archive_close(a)
    blob = blob_from_file(path, off) if off > 1 else b""
    os.remove(path)
    return blob, dec


def manifest_for(specs, mode, threads=8):
    out = []
    batch = 32
    for b0 in range(0, len(specs), batch):
        sub = specs[b0:b0 + batch]
        chunks = mc_amd.synth_chunks([s["seed"] for s in sub], [s["air"] for s in sub])

        def one(k):
            blob, dec = ref_chunk(sub[k], chunks[k], mode)
            return dict(seed=sub[k]["seed"], air=sub[k]["air"], q=sub[k]["q"], tau=sub[k]["tau"],
                        preset=sub[k]["preset"], input_xxh64=f"{ob.xxh64(chunks[k]):016x}",
                        blob_len=len(blob), blob_xxh64=f"{ob.xxh64(blob):016x}",
                        decoded_xxh64=f"{ob.xxh64(dec):016x}")

        with ThreadPoolExecutor(threads) as ex:
            out.extend(ex.map(one, range(len(sub))))
        print(f"  {mode}: {b0 + len(sub)}/{len(specs)}", flush=True)
    return out


def ref_encode_level(chunks, q=6.0, threads=8):
    """Reference blob and decoded raster of each 256^3 chunk (par path)."""
    spec = dict(q=q, preset=None)
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda c: ref_chunk(spec, c, "par"), chunks))


def level_records(coords, chunks, res):
    return [dict(coord=list(c), input_xxh64=f"{ob.xxh64(ch):016x}", blob_len=len(bl),
                 blob_xxh64=f"{ob.xxh64(bl):016x}", decoded_xxh64=f"{ob.xxh64(dec):016x}")
            for c, ch, (bl, dec) in zip(coords, chunks, res)]


def make_volume():
    """volume2048_q6 and its LOD pyramid, 2x2x2 chunk groups at a time."""
    specs = workloads.WORKLOADS["volume2048_q6"]()
    by = {x["coord"]: x for x in specs}
    cube = [(z, y, x) for z in range(2) for y in range(2) for x in range(2)]
    levels = {0: []}
    lod1 = np.zeros((1024, 1024, 1024), np.uint8)
    for gz in range(4):
        for gy in range(4):
            for gx in range(4):
                grp = [by[(2 * gz + z, 2 * gy + y, 2 * gx + x)] for z, y, x in cube]
                ch = mc_amd.synth_chunks([x["seed"] for x in grp], [x["air"] for x in grp],
                                         coords=[x["coord"] for x in grp])
                levels[0] += level_records([x["coord"] for x in grp], ch, ref_encode_level(ch))
                blk = np.zeros((512, 512, 512), np.uint8)
                for k, (z, y, x) in enumerate(cube):
                    blk[z * 256:(z + 1) * 256, y * 256:(y + 1) * 256, x * 256:(x + 1) * 256] = ch[k]
                lod1[gz * 256:(gz + 1) * 256, gy * 256:(gy + 1) * 256, gx * 256:(gx + 1) * 256] = ob.decimate(blk)
        print(f"  volume LOD 0: {len(levels[0])}/512", flush=True)
    vol = lod1
    for lod in (1, 2, 3):
        n = vol.shape[0] // 256
        coords = [(z, y, x) for z in range(n) for y in range(n) for x in range(n)]
        chunks = [np.ascontiguousarray(vol[z * 256:(z + 1) * 256, y * 256:(y + 1) * 256, x * 256:(x + 1) * 256])
                  for z, y, x in coords]
        levels[lod] = level_records(coords, chunks, ref_encode_level(chunks))
        print(f"  volume LOD {lod}: {len(coords)} chunks", flush=True)
        if lod < 3:
            vol = ob.decimate(vol)
    with open(os.path.join(HERE, "volume_manifest.json"), "w") as f:
        json.dump({"volume2048_q6": {str(k): v for k, v in levels.items()}}, f, indent=0)
    print("volume_manifest.json written")


PRIOR_CHUNKS = [  # (seed, air, q, chunk coordinate)
    (1, 0.419, 6.0, (0, 0, 0)), (2, 0.5, 2.5, (0, 0, 1)), (3, 0.3, 16.0, (0, 1, 0)), (4, 0.6, 1.0, (1, 0, 0))]
PRIOR_CTX = (5, 0.45, 4, (1, 1, 1))   # (seed, air, preset index, coordinate) through append_chunk_ctx


def make_priors():
    """The reference writing an archive under a per-volume prior override."""
    R = ob.ref()
    lo, hi = ob.custom_priors()
    fd, path = tempfile.mkstemp(suffix=".mca", dir="/tmp")
    os.close(fd)
    os.remove(path)
    a = R.mc_archive_open_dims(path.encode(), 512, 512, 512, C.c_float(6.0))
    assert R.mc_archive_set_priors(a, ob.p(lo), ob.p(hi)) == 0
    recs = []
    seeds = [c[0] for c in PRIOR_CHUNKS] + [PRIOR_CTX[0]]
    airs = [c[1] for c in PRIOR_CHUNKS] + [PRIOR_CTX[1]]
    chunks = mc_amd.synth_chunks(seeds, airs)
    for k, (seed, air, q, (cz, cy, cx)) in enumerate(PRIOR_CHUNKS):
        assert R.mc_archive_append_chunk_par(a, 0, cz, cy, cx, ob.p(chunks[k]), C.c_float(q), 1) == 0
        recs.append(dict(seed=seed, air=air, q=q, tau=0, preset=None, coord=[cz, cy, cx], path="par"))
    seed, air, preset, (cz, cy, cx) = PRIOR_CTX
    ctx = R.mc_codec_ctx_new()
    R.mc_apply_preset(ctx, preset)
    assert R.mc_archive_append_chunk_ctx(a, ctx, 0, cz, cy, cx, ob.p(chunks[-1])) == 0
    R.mc_codec_ctx_free(ctx)
    q, tau = workloads.PRESETS[preset]
    recs.append(dict(seed=seed, air=air, q=q, tau=tau, preset=preset, coord=[cz, cy, cx], path="ctx"))
    for k, r in enumerate(recs):
        off = R.mc_archive_chunk_offset(a, 0, *r["coord"])
        dec = np.zeros((256, 256, 256), np.uint8)
        R.mc_archive_decode_chunk(a, off, ob.p(dec), 1)
        r["chunk_off"] = int(off)
        r["input_xxh64"] = f"{ob.xxh64(chunks[k]):016x}"
        r["decoded_xxh64"] = f"{ob.xxh64(dec):016x}"
    R.mc_archive_close(a)
    R.mc_codec_set_priors(None, None)
    data = np.fromfile(path, np.uint8)
    os.remove(path)
    for r in recs:
        bl = blob_from_file_bytes(data, r["chunk_off"])
        r["blob_len"] = len(bl)
        r["blob_xxh64"] = f"{ob.xxh64(bl):016x}"
    out = dict(plo_u16le=lo.astype("<u2").tobytes().hex(), phi_u16le=hi.astype("<u2").tobytes().hex(),
               dims=[512, 512, 512], quality=6.0,
               chunks=recs, archive_len=int(data.size), archive_xxh64=f"{ob.xxh64(data):016x}")
    with open(os.path.join(HERE, "priors_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"priors_golden.json written ({len(recs)} chunks, archive {data.size} bytes)")


def blob_from_file_bytes(data, off):
    n = int(ob.oracle().orc_blob_len(ob.p(data[off:])))
    return data[off:off + n].tobytes()


def main():
    if not ob.ref_available():
        sys.exit("oracle/_ref/libmcref.so missing: run make -C oracle (needs the reference sources)")
    R = ob.ref()
    only = sys.argv[sys.argv.index("--only") + 1:] if "--only" in sys.argv else None

    def want(name):
        return only is None or name in only

    if want("blocks"):
        make_blocks(R)
    if "--blocks-only" in sys.argv:
        return
    if want("priors"):
        make_priors()
    if want("chunks"):
        man = {
            "plain64_q6": manifest_for(workloads.plain_chunks(64), "par"),
            "mixed256": manifest_for(workloads.mixed_chunks(256), "ctx"),
        }
        with open(os.path.join(HERE, "chunk_manifest.json"), "w") as f:
            json.dump(man, f, indent=1)
        print("chunk_manifest.json written")
    if want("volume"):
        make_volume()


if __name__ == "__main__":
    main()
