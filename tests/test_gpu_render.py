"""Device quad render and LOD decimation against the oracle (which is pinned
to the reference's renderer and builder in test_oracle_render.py).

* mchip_decimate / mchip_decimate_chunks equal the oracle's decimation (odd
  dims, absent chunks).
* mchip_render_quad over dense device levels equals the oracle's
  mc_render_quad_lod restatement on random pyramids and quads: every LOD,
  both filters, every supported composite, grid edges and invalid points.
* Renders served from an archive's LOD pyramid (mc_serve_open_levels +
  mc_serve_render_quad: touch pass, admission, decode into the arena, render
  through the slot table) equal the oracle fed the oracle-decoded levels,
  alone and interleaved with box batches in an arena small enough to evict.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob
from test_dropin_archive import _api, _tmp
from test_oracle_render import _plane_grid, _smooth_volume

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import mc_amd
    return mc_amd.Ctx(0)


def _up(ctx, arr):
    import mc_amd
    b = mc_amd.DeviceBuffer(max(1, arr.nbytes))
    ctx.h2d(b, np.ascontiguousarray(arr))
    return b


def test_decimate_dense_and_chunk_groups(ctx):
    import mc_amd
    rng = np.random.default_rng(4)
    for shape in [(1, 1, 1), (3, 5, 7), (16, 16, 16), (33, 20, 65), (64, 48, 80)]:
        v = rng.integers(0, 256, shape, dtype=np.uint8)
        v[rng.random(shape) < 0.4] = 0
        src = _up(ctx, v)
        want = ob.decimate(v)
        dst = mc_amd.DeviceBuffer(want.nbytes)
        print(f"decimate {shape}", flush=True)
        ctx.decimate(src, *shape, dst)
        got = np.empty_like(want)
        ctx.d2h(got, dst)
        src.free(); dst.free()
        assert np.array_equal(got, want), f"shape {shape}"
    # a 2x2x2 group of chunks, two of them absent
    chunks = mc_amd.synth_chunks([11, 12, 13, 14, 15, 16], [0.3, 0.5, 0.2, 0.6, 0.4, 0.7])
    present = [0, 1, 2, 4, 5, 7]
    big = np.zeros((512, 512, 512), np.uint8)
    buf = _up(ctx, chunks)
    srcs = [None] * 8
    for k, ci in zip(present, range(6)):
        z, y, x = k >> 2, (k >> 1) & 1, k & 1
        big[256 * z:256 * z + 256, 256 * y:256 * y + 256, 256 * x:256 * x + 256] = chunks[ci]
        srcs[k] = (buf, ci * mc_amd.CHUNK_BYTES)
    dst = mc_amd.DeviceBuffer(mc_amd.CHUNK_BYTES)
    print("decimate chunk group", flush=True)
    ctx.decimate_chunks(srcs, dst)
    got = np.empty((256, 256, 256), np.uint8)
    ctx.d2h(got, dst)
    buf.free(); dst.free()
    assert np.array_equal(got, ob.decimate(big))


def _render_cases(levels, n, seed):
    rng = np.random.default_rng(seed)
    nz, ny, nx = levels[0].shape
    out = []
    for k in range(n):
        gw, gh = int(rng.integers(2, 40)), int(rng.integers(2, 40))
        spacing = float(rng.uniform(2.0, 9.0))
        edge = k % 4 == 3
        lo, hi = (-10, max(nz, ny, nx) + 10) if edge else (0.25 * min(nz, ny, nx), 0.75 * min(nz, ny, nx))
        g = _plane_grid(rng, lo, hi, gw, gh, spacing, invalid=0.05 if k % 3 == 0 else 0.0)
        zoom = [1, 2, 4, 8, 0.5, 16][k % 6]
        step = zoom / spacing
        w, h = int(rng.integers(16, 96)), int(rng.integers(16, 80))
        if not edge:
            w = int(min(w, max(16, (gw - 1) / step)))
            h = int(min(h, max(16, (gh - 1) / step)))
            x0 = float((gw - 1) / 2 - w * step / 2 + rng.uniform(-0.5, 0.5))
            y0 = float((gh - 1) / 2 - h * step / 2 + rng.uniform(-0.5, 0.5))
        else:
            x0, y0 = float(rng.uniform(-2, gw * 0.5)), float(rng.uniform(-2, gh * 0.5))
        filt = 1 if k % 5 else 0
        comp = [2, 4, 0, 5, 1, 3, 7, 8, 2, 5, 7, 8][k % 12]
        t0, t1, dt = [(-4.0, 4.0, 1.0), (-2.5, 3.0, 0.75), (3.0, -3.0, 0.5), (0.0, 0.0, 1.0),
                      (-7.0, 7.0, 1.0)][k % 5]
        out.append((g, x0, y0, step, w, h, filt, comp, t0, t1, dt))
    return out


def test_render_dense_levels_match_oracle(ctx):
    import mc_amd
    vol = _smooth_volume((96, 112, 136), 9)
    levels = ob.pyramid(vol, 4)
    dims = [v.shape for v in levels]
    dev = [_up(ctx, v) for v in levels]
    seen, lit = set(), []
    try:
        for k, (g, x0, y0, step, w, h, filt, comp, t0, t1, dt) in enumerate(_render_cases(levels, 48, 23)):
            want, lod = ob.render_quad_lod(levels, g, x0, y0, step, w, h, filt, comp, t0, t1, dt)
            out = mc_amd.DeviceBuffer(w * h)
            print(f"render case {k} lod {lod} filter {filt} comp {comp} {w}x{h}", flush=True)
            got_lod = ctx.render_quad(dev, dims, g, x0, y0, step, w, h, out, filt, comp, t0, t1, dt)
            got = np.empty((h, w), np.uint8)
            ctx.d2h(got, out)
            out.free()
            assert got_lod == lod, f"case {k}"
            assert np.array_equal(got, want), f"case {k}: lod {lod} filter {filt} comp {comp}, " \
                f"{int((got != want).sum())} pixels differ"
            seen.add(lod)
            lit.append((want > 0).mean())
    finally:
        for b in dev:
            b.free()
    assert seen == {0, 1, 2, 3} and np.mean(lit) > 0.3


def _pyramid_archive(vol, q=6.0):
    """An archive whose LODs 0..3 hold vol and its decimations (oracle-encoded
    chunks), and the oracle-decoded levels."""
    L = _api()
    levels = ob.pyramid(vol, 4)
    path = _tmp("lod.mca")
    nz, ny, nx = vol.shape
    a = L.mc_archive_open_dims(path.encode(), nx, ny, nz, C.c_float(q))
    assert a
    decoded = []
    for lod, lv in enumerate(levels):
        dz, dy, dx = lv.shape
        dec = np.zeros(lv.shape, np.uint8)
        for cz in range((dz + 255) // 256):
            for cy in range((dy + 255) // 256):
                for cx in range((dx + 255) // 256):
                    ch = np.zeros((256, 256, 256), np.uint8)
                    part = lv[256 * cz:256 * cz + 256, 256 * cy:256 * cy + 256, 256 * cx:256 * cx + 256]
                    ch[:part.shape[0], :part.shape[1], :part.shape[2]] = part
                    if not ch.any():
                        continue
                    blob = ob.encode_chunk(ch, q)
                    assert L.mc_archive_append_chunk_compressed(a, lod, cz, cy, cx, blob, len(blob)) == 0
                    d = ob.decode_chunk(blob)
                    dec[256 * cz:256 * cz + part.shape[0], 256 * cy:256 * cy + part.shape[1],
                        256 * cx:256 * cx + part.shape[2]] = d[:part.shape[0], :part.shape[1], :part.shape[2]]
        decoded.append(dec)
    return a, decoded


@pytest.fixture(scope="module")
def pyramid_archive():
    vol = _smooth_volume((256, 256, 512), 12)
    a, decoded = _pyramid_archive(vol)
    yield a, decoded
    _api().mc_archive_close(a)


def _serve_renders(sv, cases, ctx):
    import mc_amd
    outs, tickets = [], []
    for (g, x0, y0, step, w, h, filt, comp, t0, t1, dt) in cases:
        o = mc_amd.DeviceBuffer(w * h)
        tickets.append(sv.render(g, x0, y0, step, w, h, o, filt, comp, t0, t1, dt))
        outs.append(o)
    lods = [sv.render_level(t) for t in tickets]
    imgs = []
    for t, o, c in zip(tickets, outs, cases):
        sv.wait(t)
        img = np.empty((c[5], c[4]), np.uint8)
        ctx.d2h(img, o)
        o.free()
        imgs.append(img)
    return imgs, lods


def test_served_renders_match_oracle(ctx, pyramid_archive):
    import mc_amd
    a, decoded = pyramid_archive
    cases = _render_cases(decoded, 16, 31)
    sv = mc_amd.Serve(a, 0, arena_bytes=64 << 20, lanes=4, nlods=4)
    try:
        print("served renders, one at a time", flush=True)
        imgs, lods = [], []
        for c in cases:
            i1, l1 = _serve_renders(sv, [c], ctx)
            imgs += i1
            lods += l1
        st = sv.stats()
    finally:
        sv.close()
    assert st["renders"] == len(cases) and st["misses"] > 0
    for k, (c, img, lod) in enumerate(zip(cases, imgs, lods)):
        want, wl = ob.render_quad_lod(decoded, *c)
        assert lod == wl, f"case {k}"
        assert np.array_equal(img, want), f"case {k}: lod {lod}, {int((img != want).sum())} pixels differ"
    assert set(lods) == {0, 1, 2, 3}


def test_served_renders_and_boxes_interleaved_small_arena(ctx, pyramid_archive):
    """Renders and box batches submitted together into an arena of 500
    slots: admissions wait for slots, blocks are evicted and decoded again,
    renders reuse blocks decoded by box requests and the other way round."""
    import mc_amd
    a, decoded = pyramid_archive
    cases = _render_cases(decoded, 24, 47)
    rng = np.random.default_rng(8)
    boxes = [[(int(rng.integers(-8, 250)), int(rng.integers(-8, 250)), int(rng.integers(-8, 500)))
              for _ in range(4)] for _ in range(12)]
    sv = mc_amd.Serve(a, 0, arena_bytes=500 * 4096, lanes=4, nlods=4)
    outs, tickets, kinds = [], [], []
    try:
        print("interleaved renders and boxes", flush=True)
        for k in range(max(len(cases), len(boxes))):
            if k < len(cases):
                g, x0, y0, step, w, h, filt, comp, t0, t1, dt = cases[k]
                o = mc_amd.DeviceBuffer(w * h)
                tickets.append(sv.render(g, x0, y0, step, w, h, o, filt, comp, t0, t1, dt))
                outs.append(o)
                kinds.append(("r", k))
            if k < len(boxes):
                o = mc_amd.DeviceBuffer(4 * 24 * 24 * 24)
                tickets.append(sv.submit(boxes[k], 24, 24, 24, o))
                outs.append(o)
                kinds.append(("b", k))
        for t in tickets:
            sv.wait(t)
        st = sv.stats()
    finally:
        sv.close()
    for (kind, k), o in zip(kinds, outs):
        if kind == "r":
            c = cases[k]
            img = np.empty((c[5], c[4]), np.uint8)
            ctx.d2h(img, o)
            want, _ = ob.render_quad_lod(decoded, *c)
            assert np.array_equal(img, want), f"render {k}"
        else:
            got = np.empty((4, 24, 24, 24), np.uint8)
            ctx.d2h(got, o)
            big = np.zeros((256 + 40, 256 + 40, 512 + 40), np.uint8)
            big[8:8 + 256, 8:8 + 256, 8:8 + 512] = decoded[0]
            for j, (z, y, x) in enumerate(boxes[k]):
                assert np.array_equal(got[j], big[z + 8:z + 32, y + 8:y + 32, x + 8:x + 32]), f"box {k}/{j}"
        o.free()
    assert st["evictions"] > 0 and st["renders"] == len(cases)
