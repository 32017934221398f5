"""The oracle's LOD decimation and quad render, pinned.

Decimation: closed-form cells (constant, single nonzero voxel, rounding half
up, all air, odd edges) and the reference's one-shot builder: mc_build's
LOD-1 chunk blob equals the oracle encoder's blob of the oracle's decimation.

Render: mc_render_pick_lod, mc_quad_spacing and mc_render_quad_lod of the
reference compiled from its own sources (oracle/_ref/libmcref_render.so),
through its dense and its block paths, against the oracle on random
pyramids, quads (with invalid control points and grid edges), filters,
composites, zooms and slabs, bit for bit."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

import oracle_bind as ob

needs_ref = pytest.mark.skipif(not ob.ref_render_available(), reason="reference build not present")


def test_decimate_closed_forms():
    v = np.zeros((4, 4, 6), np.uint8)
    v[0:2, 0:2, 0:2] = 77                       # constant cell
    v[0, 2, 2] = 9                              # one nonzero voxel of a cell
    v[2, 0, 0], v[3, 1, 1] = 1, 2               # {1, 2}: (3 + 1) // 2 = 2, half up
    v[2, 2, 0], v[2, 2, 1], v[3, 3, 1] = 1, 1, 2   # {1, 1, 2}: (4 + 1) // 3 = 1
    d = ob.decimate(v)
    assert d.shape == (2, 2, 3)
    assert d[0, 0, 0] == 77 and d[0, 1, 1] == 9 and d[1, 0, 0] == 2 and d[1, 1, 0] == 1
    assert d[0, 0, 2] == 0 and d[1, 1, 2] == 0   # all-air cells stay air
    # odd dims: the last plane decimates alone
    o = np.zeros((3, 1, 1), np.uint8)
    o[:, 0, 0] = [10, 20, 31]
    assert ob.decimate(o)[:, 0, 0].tolist() == [15, 31]
    # every cell of 255s and every cell with one 255 and seven zeros -> 255
    assert (ob.decimate(np.full((6, 6, 6), 255, np.uint8)) == 255).all()


def _smooth_volume(shape, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float32) for n in shape], indexing="ij")
    f = 128 + 60 * np.sin(2 * np.pi * z / 37 + 0.5 * np.sin(2 * np.pi * x / 71)) * np.cos(2 * np.pi * y / 53)
    f += rng.normal(0, 6, shape)
    v = np.clip(np.rint(f), 1, 255).astype(np.uint8)
    air = np.sin(x / 23.0) + np.cos(y / 29.0) + np.sin(z / 31.0) < -0.4
    v[air] = 0
    return v


@needs_ref
def test_decimate_matches_reference_build_lod1():
    import mc_amd
    L = mc_amd.api()
    vol = _smooth_volume((256, 256, 512), 3)
    data = ob.ref_build(vol, 0.5)
    path = os.path.join(tempfile.mkdtemp(prefix="mcamd_"), "b.mca")
    with open(path, "wb") as f:
        f.write(data)
    a = L.mc_archive_open_dims(path.encode(), 512, 256, 256, C.c_float(0.5))
    assert a
    co = L.mc_archive_chunk_offset(a, 1, 0, 0, 0)
    L.mc_archive_close(a)
    assert co > 1
    arr = np.frombuffer(data, np.uint8)
    ref_blob = bytes(arr[co:co + ob.oracle().orc_blob_len(ob.p(arr[co:]))])
    chunk = np.zeros((256, 256, 256), np.uint8)
    chunk[:128, :128, :256] = ob.decimate(vol)
    ours = ob.encode_chunk(chunk, 0.5)
    assert ours == ref_blob


def _plane_grid(rng, lo, hi, gw, gh, spacing, invalid=0.0):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    u = np.cross(n, rng.normal(size=3))
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    o = rng.uniform(lo, hi, 3)
    j = np.arange(gw, dtype=np.float64) - gw / 2
    i = np.arange(gh, dtype=np.float64) - gh / 2
    g = o[None, None, :] + spacing * (i[:, None, None] * v[None, None, :] + j[None, :, None] * u[None, None, :])
    g = g.astype(np.float32)
    if invalid:
        m = rng.random((gh, gw)) < invalid
        g[m] = -1.0
    return g


def _cases():
    rng = np.random.default_rng(17)
    vol = _smooth_volume((72, 88, 104), 5)
    levels = ob.pyramid(vol, 4)
    out = []
    for k in range(40):
        gw, gh = int(rng.integers(2, 24)), int(rng.integers(2, 24))
        spacing = float(rng.uniform(2.0, 9.0))
        edge = k % 4 == 3                  # planes that leave the volume or the grid
        g = _plane_grid(rng, -10 if edge else 20, 110 if edge else 55, gw, gh, spacing,
                        invalid=0.05 if k % 3 == 0 else 0.0)
        zoom = [0.5, 1, 2, 4, 8, 16][k % 6]
        step = zoom / spacing
        w, h = int(rng.integers(8, 48)), int(rng.integers(8, 40))
        if not edge:                       # about the grid's extent in pixels
            w = int(min(w, max(8, (gw - 1) / step)))
            h = int(min(h, max(8, (gh - 1) / step)))
        if edge:
            x0, y0 = float(rng.uniform(-2, gw * 0.5)), float(rng.uniform(-2, gh * 0.5))
        else:                              # the pixel window centred on the grid
            x0 = float((gw - 1) / 2 - w * step / 2 + rng.uniform(-0.5, 0.5))
            y0 = float((gh - 1) / 2 - h * step / 2 + rng.uniform(-0.5, 0.5))
        filt = 1 if k % 5 else 0
        comp = [2, 4, 5, 0, 1, 3, 7, 8, 2, 5, 7, 8][k % 12]
        t0, t1, dt = [(-4.0, 4.0, 1.0), (-2.5, 3.0, 0.75), (3.0, -3.0, 0.5), (0.0, 0.0, 1.0),
                      (-7.0, 7.0, 1.0)][k % 5]
        out.append((levels, g, x0, y0, step, w, h, filt, comp, t0, t1, dt))
    return out


@needs_ref
def test_pick_lod_and_spacing_match_reference():
    R = ob.ref_render()
    O = ob.oracle()
    for v in [0.0, 0.5, 1.0, 1.999, 2.0, 3.9, 4.0, 7.99, 8.0, 100.0]:
        for n in [1, 2, 4, 8]:
            assert O.orc_pick_lod(n, v) == R.refshim_pick_lod(n, v)
    rng = np.random.default_rng(2)
    for k in range(30):
        gw, gh = int(rng.integers(1, 40)), int(rng.integers(1, 40))
        g = _plane_grid(rng, 0, 100, gw, gh, float(rng.uniform(0.5, 20)), invalid=0.2 if k % 2 else 0.0)
        assert O.orc_quad_spacing(ob.p(g), gw, gh) == R.refshim_quad_spacing(ob.p(g), gw, gh)


@needs_ref
@pytest.mark.parametrize("blocked", [0, 1])
def test_render_quad_lod_matches_reference(blocked):
    seen, lit = set(), []
    for k, (levels, g, x0, y0, step, w, h, filt, comp, t0, t1, dt) in enumerate(_cases()):
        ours, lod = ob.render_quad_lod(levels, g, x0, y0, step, w, h, filt, comp, t0, t1, dt)
        ref = ob.ref_render_quad_lod(levels, g, x0, y0, step, w, h, filt, comp, t0, t1, dt, blocked)
        assert np.array_equal(ours, ref), f"case {k}: lod {lod} filter {filt} comp {comp}"
        seen.add(lod)
        lit.append((ours > 0).mean())
    assert seen == {0, 1, 2, 3}
    assert np.mean(lit) > 0.3                 # the cases sample material, not only air


def _block_tables(vol, drop=0.0, seed=0):
    rng = np.random.default_rng(seed)
    nz, ny, nx = vol.shape
    gbz, gby, gbx = (nz + 15) // 16, (ny + 15) // 16, (nx + 15) // 16
    pad = np.zeros((gbz * 16, gby * 16, gbx * 16), np.uint8)
    pad[:nz, :ny, :nx] = vol
    blocks = pad.reshape(gbz, 16, gby, 16, gbx, 16).transpose(0, 2, 4, 1, 3, 5).reshape(-1, 4096)
    off = (np.arange(blocks.shape[0], dtype=np.int64) * 4096)
    if drop:
        gone = rng.random(blocks.shape[0]) < drop
        off[gone] = -1
        blocks = blocks.copy()
        blocks[gone] = 0
    return off, np.ascontiguousarray(blocks).reshape(-1), blocks


def test_render_block_tables_equal_dense():
    """The oracle's block-table source equals its dense source, and an absent
    block reads as zeros (the dense volume with that block cleared)."""
    for k, (levels, g, x0, y0, step, w, h, filt, comp, t0, t1, dt) in enumerate(_cases()[:12]):
        dense, _ = ob.render_quad_lod(levels, g, x0, y0, step, w, h, filt, comp, t0, t1, dt)
        tabs = [_block_tables(v)[:2] for v in levels]
        viab, _ = ob.render_quad_lod(levels, g, x0, y0, step, w, h, filt, comp, t0, t1, dt, blocks=tabs)
        assert np.array_equal(dense, viab), f"case {k}"
        holes = []
        cleared = []
        for i, v in enumerate(levels):
            off, data, blocks = _block_tables(v, drop=0.3, seed=k * 10 + i)
            holes.append((off, data))
            nz, ny, nx = v.shape
            gbz, gby, gbx = (nz + 15) // 16, (ny + 15) // 16, (nx + 15) // 16
            full = blocks.reshape(gbz, gby, gbx, 16, 16, 16).transpose(0, 3, 1, 4, 2, 5).reshape(
                gbz * 16, gby * 16, gbx * 16)
            cleared.append(np.ascontiguousarray(full[:nz, :ny, :nx]))
        a, _ = ob.render_quad_lod(levels, g, x0, y0, step, w, h, filt, comp, t0, t1, dt, blocks=holes)
        b, _ = ob.render_quad_lod(cleared, g, x0, y0, step, w, h, filt, comp, t0, t1, dt)
        assert np.array_equal(a, b), f"case {k} with absent blocks"
