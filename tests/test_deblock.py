"""Decode-side deblocking (mc_deblock, src/mc_codec.c:561-595): the oracle
against the reference build and closed-form face cases (CPU), the GPU filter
against the oracle on decoded volumes of several shapes and qualities."""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob

# shapes whose last face on every axis has its far neighbour inside the
# raster (the reference reads outside its buffer otherwise)
SHAPES = [(64, 64, 64), (48, 80, 40), (32, 20, 70), (100, 36, 52), (16, 16, 16), (20, 18, 34)]


def _volume(shape, seed):
    """Smooth material with air regions and block-edge steps: the filter
    meets flat seams, real edges and air on both sides of faces."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    v = 120 + 60 * np.sin(z / 7.0 + 0.3 * np.sin(x / 11.0)) + 25 * np.cos(y / 9.0)
    v += ((x // 16 + y // 16 + z // 16) % 3) * rng.integers(0, 9)      # per-block offsets: seams
    v += rng.normal(0, 3, shape)
    v = np.clip(np.round(v), 1, 255)
    v[rng.random(shape) < 0.05] = 0                                   # scattered air
    v[:, :, : shape[2] // 5] = 0                                      # an air slab
    return v.astype(np.uint8)


def test_oracle_deblock_matches_reference():
    if not ob.ref_available():
        pytest.skip("reference build not present")
    R = ob.ref()
    for k, shape in enumerate(SHAPES):
        for q in (0.5, 2.5, 6.0, 16.0, 64.0):
            v = _volume(shape, 10 * k + int(q))
            want = v.copy()
            R.mc_deblock(ob.p(want), *shape, C.c_float(q))
            assert np.array_equal(ob.deblock(v, q), want), f"shape {shape} q {q}"


def test_oracle_deblock_closed_form_faces():
    # one X face between two flat blocks of 100 and 104 at q=6 (beta 24, tc 4):
    # delta = (4*4 + (100-104) + 4) >> 3 = 2 -> 102 | 102
    v = np.zeros((1, 1, 32), np.uint8)
    v[..., :16] = 100
    v[..., 16:] = 104
    out = ob.deblock(v, 6.0)
    assert out[0, 0, 15] == 102 and out[0, 0, 16] == 102
    assert np.array_equal(out[0, 0, :15], v[0, 0, :15]) and np.array_equal(out[0, 0, 17:], v[0, 0, 17:])
    # a step of beta or more is a real edge and stays
    v[..., 16:] = 100 + 24
    assert np.array_equal(ob.deblock(v, 6.0), v)
    # air on either side of the face stays air and blocks the filter
    v[..., 16:] = 104
    v[0, 0, 15] = 0
    assert np.array_equal(ob.deblock(v, 6.0), v)
    # the correction is clamped to tc = (int)(0.5 q + 1)
    v = np.zeros((1, 1, 32), np.uint8)
    v[..., :16] = 50
    v[..., 16:] = 80
    out = ob.deblock(v, 16.0)        # beta 54, tc 9; delta = (120 - 30 + 4) >> 3 = 11 -> 9
    assert out[0, 0, 15] == 59 and out[0, 0, 16] == 71


@pytest.mark.gpu
def test_gpu_deblock_matches_oracle():
    import mc_amd
    ctx = mc_amd.Ctx(0)
    L = mc_amd.load()
    L.mchip_deblock_host.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float]
    try:
        for k, shape in enumerate(SHAPES + [(256, 256, 256)]):
            for q in (0.5, 6.0, 64.0):
                v = _volume(shape, 100 + 10 * k + int(q))
                print(f"deblock {shape} q={q}", flush=True)
                got = v.copy()
                mc_amd.check(L.mchip_deblock_host(ctx.h, ob.p(got), *shape, C.c_float(q)))
                assert np.array_equal(got, ob.deblock(v, q)), f"shape {shape} q {q}"
        # the drop-in entry point on a decoded chunk
        ch = mc_amd.synth_chunks([3], [0.419])[0]
        dec = ob.decode_chunk(ob.encode_chunk(ch, 16.0)).reshape(256, 256, 256)
        got = dec.copy()
        mc_amd.api().mc_deblock.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float]
        mc_amd.api().mc_deblock(ob.p(got), 256, 256, 256, C.c_float(16.0))
        want = ob.deblock(dec, 16.0)
        assert np.array_equal(got, want) and not np.array_equal(want, dec)
        assert not np.any(got[dec == 0])      # air stays air
    finally:
        ctx.close()
