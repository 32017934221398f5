"""mchip_decode_chunks_static, the decode entry point that allocates nothing,
against mchip_decode_chunks (the allocating entry point it mirrors) and the
oracle: two calls in a row, the second with a smaller input, then the calls
it must refuse.  Runs alone in a fresh process and after the other tests."""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob


def _resident(ctx, blobs):
    import mc_amd
    offs = np.zeros(len(blobs), np.uint64)
    tot = 0
    for k, b in enumerate(blobs):
        offs[k] = tot
        tot += (len(b) + 255) // 256 * 256
    buf = np.zeros(max(tot, 256), np.uint8)
    for k, b in enumerate(blobs):
        buf[int(offs[k]):int(offs[k]) + len(b)] = np.frombuffer(b, np.uint8)
    d = mc_amd.DeviceBuffer(buf.size)
    ctx.h2d(d, buf)
    lens = np.array([len(b) for b in blobs], np.uint64)
    qs = np.array([np.frombuffer(b[:4], np.float32)[0] for b in blobs], np.float32)
    return d, offs, lens, qs


def _static(L, d_blobs, offs, lens, qs, d_out, stream):
    return L.mchip_decode_chunks_static(C.c_void_p(d_blobs.ptr), ob.p(offs), ob.p(lens), ob.p(qs), len(offs),
                                        C.c_void_p(d_out.ptr), 16777216, stream)


@pytest.mark.gpu
def test_static_decode_equals_allocating_decode_twice():
    import mc_amd
    L = mc_amd.load()
    ctx = mc_amd.Ctx(0)
    stream = L.mchip_ctx_stream(ctx.h)
    ch = mc_amd.synth_chunks([41, 42, 43], [0.5, 0.6, 0.8])
    # first call: two chunks (one at a quality no other call has used, one
    # with corrections) that fit the fixed workspace; second call: one
    # smaller chunk
    calls = [[ob.encode_chunk(ch[0], 6.5), ob.encode_chunk(ch[2], 16.0, 32)],
             [ob.encode_chunk(ch[1], 32.0, 64)]]
    assert 32 * sum(len(b) for b in calls[0]) + (2 << 20) < 128 << 20
    assert len(calls[1][0]) < min(len(b) for b in calls[0])
    d_ref = mc_amd.DeviceBuffer(3 * mc_amd.CHUNK_BYTES)
    d_out = mc_amd.DeviceBuffer(3 * mc_amd.CHUNK_BYTES)
    try:
        for k, blobs in enumerate(calls):
            n = len(blobs)
            d_blobs, offs, lens, qs = _resident(ctx, blobs)
            print(f"static decode call {k}: {n} chunks", flush=True)
            ctx.h2d(d_out, np.full(n * mc_amd.CHUNK_BYTES, 0x5A, np.uint8))   # stale output must not survive
            mc_amd.check(_static(L, d_blobs, offs, lens, qs, d_out, stream))
            ctx.decode_chunks(d_blobs, offs, lens, qs, d_ref)
            ctx.sync()
            got = np.empty((n, 256, 256, 256), np.uint8)
            want = np.empty_like(got)
            ctx.d2h(got, d_out)
            ctx.d2h(want, d_ref)
            for j in range(n):
                assert np.array_equal(got[j], want[j]), f"call {k} chunk {j}: static differs from allocating"
                assert np.array_equal(got[j], ob.decode_chunk(blobs[j]).reshape(got[j].shape)), \
                    f"call {k} chunk {j}: differs from the oracle"
            d_blobs.free()
    finally:
        d_ref.free()
        d_out.free()
        ctx.close()


@pytest.mark.gpu
def test_static_decode_refuses_what_does_not_fit():
    import mc_amd
    L = mc_amd.load()
    ctx = mc_amd.Ctx(0)
    d = mc_amd.DeviceBuffer(1 << 20)
    try:
        one = np.zeros(1, np.uint64)
        q = np.full(8, 6.0, np.float32)
        offs = np.zeros(8, np.uint64)
        lens = np.full(8, 1000, np.uint64)
        # more than 4 chunks per call
        assert L.mchip_decode_chunks_static(C.c_void_p(d.ptr), ob.p(offs), ob.p(lens), ob.p(q), 5,
                                            C.c_void_p(d.ptr), 16777216, None) != 0
        assert "at most 4" in mc_amd.last_error()
        # blobs whose lists would not fit the fixed workspace (4 x 2 MB blobs)
        big = np.full(4, 2_000_000, np.uint64)
        assert L.mchip_decode_chunks_static(C.c_void_p(d.ptr), ob.p(offs), ob.p(big), ob.p(q), 4,
                                            C.c_void_p(d.ptr), 16777216, None) != 0
        assert "fixed workspace" in mc_amd.last_error()
        # an output stride below one chunk
        assert L.mchip_decode_chunks_static(C.c_void_p(d.ptr), ob.p(one), ob.p(lens), ob.p(q), 1,
                                            C.c_void_p(d.ptr), 4096, None) != 0
    finally:
        d.free()
        ctx.close()
