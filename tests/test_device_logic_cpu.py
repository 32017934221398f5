"""The device coder logic (mchip_device.hpp / mchip_codec.hpp, the exact code
the HIP kernels run) compiled for the CPU through a host shim, checked
against the oracle: decoding of the reference's known-answer payloads and of
corrupted ones, encoding back to the reference's bytes, and the line
transforms."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_bind as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_h = None


def harness():
    global _h
    if _h is not None:
        return _h
    d = tempfile.mkdtemp(prefix="mchip_dl_")
    so = os.path.join(d, "dev_logic.so")
    src = os.path.join(ROOT, "tests", "cpu_harness", "dev_logic.cpp")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
           "-Wno-unknown-pragmas", "-I", os.path.join(ROOT, "tests", "cpu_harness"),
           "-I", os.path.join(ROOT, "matter-compressor_amd", "csrc"), src, "-o", so]
    subprocess.run(cmd, check=True)
    H = C.CDLL(so)
    vp = C.c_void_p
    H.dl_decode.argtypes = [vp, vp, vp, vp, C.c_uint32, vp, vp, vp, vp, vp, vp]
    H.dl_decode.restype = C.c_int
    H.dl_decode_levels_full.argtypes = [vp, vp, vp, vp, C.c_uint32, vp, vp, vp]
    H.dl_decode_levels_full.restype = C.c_uint32
    H.dl_encode.argtypes = [vp, vp, C.c_uint32, C.c_uint32, vp, C.c_uint32, vp, vp, C.c_uint32]
    H.dl_encode.restype = C.c_uint32
    H.dl_encode_corrections.argtypes = [vp, C.c_uint32, C.c_uint32, vp, vp, vp, vp, C.c_uint32]
    H.dl_encode_corrections.restype = C.c_uint32
    H.dl_idct_line.argtypes = [vp, vp]
    H.dl_idct_line_sparse.argtypes = [vp, vp, vp]
    H.dl_idct_line_sparse.restype = C.c_int
    H.dl_fdct_line.argtypes = [vp, vp]
    H.dl_set_variant.argtypes = [C.c_int]
    _h = H
    return H


QTAB = np.dtype([("step", "<f4", 46), ("rstep", "<f4", 46), ("sig", "<u2", 32), ("mag", "<u2", 12),
                 ("eob", "<u2", 14), ("flag", "<u2", 2), ("dc", "<u2", 8)])
assert QTAB.itemsize == 504


def tables():
    O = ob.oracle()
    scan = np.zeros(4096, np.uint16)
    O.orc_scan_order(ob.p(scan))
    s = scan.astype(int)
    l1 = (s >> 8) + ((s >> 4) & 15) + (s & 15)
    band_start = np.array([np.searchsorted(l1, 6 * b) for b in range(8)] + [4096], np.uint16)
    return scan, band_start


def qtab(q):
    O = ob.oracle()
    q = O.orc_sanitize_q(C.c_float(q)) if False else (q if (q > 0 and q <= 1024) else 8.0)
    st = np.zeros(4096, np.float32)
    rs = np.zeros(4096, np.float32)
    O.orc_step_table(C.c_float(q), ob.p(st), ob.p(rs))
    pri = np.zeros((8, 32), np.uint16)
    O.orc_priors_for_q(C.c_float(q), ob.p(pri))
    t = np.zeros(1, QTAB)
    for l in range(46):
        z = min(l, 15)
        y = min(l - z, 15)
        x = l - z - y
        t["step"][0, l] = st[z * 256 + y * 16 + x]
        t["rstep"][0, l] = rs[z * 256 + y * 16 + x]
    t["sig"][0] = pri[0]
    t["mag"][0] = pri[1][:12]
    t["eob"][0] = pri[2][:14]
    t["flag"][0] = pri[6][:2]
    t["dc"][0] = pri[7][:8]
    return t


def oracle_levels(payload, q):
    lv = np.zeros(4096, np.int16)
    dc, fl, nc = C.c_int(), C.c_int(), C.c_uint32()
    cp = np.zeros(4096, np.uint32)
    cd = np.zeros(4096, np.int32)
    a = np.frombuffer(payload, np.uint8) if payload else np.zeros(1, np.uint8)
    ob.oracle().orc_dec_block_levels(C.c_float(q), ob.p(a), len(payload), ob.p(lv), C.byref(dc), C.byref(fl),
                                     C.byref(nc), ob.p(cp), ob.p(cd))
    return lv, dc.value, fl.value, cp[: nc.value].copy(), cd[: nc.value].copy()


def device_decode(payload, q, scan, band_start):
    H = harness()
    t = qtab(q)
    lead = len(payload) % 4                      # exercise unaligned payload starts
    buf = np.zeros(len(payload) + 32, np.uint8)
    buf[lead:lead + len(payload)] = np.frombuffer(payload, np.uint8)
    pp = C.c_void_p(buf.ctypes.data + lead)
    lv = np.zeros(4096, np.int16)
    dc, fl, nc = C.c_int(), C.c_int(), C.c_uint32()
    cp = np.zeros(4096, np.uint32)
    cd = np.zeros(4096, np.int32)
    st = H.dl_decode(ob.p(t), ob.p(band_start), ob.p(scan), pp, len(payload), ob.p(lv), C.byref(dc),
                     C.byref(fl), C.byref(nc), ob.p(cp), ob.p(cd))
    if st != 0:   # list overflow: full decode (corrections are then checked by the GPU tests)
        H.dl_decode_levels_full(ob.p(t), ob.p(band_start), ob.p(scan), pp, len(payload), ob.p(lv),
                                C.byref(dc), C.byref(fl))
        return lv, dc.value, fl.value, None, None
    return lv, dc.value, fl.value, cp[: nc.value].copy(), cd[: nc.value].copy()


@pytest.fixture(params=[2, 0], ids=["lane", "per_position"])
def decoder_variant(request):
    harness().dl_set_variant(request.param)
    yield request.param
    harness().dl_set_variant(2)


def test_device_decoder_matches_oracle_on_reference_kat(kat, decoder_variant):
    scan, bs = tables()
    for i in range(len(kat["q"])):
        pay = kat["payload"][kat["offs"][i]:kat["offs"][i + 1]].tobytes()
        if not pay:
            continue
        want = oracle_levels(pay, float(kat["q"][i]))
        got = device_decode(pay, float(kat["q"][i]), scan, bs)
        assert np.array_equal(got[0], want[0]) and got[1:3] == want[1:3], f"case {i}"
        assert np.array_equal(got[3], want[3]) and np.array_equal(got[4], want[4]), f"case {i} corrections"


def test_device_decoder_matches_oracle_on_corrupt_payloads(kat, decoder_variant):
    scan, bs = tables()
    rng = np.random.default_rng(21)
    for trial in range(400):
        i = int(rng.integers(0, len(kat["q"])))
        p = bytearray(kat["payload"][kat["offs"][i]:kat["offs"][i + 1]].tobytes() or b"\x00")
        for _ in range(int(rng.integers(1, 5))):
            p[int(rng.integers(0, len(p)))] ^= int(rng.integers(1, 256))
        if trial % 3 == 0:
            p = p[: int(rng.integers(0, len(p) + 1))]
        p = bytes(p)
        q = float(kat["q"][i])
        want = oracle_levels(p, q)
        got = device_decode(p, q, scan, bs)
        assert np.array_equal(got[0], want[0]) and got[1:3] == want[1:3], f"trial {trial}"
        if got[3] is not None:
            assert np.array_equal(got[3], want[3]) and np.array_equal(got[4], want[4]), f"trial {trial}"


def test_device_encoder_reproduces_reference_payloads(kat):
    H = harness()
    scan, bs = tables()
    for i in range(len(kat["q"])):
        pay = kat["payload"][kat["offs"][i]:kat["offs"][i + 1]].tobytes()
        if not pay:
            continue
        q = float(kat["q"][i])
        lv, dc, fl, cp, cd = oracle_levels(pay, q)
        L = lv[scan.astype(int)].astype(np.int16)
        nz = np.nonzero(L)[0]
        eob = int(nz[-1]) + 1 if len(nz) else 0
        Cw = (cp.astype(np.uint32) | ((cd.astype(np.int32) & 0xFFFF).astype(np.uint32) << 16)).astype(np.uint32)
        if Cw.size == 0:
            Cw = np.zeros(1, np.uint32)
        out = np.zeros(17408, np.uint8)
        t = qtab(q)
        n = H.dl_encode(ob.p(t), ob.p(bs), dc, eob, ob.p(L), len(cp), ob.p(Cw), ob.p(out), out.size)
        assert out[:n].tobytes() == pay, f"case {i} {kat['names'][i]}"
        # a capacity below the payload: the first cap bytes are stored, nothing
        # past them, and the full length is still returned
        for cap in (4, 8, 12, 20, 36, 64, 68):
            if cap >= len(pay):
                continue
            small = np.full(len(pay) + 64, 0xAB, np.uint8)
            m = H.dl_encode(ob.p(t), ob.p(bs), dc, eob, ob.p(L), len(cp), ob.p(Cw), ob.p(small), cap)
            assert m == len(pay), f"case {i} cap {cap}: length"
            assert small[:cap].tobytes() == pay[:cap], f"case {i} cap {cap}: stored prefix"
            assert (small[cap:] == 0xAB).all(), f"case {i} cap {cap}: bytes past the capacity"


def test_device_line_transforms_match_oracle_dct():
    H = harness()
    cm = np.zeros(256, np.float32)
    ob.oracle().orc_dct_matrix(ob.p(cm))
    rng = np.random.default_rng(4)
    for trial in range(20):
        blk = rng.normal(0, 50, 4096).astype(np.float32)
        if trial % 4 == 0:
            blk[rng.random(4096) < 0.9] = 0.0
        want = np.zeros_like(blk)
        ob.oracle().orc_dct3_inv(ob.p(blk), ob.p(want))
        got = blk.reshape(16, 16, 16).copy()
        for axis in (2, 1, 0):                       # x pass, then y, then z
            g = np.moveaxis(got, axis, -1).copy()
            for idx in np.ndindex(16, 16):
                line = np.ascontiguousarray(g[idx])
                H.dl_idct_line(ob.p(line), ob.p(cm))
                g[idx] = line
            got = np.moveaxis(g, -1, axis)
        assert np.array_equal(got.ravel().view(np.uint32), want.view(np.uint32)), f"inverse trial {trial}"
        want_f = np.zeros_like(blk)
        ob.oracle().orc_dct3_fwd(ob.p(blk), ob.p(want_f))
        got = blk.reshape(16, 16, 16).copy()
        for axis in (2, 1, 0):
            g = np.moveaxis(got, axis, -1).copy()
            for idx in np.ndindex(16, 16):
                line = np.ascontiguousarray(g[idx])
                H.dl_fdct_line(ob.p(line), ob.p(cm))
                g[idx] = line
            got = np.moveaxis(g, -1, axis)
        assert np.array_equal(got.ravel().view(np.uint32), want_f.view(np.uint32)), f"forward trial {trial}"


def test_crafted_correction_payloads_decode_like_the_reference():
    """Block payloads with hand-made correction lists (dl_encode_corrections)
    decode in the oracle exactly as in the reference build: repeated
    positions apply one after another, each clamped."""
    if not ob.ref_available():
        pytest.skip("reference build not present")
    H = harness()
    R = ob.ref()
    for gaps, negs, mags, dc, q in [([5, 0, 0, 3, 0], [0, 0, 1, 1, 0], [200, 100, 150, 90, 7], 128, 6.0),
                                    ([0, 0, 4095], [1, 0, 0], [1, 300, 2], 3, 1.0)]:
        t = qtab(q)
        n = len(gaps)
        out = np.zeros(4096, np.uint8)
        k = H.dl_encode_corrections(ob.p(t), dc, n, ob.p(np.array(gaps, np.uint32)), ob.p(np.array(negs, np.int32)),
                                    ob.p(np.array(mags, np.uint32)), ob.p(out), out.size)
        pay = out[:k].tobytes()
        ctx = R.mc_codec_ctx_new()
        R.mc_codec_ctx_set_quality(ctx, C.c_float(q))
        want = np.zeros(4096, np.uint8)
        R.mc_dec_block(ctx, pay, len(pay), ob.p(want))
        R.mc_codec_ctx_free(ctx)
        assert np.array_equal(ob.dec_block(pay, q), want)
        if dc == 128:   # voxel 5: 128 +200 -> 255, +100 -> 255, -150 -> 105; voxel 8: 128 -90 +7 = 45
            assert want[5] == 105 and want[8] == 45 and want[0] == 128


def test_wave_sparse_inverse_transform_matches_oracle_under_any_wave_mask():
    """The reconstruction's inverse transform skips an input only when no line
    of the wave needs it, so a line also runs inputs that are zero for it
    (another lane's). With the masks of such a wave (every nonzero input of
    the line, plus random zero inputs), and all-zero lines skipped, the three
    passes equal the oracle's inverse DCT bit for bit, including sums that
    cancel to exactly zero before a masked-in zero input is added."""
    H = harness()
    cm = np.zeros(256, np.float32)
    ob.oracle().orc_dct_matrix(ob.p(cm))
    cmm = cm.reshape(16, 16)
    # pairs of same-parity inputs whose weights cancel for some output
    cancel = [(k, j) for n in range(8) for k in range(16) for j in range(k + 1, 16)
              if (k & 1) == (j & 1) and cmm[k, n] == -cmm[j, n]]
    assert cancel, "the DCT matrix has no cancelling pair to exercise"
    rng = np.random.default_rng(11)
    step = np.float32(3.25)
    for trial in range(24):
        lv = rng.integers(-40, 41, 4096).astype(np.float32)
        lv[rng.random(4096) < (0.97 if trial % 3 else 0.6)] = 0.0
        blk = (lv * step).astype(np.float32)
        if trial % 4 == 1:                       # exact cancellations in the x pass
            c3 = blk.reshape(16, 16, 16)
            for z, y in zip(rng.integers(0, 16, 40), rng.integers(0, 16, 40)):
                k, j = cancel[int(rng.integers(len(cancel)))]
                c3[z, y, :] = 0.0
                c3[z, y, k] = c3[z, y, j] = step * 7
        want = np.zeros_like(blk)
        ob.oracle().orc_dct3_inv(ob.p(blk), ob.p(want))
        got = blk.reshape(16, 16, 16).copy()
        extra_p = (0.0, 0.3, 1.0)[trial % 3]
        for axis in (2, 1, 0):
            g = np.moveaxis(got, axis, -1).copy()
            for idx in np.ndindex(16, 16):
                line = np.ascontiguousarray(g[idx])
                wk = ((line != 0) | (rng.random(16) < extra_p)).astype(np.uint64)
                if not H.dl_idct_line_sparse(ob.p(line), ob.p(cm), ob.p(wk)):
                    assert not line.any()
                g[idx] = line
            got = np.moveaxis(g, -1, axis)
        assert np.array_equal(got.ravel().view(np.uint32), want.view(np.uint32)), f"trial {trial}"
