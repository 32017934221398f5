"""ctypes binding of the CPU oracle (oracle/liboracle.so) and, where it was
built in this container, the reference codec compiled from its own sources
(oracle/_ref/libmcref.so).  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libmcref.so")
ENC_CAP = 4096 * 4 + 1024

_o = None
_r = None


def oracle():
    global _o
    if _o is not None:
        return _o
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError(f"{ORACLE_SO} missing: run make -C oracle")
    O = C.CDLL(ORACLE_SO)
    vp = C.c_void_p
    O.orc_enc_block.argtypes = [C.c_float, C.c_int, vp, vp, C.c_size_t]
    O.orc_enc_block.restype = C.c_uint32
    O.orc_enc_block_ex.argtypes = [C.c_float, C.c_int, vp, vp, C.c_size_t, vp, vp, vp]
    O.orc_enc_block_ex.restype = C.c_uint32
    O.orc_dec_block.argtypes = [C.c_float, vp, C.c_uint32, vp]
    O.orc_dec_block_levels.argtypes = [C.c_float, vp, C.c_uint32, vp, vp, vp, vp, vp, vp]
    O.orc_encode_chunk.argtypes = [C.c_float, C.c_int, vp, vp, C.c_size_t]
    O.orc_encode_chunk.restype = C.c_size_t
    O.orc_decode_chunk.argtypes = [vp, C.c_uint64, vp]
    O.orc_decode_blob_block.argtypes = [vp, C.c_uint64, C.c_int, C.c_int, C.c_int, vp]
    O.orc_blob_len.argtypes = [vp]
    O.orc_blob_len.restype = C.c_uint64
    O.orc_xxh64.argtypes = [vp, C.c_size_t, C.c_uint64]
    O.orc_xxh64.restype = C.c_uint64
    O.orc_blob_hash.argtypes = [vp, C.c_uint64]
    O.orc_blob_hash.restype = C.c_uint64
    O.orc_enc_fracmap.argtypes = [vp, vp, C.c_size_t]
    O.orc_enc_fracmap.restype = C.c_uint32
    O.orc_dec_fracmap.argtypes = [vp, C.c_uint32, vp]
    O.orc_dct3_fwd.argtypes = [vp, vp]
    O.orc_dct3_inv.argtypes = [vp, vp]
    O.orc_dct_matrix.argtypes = [vp]
    O.orc_scan_order.argtypes = [vp]
    O.orc_step_table.argtypes = [C.c_float, vp, vp]
    O.orc_priors_for_q.argtypes = [C.c_float, vp]
    O.orc_set_priors.argtypes = [vp, vp]
    O.orc_rc_encode_ops.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.c_size_t]
    O.orc_rc_encode_ops.restype = C.c_size_t
    O.orc_rc_decode_ops.argtypes = [C.c_int, vp, vp, vp, vp, C.c_size_t, vp]
    O.orc_decode_chunks_mt.argtypes = [C.c_int, vp, vp, vp, C.c_int]
    O.orc_encode_chunks_mt.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, C.c_int]
    O.orc_decimate.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp]
    O.orc_deblock.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_float]
    O.orc_pick_lod.argtypes = [C.c_int, C.c_float]
    O.orc_pick_lod.restype = C.c_int
    O.orc_quad_spacing.argtypes = [vp, C.c_int, C.c_int]
    O.orc_quad_spacing.restype = C.c_float
    O.orc_render_steps.argtypes = [C.c_float, C.c_float, C.c_float]
    O.orc_render_steps.restype = C.c_int
    O.orc_quad_point.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, vp, vp]
    O.orc_render_quad_lod.argtypes = [C.c_int, vp, vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                      C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, vp]
    O.orc_render_quad_lod.restype = C.c_int
    _o = O
    return O


class OrcLevel(C.Structure):
    _fields_ = [("dense", C.c_void_p), ("blk_off", C.c_void_p), ("blk_data", C.c_void_p),
                ("nz", C.c_int), ("ny", C.c_int), ("nx", C.c_int)]


def decimate(vol):
    dz, dy, dx = vol.shape
    out = np.zeros(((dz + 1) // 2, (dy + 1) // 2, (dx + 1) // 2), np.uint8)
    v = np.ascontiguousarray(vol)
    oracle().orc_decimate(p(v), dz, dy, dx, p(out))
    return out


def pyramid(vol, nlods):
    levels = [np.ascontiguousarray(vol)]
    for _ in range(nlods - 1):
        levels.append(decimate(levels[-1]))
    return levels


def render_quad_lod(levels, grid, x0, y0, step, w, h, filt=1, comp=2, t0=-4.0, t1=4.0, dt=1.0, blocks=None):
    """Oracle quad render over dense levels (or, with blocks=[(off, data)] per
    level, through 16^3 block tables).  Returns (image, level rendered)."""
    arr = (OrcLevel * len(levels))()
    keep = []
    for i, v in enumerate(levels):
        nz, ny, nx = v.shape
        if blocks is None:
            arr[i] = OrcLevel(v.ctypes.data, None, None, nz, ny, nx)
        else:
            off, data = blocks[i]
            keep += [off, data]
            arr[i] = OrcLevel(None, off.ctypes.data, data.ctypes.data, nz, ny, nx)
    g = np.ascontiguousarray(grid, np.float32)
    out = np.zeros((h, w), np.uint8)
    L = oracle().orc_render_quad_lod(len(levels), arr, p(g), g.shape[1], g.shape[0], x0, y0, step, w, h, filt,
                                     comp, t0, t1, dt, p(out))
    return out, L


def p(a):
    return a.ctypes.data_as(C.c_void_p)


def enc_block(vox, q, tau=0):
    vox = np.ascontiguousarray(vox, np.uint8).ravel()
    out = np.zeros(ENC_CAP, np.uint8)
    n = oracle().orc_enc_block(q, tau, p(vox), p(out), out.size)
    return out[:n].tobytes()


def dec_block(payload, q):
    a = np.frombuffer(payload, np.uint8) if len(payload) else np.zeros(1, np.uint8)
    out = np.zeros(4096, np.uint8)
    oracle().orc_dec_block(q, p(a), len(payload), p(out))
    return out


def encode_chunk(chunk, q, tau=0):
    chunk = np.ascontiguousarray(chunk, np.uint8)
    out = np.zeros(20 << 20, np.uint8)
    n = oracle().orc_encode_chunk(q, tau, p(chunk), p(out), out.size)
    return out[:n].tobytes()


def decode_chunk(blob):
    a = np.frombuffer(blob, np.uint8)
    out = np.zeros((256, 256, 256), np.uint8)
    oracle().orc_decode_chunk(p(a), len(blob), p(out))
    return out


def xxh64(data, seed=0):
    a = np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray) else np.ascontiguousarray(data)
    if a.size == 0:
        a = np.zeros(1, np.uint8)
        return oracle().orc_xxh64(p(a), 0, seed)
    return oracle().orc_xxh64(p(a), a.nbytes, seed)


# ---- the reference build (only in the container that holds its sources) ----
class _McBuf(C.Structure):
    _fields_ = [("p", C.c_void_p), ("len", C.c_size_t), ("cap", C.c_size_t)]


REF_RENDER_SO = os.path.join(ROOT, "oracle", "_ref", "libmcref_render.so")
_rr = None


def ref_render_available():
    return os.path.exists(REF_RENDER_SO)


def ref_render():
    """The reference's renderer and one-shot builder (oracle/ref_render_shim.c)."""
    global _rr
    if _rr is not None:
        return _rr
    R = C.CDLL(REF_RENDER_SO)
    vp = C.c_void_p
    R.refshim_render_quad_lod.argtypes = [C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                          C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                          C.c_int, vp, C.c_int]
    R.refshim_render_quad_lod.restype = C.c_int
    R.refshim_pick_lod.argtypes = [C.c_int, C.c_float]
    R.refshim_pick_lod.restype = C.c_int
    R.refshim_quad_spacing.argtypes = [vp, C.c_int, C.c_int]
    R.refshim_quad_spacing.restype = C.c_float
    R.refshim_build.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_size_t)]
    R.refshim_build.restype = vp
    R.refshim_free.argtypes = [vp]
    _rr = R
    return R


def ref_render_quad_lod(levels, grid, x0, y0, step, w, h, filt=1, comp=2, t0=-4.0, t1=4.0, dt=1.0, blocked=0):
    R = ref_render()
    vols = (C.c_void_p * len(levels))(*[v.ctypes.data for v in levels])
    dims = np.array([d for v in levels for d in v.shape], np.int32)
    g = np.ascontiguousarray(grid, np.float32)
    out = np.zeros((h, w), np.uint8)
    rc = R.refshim_render_quad_lod(len(levels), vols, p(dims), p(g), g.shape[1], g.shape[0], x0, y0, step, w, h,
                                   filt, comp, t0, t1, dt, blocked, p(out), 1)
    assert rc == 0
    return out


def ref_build(vol, q):
    """The reference's mc_build of a dense volume: archive bytes."""
    R = ref_render()
    n = C.c_size_t()
    v = np.ascontiguousarray(vol)
    ptr = R.refshim_build(p(v), v.shape[0], v.shape[1], v.shape[2], q, C.byref(n))
    assert ptr
    data = C.string_at(ptr, n.value)
    R.refshim_free(ptr)
    return data


def default_priors():
    """The trained q=1 and q=12 endpoint tables, u16 [8, 32] each: the
    interpolated tables at q <= 1 (weight 0) and at q >= 2^3.585 (weight 1,
    src/mc_codec.c:156-167) are the endpoints themselves.  Restores the
    oracle's default tables first."""
    oracle().orc_set_priors(None, None)
    lo = np.zeros((8, 32), np.uint16)
    hi = np.zeros((8, 32), np.uint16)
    oracle().orc_priors_for_q(C.c_float(0.5), p(lo))
    oracle().orc_priors_for_q(C.c_float(16.0), p(hi))
    return lo, hi


def custom_priors(seed=4242, spread=300):
    """Deterministic non-default endpoint tables: the trained tables moved by
    up to +-spread (seeded), kept inside [32, 4064]."""
    lo, hi = default_priors()
    rng = np.random.default_rng(seed)
    lo = np.clip(lo.astype(np.int64) + rng.integers(-spread, spread + 1, lo.shape), 32, 4064).astype(np.uint16)
    hi = np.clip(hi.astype(np.int64) + rng.integers(-spread, spread + 1, hi.shape), 32, 4064).astype(np.uint16)
    return lo, hi


def set_priors(lo=None, hi=None):
    """Override the oracle's endpoint tables (None restores the defaults)."""
    if lo is None:
        oracle().orc_set_priors(None, None)
    else:
        a, b = np.ascontiguousarray(lo, np.uint16), np.ascontiguousarray(hi, np.uint16)
        oracle().orc_set_priors(p(a), p(b))


def deblock(vol, q):
    """The oracle's mc_deblock of a copy of vol (z, y, x)."""
    v = np.ascontiguousarray(vol, np.uint8).copy()
    oracle().orc_deblock(p(v), v.shape[0], v.shape[1], v.shape[2], C.c_float(q))
    return v


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _r
    if _r is not None:
        return _r
    R = C.CDLL(REF_SO)
    vp = C.c_void_p
    R.mc_codec_ctx_new.restype = vp
    R.mc_codec_ctx_free.argtypes = [vp]
    R.mc_codec_ctx_set_quality.argtypes = [vp, C.c_float]
    R.mc_codec_ctx_set_max_error.argtypes = [vp, C.c_int]
    R.mc_apply_preset.argtypes = [vp, C.c_int]
    R.mc_apply_preset.restype = C.c_float
    R.mc_enc_block.argtypes = [vp, vp, C.POINTER(_McBuf), C.POINTER(C.c_uint32)]
    R.mc_dec_block.argtypes = [vp, vp, C.c_uint32, vp]
    R.mc_enc_fracmap.argtypes = [vp, vp, C.c_size_t]
    R.mc_enc_fracmap.restype = C.c_uint32
    R.mc_xxh64.argtypes = [vp, C.c_size_t, C.c_uint64]
    R.mc_xxh64.restype = C.c_uint64
    R.mc_archive_open_dims.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_float]
    R.mc_archive_open_dims.restype = vp
    R.mc_archive_close.argtypes = [vp]
    R.mc_archive_append_chunk_par.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_float, C.c_int]
    R.mc_archive_append_chunk_ctx.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]
    R.mc_archive_chunk_offset.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    R.mc_archive_chunk_offset.restype = C.c_uint64
    R.mc_archive_decode_chunk.argtypes = [vp, C.c_uint64, vp, C.c_int]
    R.mc_archive_decode_block.argtypes = [vp, C.c_uint64, C.c_int, C.c_int, C.c_int, vp]
    R.mc_archive_read_regions.argtypes = [vp, C.c_int, vp, C.c_int, C.c_long, C.c_long, C.c_long, vp,
                                          C.c_size_t, C.c_int]
    R.mc_archive_sample_boxes.argtypes = [vp, C.c_int, C.c_uint64, C.c_int, C.c_long, C.c_long, C.c_long,
                                          C.c_float, vp]
    R.mc_archive_sample_boxes.restype = C.c_int
    R.mc_archive_set_priors.argtypes = [vp, vp, vp]
    R.mc_archive_set_priors.restype = C.c_int
    R.mc_codec_set_priors.argtypes = [vp, vp]
    R.mc_deblock.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_float]
    R.mc_codec_init()
    _r = R
    return R
