"""ctypes binding of libmc_amd.so (the MI355X codec engine and the plain-C
matter-compressor drop-in API) and libmc_synth.so (workload generator).

Loading fails loudly when the built library is missing: there is no CPU
fallback for any codec entry point.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
CHUNK_BYTES = 256 * 256 * 256

_lib = None
_synth = None


class MchipError(RuntimeError):
    pass


def lib_path():
    """The in-tree build, or the library named by MC_AMD_LIB."""
    return os.environ.get("MC_AMD_LIB") or os.path.join(LIB_DIR, "libmc_amd.so")


def load():
    """Load libmc_amd.so and declare its C ABI (idempotent)."""
    global _lib
    if _lib is not None:
        return _lib
    p = lib_path()
    if not os.path.exists(p):
        raise MchipError(f"{p} is not built: run python matter-compressor_amd/build.py")
    L = C.CDLL(p)
    vp, u8p, i = C.c_void_p, C.c_void_p, C.c_int
    sig = {
        "mchip_init": (i, [i]),
        "mchip_last_error": (C.c_char_p, []),
        "mchip_set_priors": (i, [vp, vp]),
        "mchip_set_tuning": (i, [C.c_char_p, i]),
        "mchip_ctx_create": (vp, []),
        "mchip_ctx_destroy": (None, [vp]),
        "mchip_ctx_stream": (vp, [vp]),
        "mchip_ctx_sync": (i, [vp]),
        "mchip_malloc": (vp, [C.c_size_t]),
        "mchip_mem_info": (i, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
        "mchip_free": (None, [vp]),
        "mchip_h2d": (i, [vp, vp, vp, C.c_size_t]),
        "mchip_d2h": (i, [vp, vp, vp, C.c_size_t]),
        "mchip_memset": (i, [vp, vp, i, C.c_size_t]),
        "mchip_d2d": (i, [vp, vp, vp, C.c_size_t]),
        "mchip_decode_graph_create": (vp, [vp, vp, vp, vp, vp, i, vp, C.c_uint64]),
        "mchip_graph_launch": (i, [vp]),
        "mchip_encode_graph_create": (vp, [vp, vp, C.c_uint64, vp, vp, vp, i, vp, vp, vp]),
        "mchip_encode_graph_results": (i, [vp, vp, vp]),
        "mchip_graph_destroy": (None, [vp]),
        "mchip_autotune_decode": (i, [vp, vp, vp, vp, vp, i, vp, C.c_uint64, C.c_char_p, C.c_size_t]),
        "mchip_autotune_encode": (i, [vp, vp, C.c_uint64, vp, vp, vp, i, vp, vp, vp, C.c_char_p, C.c_size_t]),
        "mchip_store_create": (vp, [vp, C.c_uint64, vp, vp, vp, i, i, i]),
        "mchip_store_destroy": (None, [vp]),
        "mchip_store_device_bytes": (C.c_uint64, [vp]),
        "mchip_server_create": (vp, [vp, C.c_uint64, i]),
        "mchip_server_submit": (C.c_int64, [vp, vp, i, C.c_int64, C.c_int64, C.c_int64, vp, C.c_uint64]),
        "mchip_server_poll": (i, [vp, C.c_int64]),
        "mchip_server_wait": (i, [vp, C.c_int64, C.POINTER(C.c_double)]),
        "mchip_server_get_stats": (i, [vp, vp]),
        "mchip_server_preload": (i, [vp, C.POINTER(C.c_uint64)]),
        "mchip_server_resize": (i, [vp, C.c_uint64]),
        "mchip_server_update_blocks": (i, [vp, vp, C.c_uint64, i, vp, vp, vp, vp]),
        "mchip_server_destroy": (None, [vp]),
        "mchip_decode_chunks": (i, [vp, u8p, vp, vp, vp, i, u8p, C.c_uint64]),
        "mchip_decode_chunks_host": (i, [vp, vp, vp, i, u8p, C.c_uint64]),
        "mchip_decode_chunks_static": (i, [u8p, vp, vp, vp, i, u8p, C.c_uint64, vp]),
        "mchip_decode_chunk_merged": (i, [vp, vp, C.c_uint64, vp]),
        "mchip_merge_stats": (i, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "mchip_deblock": (i, [vp, vp, i, i, i, C.c_float]),
        "mchip_deblock_host": (i, [vp, vp, i, i, i, C.c_float]),
        "mchip_decode_blocks_host": (i, [vp, vp, vp, vp, i, u8p]),
        "mchip_encode_chunks": (i, [vp, u8p, C.c_uint64, vp, vp, vp, i, u8p, vp, vp, vp, vp]),
        "mchip_encode_chunks_host": (i, [vp, vp, vp, vp, vp, i, vp, vp, vp, vp]),
        "mchip_xxh64_ranges": (i, [vp, vp, vp, vp, vp, i, vp]),
        "mchip_render_quad": (i, [vp, i, vp, vp, vp, i, i, C.c_float, C.c_float, C.c_float, i, i, i, i,
                                  C.c_float, C.c_float, C.c_float, vp, C.POINTER(C.c_int)]),
        "mchip_decimate": (i, [vp, vp, i, i, i, vp]),
        "mchip_decimate_chunks": (i, [vp, vp, vp]),
        "mchip_store_create_levels": (vp, [vp, C.c_uint64, i, vp, vp, vp, vp]),
        "mchip_server_submit_render": (C.c_int64, [vp, vp, i, i, C.c_float, C.c_float, C.c_float, i, i, i, i,
                                                   C.c_float, C.c_float, C.c_float, vp]),
        "mchip_server_render_level": (i, [vp, C.c_int64]),
        "mchip_verify_blobs": (C.c_long, [vp, vp, vp, vp, i, vp, vp]),
        "mchip_verify_blobs_host": (C.c_long, [vp, vp, vp, i, vp, vp]),
        "mchip_ctx_enable_timing": (i, [vp, i]),
        "mchip_ctx_kernel_time": (i, [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_long)]),
        "mchip_ctx_reset_timing": (i, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def mem_info():
    """(free, total) device memory in bytes."""
    f, t = C.c_size_t(), C.c_size_t()
    check(load().mchip_mem_info(C.byref(f), C.byref(t)))
    return f.value, t.value


def set_tuning(key, value):
    check(load().mchip_set_tuning(key.encode(), int(value)))


def last_error():
    return load().mchip_last_error().decode(errors="replace")


def check(rc):
    if rc != 0:
        raise MchipError(last_error())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _ptr_array(bufs):
    arr = (C.c_void_p * len(bufs))()
    for k, b in enumerate(bufs):
        arr[k] = None if b is None else b.ctypes.data
    return arr


class DeviceBuffer:
    """Raw device allocation owned from Python."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = load().mchip_malloc(self.nbytes)
        if not self.ptr:
            raise MchipError(load().mchip_last_error().decode())

    def free(self):
        if self.ptr:
            load().mchip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Ctx:
    """An mchip context: one HIP stream + its device workspace."""

    def __init__(self, device=0):
        L = load()
        check(L.mchip_init(device))
        self.h = L.mchip_ctx_create()
        if not self.h:
            raise MchipError(L.mchip_last_error().decode())

    def close(self):
        if self.h:
            load().mchip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(load().mchip_ctx_sync(self.h))

    def h2d(self, dbuf, arr, offset=0):
        a = np.ascontiguousarray(arr)
        check(load().mchip_h2d(self.h, C.c_void_p(dbuf.ptr + offset), _ptr(a), a.nbytes))

    def d2h(self, arr, dbuf, offset=0):
        check(load().mchip_d2h(self.h, _ptr(arr), C.c_void_p(dbuf.ptr + offset), arr.nbytes))

    def xxh64_ranges(self, buf, offs, lens, seeds):
        """XXH64 of byte ranges buf[offs[i] : offs[i]+lens[i]] of a device buffer."""
        offs, lens, seeds = (np.ascontiguousarray(x, np.uint64) for x in (offs, lens, seeds))
        out = np.zeros(len(offs), np.uint64)
        check(load().mchip_xxh64_ranges(self.h, C.c_void_p(buf.ptr), _ptr(offs), _ptr(lens), _ptr(seeds),
                                        len(offs), _ptr(out)))
        return out

    def verify_blobs(self, buf, offs, lens):
        """Blob hashes and bad flags of device-resident chunk blobs."""
        offs, lens = (np.ascontiguousarray(x, np.uint64) for x in (offs, lens))
        h = np.zeros(len(offs), np.uint64)
        bad = np.zeros(len(offs), np.uint8)
        n = load().mchip_verify_blobs(self.h, C.c_void_p(buf.ptr), _ptr(offs), _ptr(lens), len(offs), _ptr(h),
                                      _ptr(bad))
        if n < 0:
            raise MchipError(last_error())
        return n, h, bad

    def verify_blobs_host(self, blobs):
        """The same for host blobs (numpy uint8 arrays)."""
        arrs = [np.ascontiguousarray(b, np.uint8) for b in blobs]
        lens = np.array([a.size for a in arrs], np.uint64)
        h = np.zeros(len(arrs), np.uint64)
        bad = np.zeros(len(arrs), np.uint8)
        n = load().mchip_verify_blobs_host(self.h, _ptr_array(arrs), _ptr(lens), len(arrs), _ptr(h), _ptr(bad))
        if n < 0:
            raise MchipError(last_error())
        return n, h, bad

    def render_quad(self, levels, dims, grid, x0, y0, step, w, h, d_out, filt=1, comp=2, t0=-4.0, t1=4.0,
                    dt=1.0):
        """Quad render over dense device levels (DeviceBuffers); returns the level rendered."""
        g = np.ascontiguousarray(grid, np.float32)
        lv = (C.c_void_p * len(levels))(*[b.ptr for b in levels])
        dd = np.ascontiguousarray(dims, np.int32).reshape(-1)
        lod = C.c_int(-1)
        check(load().mchip_render_quad(self.h, len(levels), lv, _ptr(dd), _ptr(g), g.shape[1], g.shape[0], x0, y0,
                                       step, w, h, filt, comp, t0, t1, dt, C.c_void_p(d_out.ptr), C.byref(lod)))
        return lod.value

    def decimate(self, src, dz, dy, dx, dst):
        check(load().mchip_decimate(self.h, C.c_void_p(src.ptr), dz, dy, dx, C.c_void_p(dst.ptr)))

    def decimate_chunks(self, srcs, dst, dst_off=0):
        """srcs: 8 (buffer, byte offset) pairs or None, (z, y, x) order."""
        arr = (C.c_void_p * 8)(*[None if x is None else x[0].ptr + x[1] for x in srcs])
        check(load().mchip_decimate_chunks(self.h, arr, C.c_void_p(dst.ptr + dst_off)))

    def d2d(self, dst, src, nbytes, dst_off=0, src_off=0):
        check(load().mchip_d2d(self.h, C.c_void_p(dst.ptr + dst_off), C.c_void_p(src.ptr + src_off), nbytes))

    def enable_timing(self, on=True):
        check(load().mchip_ctx_enable_timing(self.h, 1 if on else 0))

    def reset_timing(self):
        check(load().mchip_ctx_reset_timing(self.h))

    def kernel_time(self, name):
        ms, n = C.c_double(), C.c_long()
        check(load().mchip_ctx_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # ---- decode ----------------------------------------------------------
    def decode_chunks(self, d_blobs, offs, lens, qs, d_out, out_stride=CHUNK_BYTES):
        n = len(offs)
        offs = np.ascontiguousarray(offs, np.uint64)
        lens = np.ascontiguousarray(lens, np.uint64)
        qs = np.ascontiguousarray(qs, np.float32)
        check(load().mchip_decode_chunks(self.h, C.c_void_p(d_blobs.ptr), _ptr(offs), _ptr(lens), _ptr(qs), n,
                                         C.c_void_p(d_out.ptr), out_stride))

    def decode_graph(self, d_blobs, offs, lens, qs, d_out, out_stride=CHUNK_BYTES):
        """HIP graph of one batched decode (replay with graph_launch)."""
        offs = np.ascontiguousarray(offs, np.uint64)
        lens = np.ascontiguousarray(lens, np.uint64)
        qs = np.ascontiguousarray(qs, np.float32)
        g = load().mchip_decode_graph_create(self.h, C.c_void_p(d_blobs.ptr), _ptr(offs), _ptr(lens), _ptr(qs),
                                             len(offs), C.c_void_p(d_out.ptr), out_stride)
        if not g:
            raise MchipError(last_error())
        return g

    @staticmethod
    def graph_launch(g):
        check(load().mchip_graph_launch(g))

    @staticmethod
    def graph_destroy(g):
        load().mchip_graph_destroy(g)

    def autotune_decode(self, d_blobs, offs, lens, qs, d_out, out_stride=CHUNK_BYTES):
        """Time the decode variants on this batch; keeps the fastest.  Returns the timings."""
        import json as _json
        offs = np.ascontiguousarray(offs, np.uint64)
        lens = np.ascontiguousarray(lens, np.uint64)
        qs = np.ascontiguousarray(qs, np.float32)
        rep = C.create_string_buffer(16384)
        check(load().mchip_autotune_decode(self.h, C.c_void_p(d_blobs.ptr), _ptr(offs), _ptr(lens), _ptr(qs),
                                           len(offs), C.c_void_p(d_out.ptr), out_stride, rep, 16384))
        return _json.loads(rep.value.decode())

    def autotune_encode(self, d_in, qs, taus, d_out, out_offs, out_caps, q_store=None, in_stride=CHUNK_BYTES):
        """Time the encode shapes on these chunks; keeps the fastest.  Returns the timings."""
        import json as _json
        qs = np.ascontiguousarray(qs, np.float32)
        taus = np.ascontiguousarray(taus, np.int32)
        offs = np.ascontiguousarray(out_offs, np.uint64)
        caps = np.ascontiguousarray(out_caps, np.uint64)
        qst = None if q_store is None else np.ascontiguousarray(q_store, np.float32)
        rep = C.create_string_buffer(8192)
        check(load().mchip_autotune_encode(self.h, C.c_void_p(d_in.ptr), in_stride, _ptr(qs), _ptr(taus),
                                           None if qst is None else _ptr(qst), len(qs), C.c_void_p(d_out.ptr),
                                           _ptr(offs), _ptr(caps), rep, 8192))
        return _json.loads(rep.value.decode())

    def decode_chunks_host(self, blobs):
        n = len(blobs)
        arrs = [np.frombuffer(b, np.uint8) if not isinstance(b, np.ndarray) else b for b in blobs]
        lens = np.array([a.size for a in arrs], np.uint64)
        out = np.empty((n, 256, 256, 256), np.uint8)
        check(load().mchip_decode_chunks_host(self.h, _ptr_array(arrs), _ptr(lens), n, _ptr(out), CHUNK_BYTES))
        return out

    def decode_blocks_host(self, payloads, qs):
        n = len(payloads)
        arrs = [None if p is None else (np.frombuffer(p, np.uint8) if not isinstance(p, np.ndarray) else p)
                for p in payloads]
        lens = np.array([0 if a is None else a.size for a in arrs], np.uint32)
        qs = np.ascontiguousarray(qs, np.float32)
        out = np.empty((n, 4096), np.uint8)
        check(load().mchip_decode_blocks_host(self.h, _ptr_array(arrs), _ptr(lens), _ptr(qs), n, _ptr(out)))
        return out


    def encode_graph(self, d_in, qs, taus, d_out, out_offs, out_caps, q_store=None, in_stride=CHUNK_BYTES):
        """HIP graph of one encode sequence (replay with graph_launch, then
        encode_graph_results)."""
        qs = np.ascontiguousarray(qs, np.float32)
        taus = np.ascontiguousarray(taus, np.int32)
        offs = np.ascontiguousarray(out_offs, np.uint64)
        caps = np.ascontiguousarray(out_caps, np.uint64)
        qst = None if q_store is None else np.ascontiguousarray(q_store, np.float32)
        g = load().mchip_encode_graph_create(self.h, C.c_void_p(d_in.ptr), in_stride, _ptr(qs), _ptr(taus),
                                             None if qst is None else _ptr(qst), len(qs), C.c_void_p(d_out.ptr),
                                             _ptr(offs), _ptr(caps))
        if not g:
            raise MchipError(last_error())
        return g, len(qs)

    @staticmethod
    def encode_graph_results(g, n):
        lens = np.zeros(n, np.uint64)
        st = np.zeros(n, np.int32)
        check(load().mchip_encode_graph_results(g, _ptr(lens), _ptr(st)))
        return lens, st

    # ---- encode ----------------------------------------------------------
    def encode_chunks(self, d_in, qs, taus, d_out, out_offs, out_caps, q_store=None, in_stride=CHUNK_BYTES):
        """Device-resident encode; returns (blob_len[n], status[n])."""
        n = len(qs)
        qs = np.ascontiguousarray(qs, np.float32)
        taus = np.ascontiguousarray(taus, np.int32)
        offs = np.ascontiguousarray(out_offs, np.uint64)
        caps = np.ascontiguousarray(out_caps, np.uint64)
        qst = None if q_store is None else np.ascontiguousarray(q_store, np.float32)
        lens = np.zeros(n, np.uint64)
        st = np.zeros(n, np.int32)
        check(load().mchip_encode_chunks(self.h, C.c_void_p(d_in.ptr), in_stride, _ptr(qs), _ptr(taus),
                                         None if qst is None else _ptr(qst), n, C.c_void_p(d_out.ptr),
                                         _ptr(offs), _ptr(caps), _ptr(lens), _ptr(st)))
        return lens, st

    def encode_chunks_host(self, chunks, qs, taus=None, q_store=None, cap=20 << 20):
        """Encode host chunks [n,256,256,256]; returns list of blob bytes."""
        chunks = np.ascontiguousarray(chunks, np.uint8)
        n = chunks.shape[0]
        qs = np.ascontiguousarray(qs, np.float32)
        taus = np.zeros(n, np.int32) if taus is None else np.ascontiguousarray(taus, np.int32)
        qst = None if q_store is None else np.ascontiguousarray(q_store, np.float32)
        bufs = [np.empty(cap, np.uint8) for _ in range(n)]
        caps = np.full(n, cap, np.uint64)
        lens = np.zeros(n, np.uint64)
        st = np.zeros(n, np.int32)
        check(load().mchip_encode_chunks_host(self.h, _ptr(chunks), _ptr(qs), _ptr(taus),
                                              None if qst is None else _ptr(qst), n, _ptr_array(bufs),
                                              _ptr(caps), _ptr(lens), _ptr(st)))
        out = []
        for k in range(n):
            if st[k] == 2:          # capacity too small: retry this chunk with the reported length
                out.append(self.encode_chunks_host(chunks[k:k + 1], qs[k:k + 1], taus[k:k + 1],
                                                   None if qst is None else qst[k:k + 1],
                                                   cap=int(lens[k]) + 4096)[0])
                continue
            if st[k] != 0:
                raise MchipError(f"encode of chunk {k} failed with status {st[k]}")
            out.append(bufs[k][: int(lens[k])].tobytes())
        return out


# ---- synthetic workload generator -------------------------------------------
def synth():
    global _synth
    if _synth is not None:
        return _synth
    p = os.path.join(LIB_DIR, "libmc_synth.so")
    if not os.path.exists(p):
        raise MchipError(f"{p} is not built: run python matter-compressor_amd/build.py")
    S = C.CDLL(p)
    S.mcs_chunk.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p]
    S.mcs_chunks.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_int]
    _synth = S
    return S


def synth_chunks(seeds, air, coords=None, nthreads=None):
    """Generate len(seeds) 256^3 chunks (air fraction per chunk)."""
    n = len(seeds)
    seeds = np.ascontiguousarray(seeds, np.uint64)
    air = np.ascontiguousarray(np.broadcast_to(np.asarray(air, np.float64), (n,)))
    out = np.empty((n, 256, 256, 256), np.uint8)
    if coords is None:
        cz = cy = cx = None
    else:
        cc = np.ascontiguousarray(coords, np.int32).reshape(n, 3)
        cz, cy, cx = [np.ascontiguousarray(cc[:, k]) for k in range(3)]
    nt = nthreads or min(16, os.cpu_count() or 4)
    synth().mcs_chunks(n, _ptr(seeds), None if cz is None else _ptr(cz), None if cy is None else _ptr(cy),
                       None if cx is None else _ptr(cx), _ptr(air), _ptr(out), nt)
    return out


# ---- drop-in matter-compressor API (include/matter_compressor_amd.h) ----------
class McBuf(C.Structure):
    _fields_ = [("p", C.c_void_p), ("len", C.c_size_t), ("cap", C.c_size_t)]


class McBox(C.Structure):
    _fields_ = [("z0", C.c_long), ("y0", C.c_long), ("x0", C.c_long)]


_api_declared = False


def api():
    """libmc_amd.so with the drop-in mc_* entry points declared."""
    global _api_declared
    L = load()
    if _api_declared:
        return L
    vp, i, f, u64, lng = C.c_void_p, C.c_int, C.c_float, C.c_uint64, C.c_long
    sig = {
        "mc_codec_init": (None, []),
        "mc_codec_ctx_new": (vp, []),
        "mc_codec_ctx_free": (None, [vp]),
        "mc_codec_ctx_set_quality": (None, [vp, f]),
        "mc_codec_ctx_get_quality": (f, [vp]),
        "mc_codec_ctx_set_max_error": (None, [vp, i]),
        "mc_codec_ctx_get_max_error": (i, [vp]),
        "mc_apply_preset": (f, [vp, i]),
        "mc_preset_quality": (f, [i]),
        "mc_preset_tau": (i, [i]),
        "mc_preset_name": (C.c_char_p, [i]),
        "mc_codec_set_priors": (None, [vp, vp]),
        "mc_buf_put": (None, [C.POINTER(McBuf), vp, C.c_size_t]),
        "mc_enc_block": (i, [vp, vp, C.POINTER(McBuf), C.POINTER(C.c_uint32)]),
        "mc_dec_block": (None, [vp, vp, C.c_uint32, vp]),
        "mc_enc_fracmap": (C.c_uint32, [vp, vp, C.c_size_t]),
        "mc_dec_fracmap": (None, [vp, C.c_uint32, vp]),
        "mc_archive_open": (vp, [C.c_char_p, i, f]),
        "mc_archive_open_dims": (vp, [C.c_char_p, i, i, i, f]),
        "mc_archive_close": (None, [vp]),
        "mc_archive_set_block_codec": (i, [vp, C.c_uint32]),
        "mc_archive_block_codec": (C.c_uint32, [vp]),
        "mc_archive_block_blob": (i, [vp, u64, i, i, i, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)]),
        "mc_archive_append_chunk_raw": (i, [vp, i, i, i, i, vp]),
        "mc_archive_append_chunk_raw_q": (i, [vp, i, i, i, i, vp, f]),
        "mc_archive_append_chunk_ctx": (i, [vp, vp, i, i, i, i, vp]),
        "mc_archive_append_chunk_compressed": (i, [vp, i, i, i, i, vp, C.c_size_t]),
        "mc_archive_append_chunk_par": (i, [vp, i, i, i, i, vp, f, i]),
        "mc_archive_reserve_index": (i, [vp, i, i, i, i]),
        "mc_archive_chunk_coverage": (i, [vp, i, i, i, i]),
        "mc_archive_data_len": (u64, [vp]),
        "mc_archive_set_metadata": (i, [vp, vp, C.c_size_t]),
        "mc_archive_metadata": (C.c_void_p, [vp, C.POINTER(C.c_size_t)]),
        "mc_archive_chunk_offset": (u64, [vp, i, i, i, i]),
        "mc_archive_decode_block": (None, [vp, u64, i, i, i, vp]),
        "mc_archive_decode_chunk": (None, [vp, u64, vp, i]),
        "mc_archive_read_region": (None, [vp, i, lng, lng, lng, lng, lng, lng, vp, C.c_size_t, C.c_size_t, i]),
        "mc_archive_read_regions": (None, [vp, i, vp, i, lng, lng, lng, vp, C.c_size_t, i]),
        "mc_serve_open": (vp, [vp, i, u64, i]),
        "mc_serve_open_levels": (vp, [vp, i, i, vp, u64, i]),
        "mc_serve_render_quad": (C.c_int64, [vp, vp, i, i, f, f, f, i, i, i, i, f, f, f, vp]),
        "mc_serve_submit": (C.c_int64, [vp, vp, i, lng, lng, lng, vp, C.c_size_t]),
        "mc_serve_poll": (i, [vp, C.c_int64]),
        "mc_serve_wait": (i, [vp, C.c_int64, C.POINTER(C.c_double)]),
        "mc_serve_close": (None, [vp]),
        "mc_serve_device": (vp, [vp]),
        "mc_serve_preload": (i, [vp, C.POINTER(C.c_uint64)]),
        "mc_serve_refresh": (i, [vp]),
        "mc_serve_invalidate_chunk": (i, [vp, i, i, i, i]),
        "mc_serve_resize": (i, [vp, u64]),
        "mc_archive_block_present": (i, [vp, i, i, i, i]),
        "mc_archive_block_fraction": (f, [vp, i, i, i, i]),
        "mc_archive_sample_boxes": (i, [vp, i, u64, i, lng, lng, lng, f, vp]),
        "mc_archive_set_priors": (i, [vp, vp, vp]),
        "mc_metadata": (C.c_void_p, [vp, C.POINTER(C.c_size_t)]),
        "mc_chunk_compute_hash": (u64, [vp, u64]),
        "mc_verify_archive": (lng, [vp, C.c_size_t, i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _api_declared = True
    return L


SERVER_STATS = ["requests", "completed", "boxes", "block_refs", "hits", "misses", "evictions", "deferred",
                "arena_slots", "renders", "arena_grows", "arena_mapped_bytes", "arena_budget_bytes",
                "updated_blocks"]


def server_stats(server_handle):
    """mchip_server_stats of a device serving loop as a dict."""
    buf = (C.c_uint64 * len(SERVER_STATS))()
    check(load().mchip_server_get_stats(server_handle, buf))
    return dict(zip(SERVER_STATS, [int(v) for v in buf]))


class Serve:
    """mc_serve_*: HBM-resident serving of an archive LOD (box batches), or of
    consecutive LODs (boxes of the first, quad renders over all)."""

    def __init__(self, archive, lod=0, arena_bytes=0, lanes=8, nlods=None, dims=None):
        L = api()
        if nlods is None:
            self.h = L.mc_serve_open(archive, lod, arena_bytes, lanes)
        else:
            d = None if dims is None else np.ascontiguousarray(dims, np.int32).reshape(-1)
            self.h = L.mc_serve_open_levels(archive, lod, nlods, None if d is None else _ptr(d), arena_bytes, lanes)
        if not self.h:
            raise MchipError(f"mc_serve_open failed: {last_error()}")

    def render(self, grid, x0, y0, step, w, h, d_out, filt=1, comp=2, t0=-4.0, t1=4.0, dt=1.0, offset=0):
        g = np.ascontiguousarray(grid, np.float32)
        t = api().mc_serve_render_quad(self.h, _ptr(g), g.shape[1], g.shape[0], x0, y0, step, w, h, filt, comp,
                                       t0, t1, dt, C.c_void_p(d_out.ptr + offset))
        if t < 0:
            raise MchipError(f"mc_serve_render_quad failed: {last_error()}")
        return t

    def render_level(self, ticket):
        return load().mchip_server_render_level(api().mc_serve_device(self.h), ticket)

    def submit(self, origins, dz, dy, dx, d_out, batch_stride=None, offset=0):
        o = np.ascontiguousarray(origins, np.int64).reshape(-1, 3)
        n = o.shape[0]
        stride = dz * dy * dx if batch_stride is None else batch_stride
        boxes = (McBox * n)(*[McBox(int(a), int(b), int(c)) for a, b, c in o])
        t = api().mc_serve_submit(self.h, boxes, n, dz, dy, dx, C.c_void_p(d_out.ptr + offset), stride)
        if t < 0:
            raise MchipError(f"mc_serve_submit failed: {last_error()}")
        return t

    def poll(self, ticket):
        return api().mc_serve_poll(self.h, ticket)

    def wait(self, ticket):
        lat = C.c_double(0.0)
        if api().mc_serve_wait(self.h, ticket, C.byref(lat)) != 0:
            raise MchipError(f"serving request failed: {last_error()}")
        return lat.value

    def preload(self):
        n = C.c_uint64(0)
        if api().mc_serve_preload(self.h, C.byref(n)) != 0:
            raise MchipError(f"mc_serve_preload failed: {last_error()}")
        return n.value

    def stats(self):
        return server_stats(api().mc_serve_device(self.h))

    def refresh(self):
        n = api().mc_serve_refresh(self.h)
        if n < 0:
            raise MchipError(f"mc_serve_refresh failed: {last_error()}")
        return n

    def invalidate_chunk(self, lod, cz, cy, cx):
        return api().mc_serve_invalidate_chunk(self.h, lod, cz, cy, cx)

    def resize(self, arena_bytes):
        if api().mc_serve_resize(self.h, arena_bytes) != 0:
            raise MchipError(f"mc_serve_resize failed: {last_error()}")

    def close(self):
        if self.h:
            api().mc_serve_close(self.h)
            self.h = None
