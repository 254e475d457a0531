"""Python mirror of the reference's ctypes wrapper (src/interfaces/python/bwc.py)
over the MI355X libbwc.so.

Same function names and argument meaning as the reference wrapper
(init_stream, alloc_coder, alloc_decoder, free_codec, set_codeblocks,
set_decomp, set_qm, set_tiles, set_precincts, create_compression, compress,
create_decompression, decompress), with argument types matching the C API
(include/bwc.h). The library must have been built (python -m
bigwhoop_amd.build); loading fails loudly otherwise — there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
from enum import IntEnum
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libbwc.so"


class DWT(IntEnum):
    CDF_9_7 = 0
    LeGall_5_3 = 1
    Haar = 2


class PROG(IntEnum):
    LRCP = 0
    RLCP = 1


class TILE(IntEnum):
    SIZEOF = 0
    NUMBEROF = 1


class PREC(IntEnum):
    single = 4
    double = 8


class MODE(IntEnum):
    comp = 0
    decomp = 1


_lib = None


def lib():
    """Load libbwc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m bigwhoop_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        vp, u8, u16, u32, u64 = C.c_void_p, C.c_uint8, C.c_uint16, C.c_uint32, C.c_uint64
        sig = {
            "bwc_init_stream": (vp, [vp, vp, C.c_int]),
            "bwc_set_com": (C.c_ubyte, [vp, C.c_char_p, u16]),
            "bwc_set_aux": (C.c_ubyte, [vp, C.c_char_p, u32]),
            "bwc_alloc_coder": (vp, [u64, u64, u64, u64, u8, C.c_int]),
            "bwc_alloc_decoder": (vp, []),
            "bwc_free_codec": (None, [vp]),
            "bwc_set_error_resilience": (None, [vp]),
            "bwc_set_decomp": (None, [vp, u8, u8, u8, u8]),
            "bwc_set_precincts": (None, [vp, u8, u8, u8, u8]),
            "bwc_set_codeblocks": (None, [vp, u8, u8, u8, u8]),
            "bwc_set_qm": (None, [vp, u8]),
            "bwc_set_tiles": (None, [vp, u64, u64, u64, u64, C.c_int]),
            "set_kernels": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
            "set_progression": (None, [vp, C.c_int]),
            "set_quant_style": (None, [vp, C.c_int]),
            "set_quant_step_size": (None, [vp, C.c_double]),
            "bwc_create_compression": (C.c_ubyte, [vp, vp, C.c_char_p]),
            "bwc_compress": (C.c_size_t, [vp, vp]),
            "bwc_create_decompression": (C.c_ubyte, [vp, vp, u8]),
            "bwc_decompress": (C.c_ubyte, [vp, vp]),
            "bwc_open_header": (vp, [vp]),
            "bwc_close_header": (None, [vp]),
            "bwc_amd_version": (C.c_int, []),
            "bwc_amd_device_count": (C.c_int, []),
            "bwc_amd_host_alloc": (vp, [C.c_size_t]),
            "bwc_amd_host_free": (None, [vp]),
            "bwc_amd_last_error": (C.c_char_p, []),
            "bwc_amd_reset_gain_lut": (None, []),
            "bwc_amd_compress_batch": (C.c_int, [C.POINTER(vp), C.POINTER(vp), C.c_int, C.POINTER(C.c_size_t),
                                                 C.c_int]),
            "bwc_amd_set_tuning": (C.c_int, [C.c_char_p, C.c_int]),
            "bwc_amd_get_tuning": (C.c_int, [C.c_char_p, C.POINTER(C.c_int)]),
            "bwc_amd_tuning_source": (C.c_char_p, []),
            "bwc_amd_ncblk": (u64, [vp, u64, u8]),
            "bwc_amd_param_dims": (C.c_int, [vp, u64, C.POINTER(u64)]),
            "bwc_amd_stage_dwt": (C.c_int, [vp, u64, u8, vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "bwc_amd_stage_dwt_noalloc": (C.c_int, [vp, u64, u8, vp, vp, vp, vp]),
            "bwc_amd_stage_idwt": (C.c_int, [vp, u64, u8, C.POINTER(C.c_double)]),
            "bwc_amd_stage_t1": (C.c_int, [vp, u64, u8, C.POINTER(C.c_double), C.POINTER(u8), C.POINTER(u8),
                                           C.POINTER(u8), C.POINTER(u32), C.POINTER(u16), C.POINTER(u8), u64,
                                           C.POINTER(u64), C.POINTER(u16)]),
            "bwc_amd_last_timings": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
            "bwc_amd_cblk_info": (C.c_int, [vp, u64, u8, C.POINTER(u32), C.POINTER(C.c_double), C.POINTER(u32)]),
            "bwc_amd_begin_records": (None, [vp]),
            "bwc_amd_set_records": (C.c_int, [vp, u64, u8, C.POINTER(u8), C.POINTER(u8), C.POINTER(u8),
                                              C.POINTER(u32), C.POINTER(u16), C.POINTER(u8), C.POINTER(u64),
                                              C.POINTER(C.c_double)]),
            "bwc_amd_encode_records": (C.c_size_t, [vp, vp]),
            "bwc_amd_set_pass_skip": (C.c_int, [vp, C.c_int, C.c_int]),
            "bwc_amd_set_skip_model": (C.c_int, [vp, C.POINTER(C.c_float), C.c_size_t]),
            "bwc_amd_skip_stats": (C.c_int, [vp, C.POINTER(u64)]),
            "bwc_amd_cblk_cuts": (C.c_int, [vp, u64, u8, C.POINTER(C.c_int16)]),
            "bwc_amd_rc_trace": (C.c_int, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(C.c_int64), C.c_int]),
            "bwc_amd_graph_counts": (C.c_int, [vp, C.POINTER(u64)]),
            "bwc_amd_batch_queued": (u64, [vp]),
            "bwc_amd_stuffed_bytes": (C.c_int, [C.POINTER(u32), C.POINTER(u64), C.POINTER(u64), C.c_int,
                                                C.POINTER(u64)]),
            "bwc_amd_collect_features": (C.c_int, [vp, C.c_int]),
            "bwc_amd_cblk_features": (C.c_int, [vp, u64, u8, C.POINTER(C.c_float)]),
            "bwc_amd_cblk_decisions": (C.c_int, [vp, u64, u8, C.POINTER(u32)]),
            "bwc_amd_skipnet_train": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), u64,
                                                C.POINTER(C.c_float), C.c_int, C.c_float, u64,
                                                C.POINTER(C.c_float)]),
            "bwc_amd_skipnet_grad": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), u64,
                                               C.POINTER(C.c_float), C.POINTER(C.c_float),
                                               C.POINTER(C.c_float)]),
            "bwc_amd_skipnet_predict": (C.c_int, [C.POINTER(C.c_float), u64, C.POINTER(C.c_float),
                                                  C.POINTER(C.c_float)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    e = lib().bwc_amd_last_error()
    return e.decode() if e else ""


# ---- reference wrapper names -------------------------------------------------
def init_stream(inpbuf, outbuf, mode):
    return lib().bwc_init_stream(inpbuf, outbuf, int(mode))


def alloc_coder(nX, nY, nZ, nTS, nPar, prec):
    return lib().bwc_alloc_coder(nX, nY, nZ, nTS, nPar, int(prec))


def alloc_decoder():
    return lib().bwc_alloc_decoder()


def free_codec(codec):
    lib().bwc_free_codec(codec)


def set_codeblocks(codec, cbX, cbY, cbZ, cbTS):
    lib().bwc_set_codeblocks(codec, cbX, cbY, cbZ, cbTS)


def set_decomp(codec, decompX, decompY, decompZ, decompTS):
    lib().bwc_set_decomp(codec, decompX, decompY, decompZ, decompTS)


def set_qm(codec, Qm):
    lib().bwc_set_qm(codec, Qm)


def set_tiles(codec, tilesX, tilesY, tilesZ, tilesTS, instr):
    lib().bwc_set_tiles(codec, tilesX, tilesY, tilesZ, tilesTS, int(instr))


def set_precincts(codec, pX, pY, pZ, pTS):
    lib().bwc_set_precincts(codec, pX, pY, pZ, pTS)


def set_error_resilience(codec):
    lib().bwc_set_error_resilience(codec)


def create_compression(codec, stream, rate_control):
    buf = C.create_string_buffer(rate_control.encode() if isinstance(rate_control, str) else rate_control)
    return lib().bwc_create_compression(codec, stream, buf)


def compress(codec, stream):
    return lib().bwc_compress(codec, stream)


def create_decompression(codec, stream, layer):
    return lib().bwc_create_decompression(codec, stream, layer)


def decompress(codec, stream):
    return lib().bwc_decompress(codec, stream)


# ---- array-level helpers -----------------------------------------------------
def _shape4(shape):
    s = [int(x) for x in shape] + [1] * (4 - len(shape))
    return s[:4]


TUNING_KNOBS = ("dwt_unroll", "dwt_seg_log2", "dwt_seg_min", "mq_lanes_full", "mq_lanes_limited")


def set_tuning(name: str, value: int):
    """Set one kernel launch-shape knob (include/bwc_amd.h); no result changes."""
    if lib().bwc_amd_set_tuning(name.encode(), int(value)):
        raise ValueError(f"tuning knob {name}={value} rejected: " + last_error())


def get_tuning(name: str = None):
    """One knob's value, or all of them as a dict."""
    if name is None:
        return {k: get_tuning(k) for k in TUNING_KNOBS}
    v = C.c_int()
    if lib().bwc_amd_get_tuning(name.encode(), C.byref(v)):
        raise ValueError(f"unknown tuning knob {name}")
    return v.value


def compress_batch(codecs, ptrs, workers: int = 0):
    """Compress each Codec from its host or device address, concurrently
    (bwc_amd_compress_batch); returns the codestream sizes (each codec's
    .out holds its codestream until its next call)."""
    n = len(codecs)
    if len(ptrs) != n:
        raise ValueError("one input address per codec")
    for c, p in zip(codecs, ptrs):
        _Stream.from_address(c.stream).inp = p
    hs = (C.c_void_p * n)(*[c.h for c in codecs])
    ss = (C.c_void_p * n)(*[c.stream for c in codecs])
    sizes = (C.c_size_t * n)()
    if lib().bwc_amd_compress_batch(hs, ss, n, sizes, int(workers)):
        raise RuntimeError("bwc_amd_compress_batch failed: " + last_error())
    return [int(x) for x in sizes]


def tuning_source() -> str:
    """The tuning file the library read ("" when none)."""
    return (lib().bwc_amd_tuning_source() or b"").decode()


class Codec:
    """One configured compression job: owns the bwc_codec and its stream.

    ``shape`` is (nX, nY, nZ, nTS) with x fastest; settings follow the C
    setters and are applied in the order kernels (set_kernels), progression
    (set_progression), quantisation style (set_quant_style), decomp,
    precincts, codeblocks, qm, tiles, error resilience, quantisation step
    size (set_quant_step_size), then the stream's AUX and COM segments.
    """

    def __init__(self, shape, npar=1, prec=8, rate="2", decomp=None, precincts=None, cb=None, qm=None,
                 tiles=None, erres=False, kernels=None, prog=None, qstyle=None, qstep=None, com=None, aux=None,
                 fresh_gain_table=True, pinned=False):
        """``pinned``: the codestream buffer is page-locked host memory
        (bwc_amd_host_alloc), so packet bodies reach it by DMA."""
        L = lib()
        if fresh_gain_table:
            L.bwc_amd_reset_gain_lut()
        n = _shape4(shape)
        self.shape = n
        self.npar = npar
        self.prec = prec
        self.nbytes = int(np.prod(n)) * npar * prec
        self.h = L.bwc_alloc_coder(n[0], n[1], n[2], n[3], npar, prec)
        if not self.h:
            raise RuntimeError("bwc_alloc_coder failed")
        if kernels is not None:
            L.set_kernels(self.h, *[int(x) for x in kernels])
        if prog is not None:
            L.set_progression(self.h, int(prog))
        if qstyle is not None:
            L.set_quant_style(self.h, int(qstyle))
        if decomp is not None:
            L.bwc_set_decomp(self.h, *[int(x) for x in decomp])
        if precincts is not None:
            L.bwc_set_precincts(self.h, *[int(x) for x in precincts])
        if cb is not None:
            L.bwc_set_codeblocks(self.h, *[int(x) for x in cb])
        if qm is not None:
            L.bwc_set_qm(self.h, int(qm))
        if tiles is not None:
            L.bwc_set_tiles(self.h, *[int(x) for x in tiles], 0)
        if erres:
            L.bwc_set_error_resilience(self.h)
        if qstep is not None:
            L.set_quant_step_size(self.h, float(qstep))
        nout = self.nbytes * 2 + (1 << 20)
        self._pinned = None
        if pinned:
            self._pinned = L.bwc_amd_host_alloc(nout)
            if not self._pinned:
                raise RuntimeError("bwc_amd_host_alloc failed: " + last_error())
            self.out = np.ctypeslib.as_array((C.c_uint8 * nout).from_address(self._pinned))
        else:
            self.out = np.zeros(nout, np.uint8)
        self.stream = L.bwc_init_stream(None, self.out.ctypes.data_as(C.c_void_p), 0)
        for text, setter in ((aux, L.bwc_set_aux), (com, L.bwc_set_com)):
            if text is not None:
                b = text.encode() if isinstance(text, str) else bytes(text)
                if setter(self.stream, b, len(b)):
                    raise RuntimeError("bwc_set_aux / bwc_set_com failed")
        self._rate = C.create_string_buffer(rate.encode())
        if L.bwc_create_compression(self.h, self.stream, self._rate):
            raise RuntimeError("bwc_create_compression failed: " + last_error())

    def close(self):
        if getattr(self, "h", None):
            lib().bwc_free_codec(self.h)
            self.h = None
        if getattr(self, "stream", None):
            # the library never frees AUX / COM spans (the caller owns them)
            st = _Stream.from_address(self.stream)
            for span in (st.aux, st.com):
                if span:
                    C.CDLL(None).free(C.c_void_p(C.c_void_p.from_address(span + 24).value))
                    C.CDLL(None).free(C.c_void_p(span))
            C.CDLL(None).free(C.c_void_p(self.stream))
            self.stream = None
        if getattr(self, "_pinned", None):
            self.out = None
            lib().bwc_amd_host_free(self._pinned)
            self._pinned = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def compress_into(self, ptr: int) -> int:
        """Compress from a host or HIP device address (e.g. tensor.data_ptr())
        into the codec's output buffer; returns the codestream size
        (self.out[:size] holds it until the next call)."""
        st = _Stream.from_address(self.stream)
        st.inp = ptr
        n = lib().bwc_compress(self.h, self.stream)
        if n == 0:
            raise RuntimeError("bwc_compress failed: " + last_error())
        return int(n)

    def compress_ptr(self, ptr: int) -> bytes:
        """Compress from a host or HIP device address; returns the codestream."""
        n = self.compress_into(ptr)
        return self.out[:n].tobytes()

    def compress(self, data: np.ndarray) -> bytes:
        data = np.ascontiguousarray(data)
        if data.nbytes != self.nbytes:
            raise ValueError("input size does not match the codec")
        self._keep = data
        return self.compress_ptr(data.ctypes.data)

    def timings(self):
        ms = (C.c_double * 10)()
        la = (C.c_int * 4)()
        lib().bwc_amd_last_timings(self.h, ms, la)
        return list(ms), list(la)

    # ---- pass skipping ----
    def set_pass_skip(self, mode: int, param: int = 0):
        if lib().bwc_amd_set_pass_skip(self.h, mode, param):
            raise RuntimeError("bwc_amd_set_pass_skip failed: " + last_error())

    def set_skip_model(self, weights):
        w = np.ascontiguousarray(weights, dtype=np.float32)
        if lib().bwc_amd_set_skip_model(self.h, _p(w, C.c_float), w.size):
            raise RuntimeError("bwc_amd_set_skip_model failed: " + last_error())

    def skip_stats(self):
        s = (C.c_uint64 * 8)()
        lib().bwc_amd_skip_stats(self.h, s)
        keys = ["codeblocks", "passes_coded", "passes_complete", "extra_rounds", "extended", "fallbacks"]
        return dict(zip(keys, list(s)[:6]))

    def cuts(self, tile=0, param=0):
        out = np.zeros(self.ncblk(tile, param), np.int16)
        if lib().bwc_amd_cblk_cuts(self.h, tile, param, _p(out, C.c_int16)):
            raise RuntimeError("bwc_amd_cblk_cuts failed")
        return out

    def batch_queued(self):
        """Tier-1 calls of this codec coded through a batch's global codeblock queue."""
        return int(lib().bwc_amd_batch_queued(self.h))

    def graph_counts(self):
        """(graphs recorded, graph launches) of this codec's pipeline graphs."""
        out = (C.c_uint64 * 2)()
        if lib().bwc_amd_graph_counts(self.h, out):
            return (0, 0)
        return int(out[0]), int(out[1])

    def rc_trace(self, cap=4096):
        """(layer, threshold, estimate) of every probe the device PCRD
        bisection of the last compression's last tile made."""
        lay = np.zeros(cap, np.uint32)
        slo = np.zeros(cap, np.uint32)
        est = np.zeros(cap, np.int64)
        n = lib().bwc_amd_rc_trace(self.h, _p(lay, C.c_uint32), _p(slo, C.c_uint32), _p(est, C.c_int64), cap)
        if n < 0:
            raise RuntimeError("bwc_amd_rc_trace: no compression on this codec")
        n = min(n, cap)
        return list(zip(lay[:n].tolist(), slo[:n].tolist(), est[:n].tolist()))

    def collect_features(self, on: bool = True):
        lib().bwc_amd_collect_features(self.h, 1 if on else 0)

    def features(self, tile=0, param=0):
        out = np.zeros((self.ncblk(tile, param), 16), np.float32)
        if lib().bwc_amd_cblk_features(self.h, tile, param, _p(out, C.c_float)):
            raise RuntimeError("bwc_amd_cblk_features failed: " + last_error())
        return out

    # ---- single stages (bwc_amd.h) ----
    def decisions(self, tile=0, param=0):
        """MQ decisions per codeblock of the last compress (collect_features on)."""
        out = np.zeros(self.ncblk(tile, param), np.uint32)
        if lib().bwc_amd_cblk_decisions(self.h, tile, param, _p(out, C.c_uint32)):
            raise RuntimeError("bwc_amd_cblk_decisions failed: " + last_error())
        return out

    def ncblk(self, tile=0, param=0):
        return lib().bwc_amd_ncblk(self.h, tile, param)

    def param_dims(self, tile=0):
        d = (C.c_uint64 * 4)()
        if lib().bwc_amd_param_dims(self.h, tile, d):
            raise RuntimeError(last_error())
        return tuple(int(x) for x in d)

    def geometry(self, tile=0, param=0):
        n = self.ncblk(tile, param)
        box = np.zeros((n, 8), np.uint32)
        sb = np.zeros((n, 3), np.float64)
        meta = np.zeros((n, 4), np.uint32)
        lib().bwc_amd_cblk_info(self.h, tile, param, _p(box, C.c_uint32), _p(sb, C.c_double), _p(meta, C.c_uint32))
        return box, sb, meta

    def stage_dwt(self, data, tile=0, param=0):
        data = np.ascontiguousarray(data)
        out = np.zeros(int(np.prod(self.param_dims(tile))), np.float64)
        mm = np.zeros(4, np.float64)
        if lib().bwc_amd_stage_dwt(self.h, tile, param, data.ctypes.data_as(C.c_void_p), _p(out, C.c_double),
                                   _p(mm, C.c_double)):
            raise RuntimeError("stage_dwt failed: " + last_error())
        return out, mm

    def stage_idwt(self, coeffs, tile=0, param=0):
        c = np.ascontiguousarray(coeffs, dtype=np.float64).copy()
        if lib().bwc_amd_stage_idwt(self.h, tile, param, _p(c, C.c_double)):
            raise RuntimeError("stage_idwt failed: " + last_error())
        return c

    def stage_t1(self, coeffs, tile=0, param=0):
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
        n = self.ncblk(tile, param)
        K = np.zeros(n, np.uint8)
        Km = np.zeros(n, np.uint8)
        Z = np.zeros(n, np.uint8)
        Lr = np.zeros((n, 190), np.uint32)
        S = np.zeros((n, 191), np.uint16)
        cap = coeffs.size * 10 + 4096
        data = np.zeros(cap, np.uint8)
        off = np.zeros(n + 1, np.uint64)
        sl = np.zeros(2, np.uint16)
        rc = lib().bwc_amd_stage_t1(self.h, tile, param, _p(coeffs, C.c_double), _p(K, C.c_uint8),
                                    _p(Km, C.c_uint8), _p(Z, C.c_uint8), _p(Lr, C.c_uint32), _p(S, C.c_uint16),
                                    _p(data, C.c_uint8), cap, _p(off, C.c_uint64), _p(sl, C.c_uint16))
        if rc:
            raise RuntimeError("stage_t1 failed: " + last_error())
        return dict(K=K, Kmsbs=Km, Z=Z, L=Lr, S=S, data=data[: int(off[-1])], off=off, slopes=tuple(int(x) for x in sl))

    def begin_records(self):
        lib().bwc_amd_begin_records(self.h)

    def set_records(self, tile, param, K, Kmsbs, Z, L, S, data, off, mm):
        K, Kmsbs, Z = [np.ascontiguousarray(a, np.uint8) for a in (K, Kmsbs, Z)]
        L = np.ascontiguousarray(L, np.uint32)
        S = np.ascontiguousarray(S, np.uint16)
        data = np.ascontiguousarray(np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray)
                                    else data, np.uint8)
        if data.size == 0:
            data = np.zeros(1, np.uint8)
        off = np.ascontiguousarray(off, np.uint64)
        mm = np.ascontiguousarray(mm, np.float64)
        self._rec_keep = (K, Kmsbs, Z, L, S, data, off, mm)
        if lib().bwc_amd_set_records(self.h, tile, param, _p(K, C.c_uint8), _p(Kmsbs, C.c_uint8), _p(Z, C.c_uint8),
                                     _p(L, C.c_uint32), _p(S, C.c_uint16), _p(data, C.c_uint8), _p(off, C.c_uint64),
                                     _p(mm, C.c_double)):
            raise RuntimeError("set_records failed")

    def encode_records(self) -> bytes:
        n = lib().bwc_amd_encode_records(self.h, self.stream)
        if n == 0:
            raise RuntimeError("encode_records failed")
        return self.out[:n].tobytes()


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class _Stream(C.Structure):
    _fields_ = [("aux", C.c_void_p), ("com", C.c_void_p), ("inp", C.c_void_p), ("out", C.c_void_p),
                ("mode", C.c_int)]


def compress_array(data, shape, npar=1, rate="2", **kw) -> bytes:
    """One-shot compression of a numpy array (float32 or float64)."""
    prec = 4 if np.asarray(data).dtype == np.float32 else 8
    c = Codec(shape, npar=npar, prec=prec, rate=rate, **kw)
    try:
        return c.compress(data)
    finally:
        c.close()


class _GlInf(C.Structure):
    _fields_ = [("nX", C.c_uint64), ("nY", C.c_uint64), ("nZ", C.c_uint64), ("nTS", C.c_uint64),
                ("nPar", C.c_uint8), ("data_prec", C.c_int), ("codec_prec", C.c_int)]


def header_info(cs: bytes) -> dict:
    """bwc_open_header on a codestream: dimensions, parameters, precisions,
    layers and their rates."""
    L = lib()
    # bwc_open_header only reads the buffer, so the codestream is passed in
    # place (a copy of a large codestream would cost more than the parse)
    buf = np.frombuffer(cs, np.uint8)
    h = L.bwc_open_header(buf.ctypes.data_as(C.c_void_p))
    if not h:
        raise RuntimeError("bwc_open_header failed: " + last_error())
    try:
        inf = _GlInf.from_address(h)
        # bwc_header: info, then bwc_gl_ctrl; nLayers and bitrate are read
        # through the header layout mirrored in include/bwc.h
        ctrl = h + C.sizeof(_GlInf)
        out = dict(shape=(inf.nX, inf.nY, inf.nZ, inf.nTS), npar=inf.nPar, data_prec=inf.data_prec,
                   codec_prec=inf.codec_prec)
        off = _ctrl_offsets()
        nl = C.c_uint8.from_address(ctrl + off["nLayers"]).value
        rp = C.c_void_p.from_address(ctrl + off["bitrate"]).value
        out["rates"] = [C.c_float.from_address(rp + 4 * i).value for i in range(nl)] if rp else []
        return out
    finally:
        L.bwc_close_header(C.c_void_p(h))


def _ctrl_offsets():
    class Ctrl(C.Structure):
        _fields_ = [("CSsgc", C.c_uint16), ("error_resilience", C.c_ubyte), ("tileSizeX", C.c_uint64),
                    ("tileSizeY", C.c_uint64), ("tileSizeZ", C.c_uint64), ("tileSizeTS", C.c_uint64),
                    ("nTiles", C.c_uint64), ("precSizeX", C.c_uint8), ("precSizeY", C.c_uint8),
                    ("precSizeZ", C.c_uint8), ("precSizeTS", C.c_uint8), ("cbX", C.c_uint8), ("cbY", C.c_uint8),
                    ("cbZ", C.c_uint8), ("cbTS", C.c_uint8), ("KernelX", C.c_int), ("KernelY", C.c_int),
                    ("KernelZ", C.c_int), ("KernelTS", C.c_int), ("decompX", C.c_uint8), ("decompY", C.c_uint8),
                    ("decompZ", C.c_uint8), ("decompTS", C.c_uint8), ("nDecomp", C.c_uint8),
                    ("qt_exponent", C.c_uint32), ("qt_mantissa", C.c_uint32), ("Qm", C.c_uint8),
                    ("bitrate", C.c_void_p), ("nLayers", C.c_uint8)]
    return {n: getattr(Ctrl, n).offset for n, _ in Ctrl._fields_}


def decompress_bytes(cs: bytes, layer: int = 0, out=None):
    """bwc_create_decompression + bwc_decompress: the decoded field as a flat
    array (float64 or float32, [param][t][z][y][x]), and the header info.
    `out` may be a device tensor (e.g. torch) of the right size to decode
    into instead (then it is returned)."""
    L = lib()
    info = header_info(cs)
    n = int(np.prod(info["shape"])) * info["npar"]
    dt = np.float64 if info["data_prec"] == 8 else np.float32
    buf = np.frombuffer(cs, np.uint8).copy()
    codec = L.bwc_alloc_decoder()
    host = None
    if out is None:
        host = np.zeros(n, dt)
        optr = host.ctypes.data
    else:
        optr = out.data_ptr()
    stream = L.bwc_init_stream(buf.ctypes.data_as(C.c_void_p), C.c_void_p(optr), 1)
    try:
        if L.bwc_create_decompression(codec, stream, layer):
            raise RuntimeError("bwc_create_decompression failed: " + last_error())
        if L.bwc_decompress(codec, stream):
            raise RuntimeError("bwc_decompress failed: " + last_error())
    finally:
        L.bwc_free_codec(codec)
        st = _Stream.from_address(stream)
        for span in (st.aux, st.com):
            if span:
                C.CDLL(None).free(C.c_void_p(C.c_void_p.from_address(span + 24).value))
                C.CDLL(None).free(C.c_void_p(span))
        C.CDLL(None).free(C.c_void_p(stream))
    return (host if out is None else out), info
