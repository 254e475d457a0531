"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it has been
built, to the reference driver (oracle/_ref/ref_dump).

Test infrastructure only: the oracle is the checker the HIP path is compared
against; nothing in the product imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_DUMP = ROOT / "oracle" / "_ref" / "ref_dump"

u8p = C.POINTER(C.c_uint8)


class OcParams(C.Structure):
    _fields_ = [
        ("n", C.c_uint64 * 4),
        ("nPar", C.c_uint8),
        ("data_prec", C.c_uint8),
        ("have_decomp", C.c_uint8),
        ("decomp", C.c_uint8 * 4),
        ("have_prec", C.c_uint8),
        ("prec", C.c_uint8 * 4),
        ("have_cb", C.c_uint8),
        ("cb", C.c_uint8 * 4),
        ("have_qm", C.c_uint8),
        ("qm", C.c_uint8),
        ("have_tiles", C.c_uint8),
        ("tiles", C.c_uint64 * 4),
        ("erres", C.c_uint8),
        ("rate", C.c_char_p),
        ("have_qstyle", C.c_uint8),
        ("qstyle", C.c_uint8),
        ("have_qt", C.c_uint8),
        ("qt_exp", C.c_uint8),
        ("qt_mant", C.c_uint16),
        ("have_prog", C.c_uint8),
        ("prog", C.c_uint8),
        ("have_kern", C.c_uint8),
        ("kern", C.c_uint8 * 4),
        ("nrates", C.c_uint8),
        ("rates", C.POINTER(C.c_float)),
        ("have_qstep", C.c_uint8),
        ("qstep", C.c_double),
        ("aux", C.c_char_p),
        ("naux", C.c_uint32),
        ("com", C.c_char_p),
        ("ncom", C.c_uint16),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            subprocess.run(["make", "-C", str(ROOT / "oracle"), "liboracle.so"], check=True,
                           stdout=subprocess.DEVNULL)
        L = C.CDLL(str(ORACLE_SO))
        L.oc_create.restype = C.c_void_p
        L.oc_create.argtypes = [C.POINTER(OcParams)]
        L.oc_destroy.argtypes = [C.c_void_p]
        L.oc_compress.restype = C.c_size_t
        L.oc_compress.argtypes = [C.c_void_p, C.c_void_p, u8p, C.c_size_t]
        L.oc_ntiles.restype = C.c_uint64
        L.oc_ntiles.argtypes = [C.c_void_p]
        L.oc_nlayers.restype = C.c_uint8
        L.oc_nlayers.argtypes = [C.c_void_p]
        L.oc_param_size.restype = C.c_uint64
        L.oc_param_size.argtypes = [C.c_void_p, C.c_uint64]
        L.oc_param_dims.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]
        L.oc_ncblk.restype = C.c_uint64
        L.oc_ncblk.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32]
        L.oc_cblk_geometry.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_double), C.POINTER(C.c_uint32)]
        L.oc_fill_normalize.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
        L.oc_dwt_forward.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]
        L.oc_dwt_inverse.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]
        L.oc_t1_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]
        L.oc_cblk_record.restype = C.c_uint64
        L.oc_cblk_record.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64, u8p, C.POINTER(C.c_uint16),
                                     C.POINTER(C.c_uint64), C.POINTER(u8p)]
        L.oc_tile_slopes.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_uint16)]
        L.oc_t2_encode.restype = C.c_int
        L.oc_t2_encode.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_int16)]
        L.oc_rc_trace.restype = C.c_int
        L.oc_rc_trace.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_int64), C.c_int]
        L.oc_assemble.restype = C.c_size_t
        L.oc_assemble.argtypes = [C.c_void_p, u8p, C.c_size_t]
        L.oc_t1_probe_block.restype = C.c_int
        L.oc_t1_probe_block.argtypes = [C.POINTER(C.c_double), C.c_uint64, C.c_uint64, C.c_uint64,
                                        C.POINTER(C.c_uint64), C.c_double, C.c_double, C.c_double, C.c_uint8,
                                        C.c_uint16, C.c_uint8, C.c_int, C.POINTER(C.c_uint8),
                                        C.POINTER(C.c_uint16), C.POINTER(C.c_uint64), C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_uint32)]
        L.oc_t1_encode_block.restype = C.c_int
        L.oc_t1_encode_block.argtypes = [C.POINTER(C.c_double), C.c_uint64, C.c_uint64, C.c_uint64,
                                         C.POINTER(C.c_uint64), C.c_double, C.c_double, C.c_double, C.c_uint8,
                                         C.c_uint16, C.c_uint8, u8p, C.POINTER(C.c_uint16),
                                         C.POINTER(C.c_uint64), u8p, C.c_uint64]
        L.oc_mq_encode_symbols.restype = C.c_uint64
        L.oc_mq_encode_symbols.argtypes = [u8p, u8p, C.c_uint64, u8p, C.c_uint64]
        L.oc_mq_decode_symbols.argtypes = [u8p, C.c_uint64, u8p, C.c_uint64, u8p]
        L.oc_gain_lut.argtypes = [C.c_int, C.POINTER(C.c_double)]
        L.oc_slope2log.restype = C.c_uint16
        L.oc_slope2log.argtypes = [C.c_double]
        L.oc_decompress.restype = C.c_int
        L.oc_decompress.argtypes = [u8p, C.c_size_t, C.c_uint8, C.c_void_p, C.POINTER(C.c_uint64)]
        _lib = L
    return _lib


def ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def make_params(shape, npar=1, prec=8, rate="2", decomp=None, cb=None, precincts=None, qm=None,
                tiles=None, erres=False, kernels=None, prog=None, qstyle=None, qstep=None, com=None, aux=None):
    """Oracle settings; the setters apply in the order kernels, progression,
    quantisation style, decomposition, precincts, codeblocks, Qm, tiles,
    error resilience, quantisation step size, then AUX and COM."""
    p = OcParams()
    n = list(shape) + [1] * (4 - len(shape))
    for i in range(4):
        p.n[i] = int(n[i])
    p.nPar = npar
    p.data_prec = prec
    if decomp is not None:
        p.have_decomp = 1
        for i in range(4):
            p.decomp[i] = decomp[i]
    if cb is not None:
        p.have_cb = 1
        for i in range(4):
            p.cb[i] = cb[i]
    if precincts is not None:
        p.have_prec = 1
        for i in range(4):
            p.prec[i] = precincts[i]
    if qm is not None:
        p.have_qm = 1
        p.qm = qm
    if tiles is not None:
        p.have_tiles = 1
        for i in range(4):
            p.tiles[i] = tiles[i]
    p.erres = 1 if erres else 0
    p.rate = rate.encode() if rate is not None else None
    if kernels is not None:
        p.have_kern = 1
        for i in range(4):
            p.kern[i] = kernels[i]
    if prog is not None:
        p.have_prog = 1
        p.prog = prog
    if qstyle is not None:
        p.have_qstyle = 1
        p.qstyle = qstyle
    if qstep is not None:
        p.have_qstep = 1
        p.qstep = float(qstep)
    if aux is not None:
        p.aux = aux.encode() if isinstance(aux, str) else bytes(aux)
        p.naux = len(p.aux)
    if com is not None:
        p.com = com.encode() if isinstance(com, str) else bytes(com)
        p.ncom = len(p.com)
    return p


class Oracle:
    """One oracle codec (geometry + stage state) for a compression job."""

    def __init__(self, shape, **kw):
        self.params = make_params(shape, **kw)
        self._rate_keep = self.params.rate
        self.h = lib().oc_create(C.byref(self.params))
        if not self.h:
            raise ValueError("oracle rejected the parameters")
        self.npar = self.params.nPar
        self.prec = self.params.data_prec

    def close(self):
        if self.h:
            lib().oc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def compress(self, data: np.ndarray) -> bytes:
        data = np.ascontiguousarray(data)
        cap = data.nbytes * 2 + (1 << 20)
        out = np.zeros(cap, dtype=np.uint8)
        n = lib().oc_compress(self.h, data.ctypes.data_as(C.c_void_p), ptr(out, C.c_uint8), cap)
        if n == 0:
            raise RuntimeError("oracle compression failed")
        return out[:n].tobytes()

    # ---- stages ----
    @property
    def ntiles(self):
        return lib().oc_ntiles(self.h)

    @property
    def nlayers(self):
        return lib().oc_nlayers(self.h)

    def param_dims(self, tile=0):
        d = (C.c_uint64 * 4)()
        lib().oc_param_dims(self.h, tile, d)
        return tuple(int(x) for x in d)

    def ncblk(self, tile=0, param=0):
        return lib().oc_ncblk(self.h, tile, param)

    def geometry(self, tile=0, param=0):
        n = self.ncblk(tile, param)
        boxes = np.zeros((n, 8), np.uint64)
        sb = np.zeros((n, 3), np.float64)
        meta = np.zeros((n, 4), np.uint32)
        b = (C.c_uint64 * 8)()
        s = (C.c_double * 3)()
        m = (C.c_uint32 * 4)()
        for i in range(n):
            lib().oc_cblk_geometry(self.h, tile, param, i, b, s, m)
            boxes[i] = list(b)
            sb[i] = list(s)
            meta[i] = list(m)
        return boxes, sb, meta

    def fill_normalize(self, data, tile=0, param=0):
        dims = self.param_dims(tile)
        work = np.zeros(int(np.prod(dims)), np.float64)
        mm = (C.c_double * 4)()
        lib().oc_fill_normalize(self.h, tile, param, np.ascontiguousarray(data).ctypes.data_as(C.c_void_p),
                                ptr(work, C.c_double), mm)
        return work, tuple(mm)

    def dwt_forward(self, work, tile=0, param=0):
        lib().oc_dwt_forward(self.h, tile, param, ptr(work, C.c_double))
        return work

    def dwt_inverse(self, work, tile=0, param=0):
        lib().oc_dwt_inverse(self.h, tile, param, ptr(work, C.c_double))
        return work

    def t1_encode(self, work, tile=0, param=0):
        lib().oc_t1_encode(self.h, tile, param, ptr(work, C.c_double))

    def cblk_record(self, cb, tile=0, param=0):
        kkz = (C.c_uint8 * 3)()
        S = np.zeros(191, np.uint16)
        L = np.zeros(190, np.uint64)
        d = u8p()
        nb = lib().oc_cblk_record(self.h, tile, param, cb, kkz, ptr(S, C.c_uint16), ptr(L, C.c_uint64),
                                  C.byref(d))
        K, Kmsbs, Z = kkz
        data = C.string_at(d, nb) if nb else b""
        return dict(K=K, Kmsbs=Kmsbs, Z=Z, S=S[: Z + 1].copy(), L=L[:Z].copy(), data=data)

    def tile_slopes(self, tile=0):
        mm = (C.c_uint16 * 2)()
        lib().oc_tile_slopes(self.h, tile, mm)
        return int(mm[0]), int(mm[1])

    def t2_encode(self, tile=0):
        n = sum(self.ncblk(tile, p) for p in range(self.npar))
        cuts = np.zeros(max(1, n * self.nlayers), np.int16)
        lib().oc_t2_encode(self.h, tile, ptr(cuts, C.c_int16))
        return cuts[: n * self.nlayers].reshape(n, self.nlayers)

    def assemble(self, cap):
        out = np.zeros(cap, np.uint8)
        n = lib().oc_assemble(self.h, ptr(out, C.c_uint8), cap)
        return out[:n].tobytes()


def rc_trace(cap=4096):
    """(layer, threshold, estimate) of every probe of the oracle's last
    tile's PCRD bisection (this thread's last compression)."""
    lay = np.zeros(cap, np.uint32)
    slo = np.zeros(cap, np.uint32)
    est = np.zeros(cap, np.int64)
    n = lib().oc_rc_trace(ptr(lay, C.c_uint32), ptr(slo, C.c_uint32), ptr(est, C.c_int64), cap)
    return list(zip(lay[:n].tolist(), slo[:n].tolist(), est[:n].tolist()))


def oracle_compress(data, shape, **kw) -> bytes:
    o = Oracle(shape, **kw)
    try:
        return o.compress(data)
    finally:
        o.close()


def t1_encode_block(buf, dims, box, step_eff, step, gain, hb, kmax, erres=False):
    """Encode one codeblock out of a float64 buffer with dims (w, h, d)."""
    buf = np.ascontiguousarray(buf, dtype=np.float64)
    b = (C.c_uint64 * 8)(*[int(x) for x in box])
    kkz = (C.c_uint8 * 3)()
    S = np.zeros(191, np.uint16)
    L = np.zeros(190, np.uint64)
    n = int(np.prod([box[4 + i] - box[i] for i in range(4)]))
    cap = n * 9 + 1024
    data = np.zeros(cap, np.uint8)
    nb = lib().oc_t1_encode_block(ptr(buf, C.c_double), dims[0], dims[1], dims[2], b, step_eff, step, gain, hb,
                                  kmax, 1 if erres else 0, kkz, ptr(S, C.c_uint16), ptr(L, C.c_uint64),
                                  ptr(data, C.c_uint8), cap)
    if nb < 0:
        raise RuntimeError("codeblock larger than buffer")
    K, Kmsbs, Z = kkz
    return dict(K=K, Kmsbs=Kmsbs, Z=Z, S=S[: Z + 1].copy(), L=L[:Z].copy(), data=data[:nb].tobytes())


def t1_probe_block(buf, dims, box, step_eff, step, gain, hb, kmax, zlim=0, erres=False):
    """One codeblock coded to `zlim` passes (0: all), with the pass-limit
    probes: per-pass integer distortion sums `dm`, truncation-search read
    extents `reach`, bytes emitted before termination `nfinal`, and the
    bit-length histogram `hist` of the quantised magnitudes."""
    buf = np.ascontiguousarray(buf, dtype=np.float64)
    b = (C.c_uint64 * 8)(*[int(x) for x in box])
    kkz = (C.c_uint8 * 3)()
    S = np.zeros(191, np.uint16)
    L = np.zeros(190, np.uint64)
    dm = np.zeros(190, np.int64)
    reach = np.zeros(190, np.int64)
    nfinal = C.c_int64(0)
    hist = np.zeros(65, np.uint32)
    lib().oc_t1_probe_block(ptr(buf, C.c_double), dims[0], dims[1], dims[2], b, step_eff, step, gain, hb, kmax,
                            1 if erres else 0, int(zlim), kkz, ptr(S, C.c_uint16), ptr(L, C.c_uint64),
                            ptr(dm, C.c_int64), ptr(reach, C.c_int64), C.byref(nfinal), ptr(hist, C.c_uint32))
    K, Kmsbs, Z = kkz
    return dict(K=K, Z=Z, S=S[: Z + 1].copy(), L=L[:Z].copy(), dm=dm[:Z].copy(), reach=reach[:Z].copy(),
                nfinal=int(nfinal.value), hist=hist.copy())


# ---------------------------------------------------------------------------
# reference driver (only where oracle/_ref was built from the reference source)
# ---------------------------------------------------------------------------
def have_ref():
    return REF_DUMP.exists()


def _opts(decomp=None, cb=None, precincts=None, qm=None, tiles=None, erres=False, kernels=None, prog=None,
          qstyle=None, qstep=None, com=None, aux=None):
    o = []
    if kernels is not None:
        o += ["--kernels", ",".join(str(int(x)) for x in kernels)]
    if prog is not None:
        o += ["--prog", str(int(prog))]
    if qstyle is not None:
        o += ["--qstyle", str(int(qstyle))]
    if qstep is not None:
        o += ["--qstep", repr(float(qstep))]
    if com is not None:
        o += ["--com", com]
    if aux is not None:
        o += ["--aux", aux]
    if decomp is not None:
        o += ["--decomp", ",".join(str(int(x)) for x in decomp)]
    if precincts is not None:
        o += ["--prec", ",".join(str(int(x)) for x in precincts)]
    if cb is not None:
        o += ["--cb", ",".join(str(int(x)) for x in cb)]
    if qm is not None:
        o += ["--qm", str(int(qm))]
    if tiles is not None:
        o += ["--tiles", ",".join(str(int(x)) for x in tiles)]
    if erres:
        o += ["--erres"]
    return o


def ref_run(mode, data, shape, npar=1, prec=8, rate="2", **kw):
    """Run the reference driver; returns dict of output files' bytes."""
    n = list(shape) + [1] * (4 - len(shape))
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "in.raw")
        np.ascontiguousarray(data).tofile(inp)
        pre = os.path.join(td, "o")
        cmd = [str(REF_DUMP), mode, inp] + [str(int(x)) for x in n] + [str(npar), str(prec), rate, pre] + _opts(**kw)
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        out = {}
        for ext in (".bwc", ".dwt", ".t1", ".geo"):
            if os.path.exists(pre + ext):
                out[ext[1:]] = Path(pre + ext).read_bytes()
        return out


def ref_decode(cs, shape, npar=1, prec=8):
    """Decode a codestream with the reference decoder (bwc_create_decompression
    + bwc_decompress in oracle/ref_dump.c); returns the output bytes."""
    n = list(shape) + [1] * (4 - len(shape))
    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "o")
        Path(pre + ".bwc").write_bytes(cs)
        cmd = [str(REF_DUMP), "decode", "-"] + [str(int(x)) for x in n] + [str(npar), str(prec), "2", pre]
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        return Path(pre + ".dec").read_bytes()


def oracle_decompress(cs, layer=0):
    """The oracle's decoder: (decoded values as a flat array, dims (nX, nY,
    nZ, nTS, nPar, prec))."""
    L = lib()
    buf = np.frombuffer(cs, np.uint8).copy()
    dims = (C.c_uint64 * 6)()
    # the header gives the output size; decode once into a buffer large
    # enough for any field the header can describe here
    hdr = _header_dims(cs)
    nout = int(np.prod(hdr[:5])) * hdr[5]
    out = np.zeros(max(nout, 1), np.uint8)
    rc = L.oc_decompress(buf.ctypes.data_as(C.POINTER(C.c_uint8)), buf.size, layer, out.ctypes.data, dims)
    if rc != 0:
        raise RuntimeError("oc_decompress failed")
    d = tuple(int(x) for x in dims)
    return out.view(np.float64 if d[5] == 8 else np.float32), d


def _header_dims(cs):
    """(nX, nY, nZ, nTS, nPar, data_prec) from the SGI marker segment."""
    i = 10   # after SOC and the codestream length
    assert cs[i:i + 2] == b"\xff\x51", "no SGI marker segment after SOC"
    v = [int.from_bytes(cs[i + 4 + 8 * a:i + 12 + 8 * a], "big") for a in range(4)]
    return v + [cs[i + 36], cs[i + 37]]


def parse_t1(blob, ncblk_per_param):
    """Parse the per-codeblock tier-1 records oracle/ref_dump.c writes."""
    recs = []
    slopes = []
    off = 0
    for n in ncblk_per_param:
        rr = []
        for _ in range(n):
            K, Kmsbs, Z, _pad = blob[off:off + 4]
            off += 4
            S = np.frombuffer(blob, np.uint16, Z + 1, off)
            off += 2 * (Z + 1)
            L = np.frombuffer(blob, np.uint64, Z, off)
            off += 8 * Z
            nb = int(np.frombuffer(blob, np.uint64, 1, off)[0])
            off += 8
            data = blob[off:off + nb]
            off += nb
            rr.append(dict(K=K, Kmsbs=Kmsbs, Z=Z, S=S.copy(), L=L.copy(), data=bytes(data)))
        slopes.append(tuple(int(x) for x in np.frombuffer(blob, np.uint16, 2, off)))
        off += 4
        recs.append(rr)
    return recs, slopes


def parse_geo(blob):
    rec = np.dtype([("box", "<u8", 8), ("sb", "<f8", 3), ("meta", "<u4", 4)])
    return np.frombuffer(blob, rec)
