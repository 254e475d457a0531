"""Host-side parts of libbwc without a GPU: the C-ABI library loads and exports
every declared symbol, its codec geometry equals the oracle's, and its tier-2
+ codestream writer, fed with the oracle's tier-1 records, reproduces the
reference codestreams byte for byte (golden fixtures)."""
import re
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as ol
from bigwhoop_amd import bwc
from golden_cases import GOLDEN, cases, codec_kw, make_input, sha

ROOT = Path(__file__).resolve().parents[1]
CASES = cases()


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt):
            n = m.group(1)
            if n.startswith(("bwc_", "create_codec", "configure_codec", "set_")):
                names.add(n)
    return names


def test_library_exports_declared_symbols():
    import ctypes
    L = ctypes.CDLL(str(bwc.LIB_PATH))
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def _oracle_records(o, data, t, p):
    work, mm = o.fill_normalize(data, t, p)
    o.dwt_forward(work, t, p)
    o.t1_encode(work, t, p)
    n = o.ncblk(t, p)
    K = np.zeros(n, np.uint8)
    Km = np.zeros(n, np.uint8)
    Z = np.zeros(n, np.uint8)
    L = np.zeros((n, 190), np.uint32)
    S = np.zeros((n, 191), np.uint16)
    chunks, off = [], [0]
    for i in range(n):
        r = o.cblk_record(i, t, p)
        K[i], Km[i], Z[i] = r["K"], r["Kmsbs"], r["Z"]
        L[i, : r["Z"]] = r["L"]
        S[i, : r["Z"] + 1] = r["S"]
        chunks.append(r["data"])
        off.append(off[-1] + len(r["data"]))
    return K, Km, Z, L, S, b"".join(chunks), np.array(off, np.uint64), mm


# BWC_AMD_RC_THREADS: 0 = the sequential in-place probes, 1 = side-effect-free
# probes on one thread, 7 = the bisection's next three levels on threads;
# "nc": tag trees encoded in every probe instead of the per-layer bit cache
@pytest.mark.parametrize("threads", ["0", "1", "7", "7nc", "0nc"])
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_geometry_and_tier2_against_golden(case, threads, monkeypatch):
    monkeypatch.setenv("BWC_AMD_RC_THREADS", threads.replace("nc", ""))
    monkeypatch.setenv("BWC_AMD_RC_NOCACHE", "1" if threads.endswith("nc") else "0")
    data = make_input(case)
    shape = case["gen"]["shape"]
    kw = codec_kw(case)
    o = ol.Oracle(shape, npar=case["npar"], prec=case["prec"], rate=case["rate"], **kw)
    c = bwc.Codec(shape, npar=case["npar"], prec=case["prec"], rate=case["rate"], **kw)
    try:
        assert o.ntiles == len([1 for _ in range(o.ntiles)])
        c.begin_records()
        for t in range(o.ntiles):
            assert c.param_dims(t) == o.param_dims(t)
            for p in range(case["npar"]):
                ob, osb, om = o.geometry(t, p)
                cbx, csb, cm = c.geometry(t, p)
                np.testing.assert_array_equal(cbx.astype(np.uint64), ob)
                np.testing.assert_array_equal(csb, osb)
                np.testing.assert_array_equal(cm, om)
                c.set_records(t, p, *_oracle_records(o, data, t, p))
        cs = c.encode_records()
        assert len(cs) == case["codestream_size"]
        assert sha(cs) == case["codestream_sha256"]
    finally:
        c.close()
        o.close()


@pytest.mark.parametrize("threads", ["0", "7", "7nc"])
def test_rate_control_modes_agree_with_oracle_multilayer(threads, monkeypatch):
    """A larger multi-layer case than the golden ones: oracle records of a
    96^3 field at four rates through the host tier-2 in each probe mode must
    give the oracle's codestream."""
    from bigwhoop_amd.fields import grf
    monkeypatch.setenv("BWC_AMD_RC_THREADS", threads.replace("nc", ""))
    monkeypatch.setenv("BWC_AMD_RC_NOCACHE", "1" if threads.endswith("nc") else "0")
    shape = (96, 96, 96)
    kw = dict(rate="0.5,1,2,4", decomp=(4, 4, 4, 0), cb=(5, 5, 3, 0))
    data = grf(shape, slope=-2.2, seed=21)
    o = ol.Oracle(shape, **kw)
    c = bwc.Codec(shape, **kw)
    try:
        ref = o.compress(data)
        c.begin_records()
        c.set_records(0, 0, *_oracle_records(o, data, 0, 0))
        assert c.encode_records() == ref
    finally:
        c.close()
        o.close()


@pytest.mark.parametrize("n, ok", [(33554434, True), (48412990, False), (48412992, True), (16777217, True),
                                   (16777219, False)])
def test_float_rounded_extents(n, ok):
    """Resolution boxes are computed in single precision (libbwc.c:1539-1546).
    A field extent that rounds down to a float (2^25 + 2, 2^24 + 1) leaves a
    shorter box and is coded as the reference codes it; one that rounds up
    (48412990 -> 48412992, 2^24 + 3 -> 2^24 + 4) would take the transform past
    the samples and is refused by the library and the oracle alike."""
    kw = dict(rate="2", decomp=(1, 0, 0, 0), cb=(6, 0, 0, 0))
    assert (int(np.float32(n)) <= n) == ok
    if ok:
        bwc.Codec((n, 1, 1, 1), **kw).close()
        ol.Oracle((n, 1, 1, 1), **kw).close()
    else:
        with pytest.raises(RuntimeError, match="single precision"):
            bwc.Codec((n, 1, 1, 1), **kw)
        with pytest.raises(ValueError, match="rejected"):
            ol.Oracle((n, 1, 1, 1), **kw)


def test_batch_rejects_a_codec_twice():
    """bwc_amd_compress_batch refuses a codec that appears twice (its
    workspace and records serve one compression at a time), before any GPU
    work."""
    kw = dict(rate="2", decomp=(1, 1, 0, 0), cb=(4, 4, 0, 0))
    a, b = bwc.Codec((32, 32), **kw), bwc.Codec((32, 32), **kw)
    x = np.zeros(32 * 32)
    try:
        with pytest.raises(RuntimeError, match="appears twice"):
            bwc.compress_batch([a, b, a], [x.ctypes.data] * 3)
    finally:
        a.close()
        b.close()
