"""GPU parity: every stage of the HIP path against the CPU oracle and the
reference's golden vectors, through the C-ABI (bigwhoop_amd/lib/libbwc.so).

Bars: DWT coefficients bit-identical (0 ULP), tier-1 records (K, Kmsbs, Z,
L, S and MQ bytes) bit-identical, codestreams byte-identical."""
import hashlib
import sys

import numpy as np
import pytest

import oracle_lib as ol
from bigwhoop_amd import bwc
from golden_cases import cases, codec_kw, load, make_input, sha

pytestmark = pytest.mark.gpu
CASES = cases()
STAGE = [c for c in CASES if any(f.endswith(".t1.npz") for f in c["files"])]


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if bwc.lib().bwc_amd_device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need the MI355X")


def _codec(case):
    return bwc.Codec(case["gen"]["shape"], npar=case["npar"], prec=case["prec"], rate=case["rate"], **codec_kw(case))


@pytest.mark.parametrize("path", ["fast", "generic"])
@pytest.mark.parametrize("case", STAGE, ids=[c["name"] for c in STAGE])
def test_dwt_bit_exact(case, path, monkeypatch):
    # "fast": LDS X pass + streaming strided passes; "generic": in-place LDS kernel only
    monkeypatch.setenv("BWC_AMD_DWT", "1" if path == "fast" else "0")
    _say("dwt", path, case["name"])
    data = make_input(case)
    c = _codec(case)
    o = ol.Oracle(case["gen"]["shape"], npar=case["npar"], prec=case["prec"], rate=case["rate"], **codec_kw(case))
    allc = []
    try:
        for t in range(o.ntiles):
            for p in range(case["npar"]):
                g, mm = c.stage_dwt(data, t, p)
                w, omm = o.fill_normalize(data, t, p)
                o.dwt_forward(w, t, p)
                np.testing.assert_array_equal(mm, np.array(omm))
                assert g.view(np.uint64).tolist() == w.view(np.uint64).tolist() or \
                    np.array_equal(g.view(np.uint64), w.view(np.uint64))
                allc.append(g)
        dw = load(case, "dwt")
        assert hashlib.sha256(np.concatenate(allc).tobytes()).digest() == dw["sha256"].tobytes()
    finally:
        c.close()
        o.close()


# line lengths from 2 upwards, odd and even, tile offsets that make the
# lowpass start on odd positions, 1-4 transformed axes; X lines up to 1024
# samples and an untransformed axis between two transformed ones for the
# fused X-Y kernel
DWT_SHAPES = [
    dict(shape=(37, 5, 9), decomp=(3, 2, 3, 0)),
    dict(shape=(3, 70, 2), decomp=(1, 4, 1, 0)),
    dict(shape=(130, 1, 1), decomp=(5, 0, 0, 0)),
    dict(shape=(2, 3, 67), decomp=(1, 1, 5, 0)),
    dict(shape=(21, 19, 17, 6), decomp=(2, 2, 2, 1)),
    dict(shape=(45, 33, 29), decomp=(3, 3, 3, 0), tiles=(24, 24, 24, 1)),
    dict(shape=(600, 7, 3), decomp=(4, 1, 1, 0)),
    dict(shape=(301, 40, 3), decomp=(3, 3, 1, 0), tiles=(150, 24, 3, 1)),
    dict(shape=(1024, 11, 2), decomp=(1, 2, 1, 0)),
    dict(shape=(96, 150, 5, 4), decomp=(3, 3, 0, 1)),
]


def _dwt_path(path, monkeypatch):
    """fast: LDS X pass, streaming strided passes, X and Y fused on large
    boxes only; xy: the fused X-Y kernel on every box it applies to (3-D
    boxes with X lines of at most 1024 samples); generic: the in-place LDS
    kernel only"""
    monkeypatch.setenv("BWC_AMD_DWT", "0" if path == "generic" else "1")
    before = bwc.get_tuning("dwt_xy_min_log2")
    bwc.set_tuning("dwt_xy_min_log2", 0 if path == "xy" else 24)
    return before


@pytest.mark.parametrize("path", ["fast", "xy", "generic"])
@pytest.mark.parametrize("spec", DWT_SHAPES, ids=[str(s["shape"]) for s in DWT_SHAPES])
def test_dwt_shapes_bit_exact(spec, path, monkeypatch):
    before = _dwt_path(path, monkeypatch)
    _say("dwt shapes", path, spec)
    from bigwhoop_amd.fields import grf
    shape = spec["shape"]
    data = grf(shape, seed=7)
    kw = dict(decomp=spec["decomp"])
    if "tiles" in spec:
        kw["tiles"] = spec["tiles"]
    c = bwc.Codec(shape, rate="2", **kw)
    o = ol.Oracle(shape, rate="2", **kw)
    try:
        for t in range(o.ntiles):
            g, _ = c.stage_dwt(data, t, 0)
            w, _ = o.fill_normalize(data, t, 0)
            o.dwt_forward(w, t, 0)
            assert np.array_equal(g.view(np.uint64), w.view(np.uint64)), (t, spec)
    finally:
        bwc.set_tuning("dwt_xy_min_log2", before)
        c.close()
        o.close()


@pytest.mark.parametrize("path", ["fast", "lane"])
@pytest.mark.parametrize("case", STAGE, ids=[c["name"] for c in STAGE])
def test_tier1_records_bit_exact(case, path, monkeypatch):
    # "fast": wavefront symbol generation + MQ kernel; "lane": one lane per codeblock
    monkeypatch.setenv("BWC_AMD_T1", "1" if path == "lane" else "0")
    _say("t1", path, case["name"])
    data = make_input(case)
    c = _codec(case)
    o = ol.Oracle(case["gen"]["shape"], npar=case["npar"], prec=case["prec"], rate=case["rate"], **codec_kw(case))
    try:
        for t in range(o.ntiles):
            for p in range(case["npar"]):
                w, _ = o.fill_normalize(data, t, p)
                o.dwt_forward(w, t, p)
                g = c.stage_t1(w.copy(), t, p)
                o.t1_encode(w, t, p)
                for i in range(o.ncblk(t, p)):
                    r = o.cblk_record(i, t, p)
                    assert (g["K"][i], g["Kmsbs"][i], g["Z"][i]) == (r["K"], r["Kmsbs"], r["Z"]), (t, p, i)
                    Z = r["Z"]
                    np.testing.assert_array_equal(g["L"][i, :Z], r["L"].astype(np.uint32))
                    np.testing.assert_array_equal(g["S"][i, : Z + 1], r["S"])
                    d = g["data"][int(g["off"][i]): int(g["off"][i + 1])].tobytes()
                    assert d == r["data"], (t, p, i)
    finally:
        c.close()
        o.close()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_codestream_matches_reference(case):
    _say("compress", case["name"])
    data = make_input(case)
    c = _codec(case)
    try:
        cs = c.compress(data)
        assert len(cs) == case["codestream_size"]
        assert sha(cs) == case["codestream_sha256"]
        # a second call on the same codec gives the same bytes
        assert c.compress(data) == cs
    finally:
        c.close()


@pytest.mark.parametrize("n", [256])
def test_headline_settings_match_oracle_at_scale(n):
    """256^3 f64 k^-5/3 GRF at the headline settings (5 levels, cb 32x32x8, 1
    layer at 2 bpv, and 4 layers 1,2,4,8): GPU codestream == CPU oracle's."""
    from bigwhoop_amd.fields import grf

    x = grf((n, n, n), seed=3)
    for rate in ("2", "1,2,4,8"):
        _say("scale", n, rate)
        kw = dict(rate=rate, decomp=(5, 5, 5, 0), cb=(5, 5, 3, 0))
        c = bwc.Codec((n, n, n), **kw)
        try:
            got = c.compress(x)
        finally:
            c.close()
        ref = ol.oracle_compress(x, (n, n, n), **kw)
        assert len(got) == len(ref)
        assert got == ref


@pytest.mark.parametrize("qm", [50, 56, 60, 62])
def test_large_qm_matches_oracle(qm):
    """Quantised magnitudes near and past 2^64 (the bitplane kernel's bit
    count takes the largest quotient below 2^64): codestream against the
    oracle at large Qm."""
    from bigwhoop_amd.fields import grf
    shape = (40, 36, 20)
    kw = dict(rate="4", decomp=(3, 3, 2, 0), cb=(4, 4, 3, 0), qm=qm)
    _say("large qm", qm)
    x = grf(shape, seed=3).reshape(-1)
    ref = ol.oracle_compress(x, shape, **kw)
    c = bwc.Codec(shape, **kw)
    try:
        assert c.compress(x) == ref
    finally:
        c.close()
