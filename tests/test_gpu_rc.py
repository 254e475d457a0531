"""PCRD rate control on the device (pcrd.hip) against the CPU oracle and the
host restatement.

* Every probe of the bisection (layer, slope threshold, size estimate) equals
  the oracle's, which follows tier2.c:1115-1254, on every golden case, and
  the codestream equals the reference's.
* The learned pass-skip encoder with the check on the device gives the same
  codestream, the same extra rounds and the same re-coded codeblocks as with
  the check on the host (BWC_AMD_RC=host), at several margins."""
import sys

import numpy as np
import pytest

import oracle_lib as ol
from bigwhoop_amd import bwc, skipnet
from bigwhoop_amd.fields import grf
from golden_cases import cases, codec_kw, make_input, sha

pytestmark = pytest.mark.gpu
CASES = cases()


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if bwc.lib().bwc_amd_device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need the MI355X")


def _kw(case):
    return dict(npar=case["npar"], prec=case["prec"], rate=case["rate"], **codec_kw(case))


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_probes_match_oracle(case, monkeypatch):
    monkeypatch.delenv("BWC_AMD_RC", raising=False)
    _say("rc probes", case["name"])
    data = make_input(case)
    c = bwc.Codec(case["gen"]["shape"], **_kw(case))
    try:
        cs = c.compress(data)
        got = c.rc_trace()
    finally:
        c.close()
    assert len(cs) == case["codestream_size"] and sha(cs) == case["codestream_sha256"]
    ref = ol.oracle_compress(data, case["gen"]["shape"], **_kw(case))
    assert ref == cs
    want = ol.rc_trace()
    assert got == want


@pytest.mark.parametrize("rate", ["2", "0.5,3", "1,2,4,8"])
def test_probes_match_oracle_at_headline_settings(rate, monkeypatch):
    monkeypatch.delenv("BWC_AMD_RC", raising=False)
    shape = (96, 80, 72)
    x = grf(shape, seed=11)
    kw = dict(rate=rate, decomp=(5, 5, 5, 0), cb=(5, 5, 3, 0))
    _say("rc probes headline settings", rate)
    c = bwc.Codec(shape, **kw)
    try:
        cs = c.compress(x)
        got = c.rc_trace()
    finally:
        c.close()
    assert cs == ol.oracle_compress(x, shape, **kw)
    assert got == ol.rc_trace()
    assert len(got) >= 8 * len(rate.split(","))


def _learned(shape, data, kw, margin, mode, monkeypatch):
    if mode == "host":
        monkeypatch.setenv("BWC_AMD_RC", "host")
    else:
        monkeypatch.delenv("BWC_AMD_RC", raising=False)
    c = bwc.Codec(shape, **kw)
    try:
        if margin < 0:   # fixed limit of -margin passes: many re-coding rounds
            c.set_pass_skip(1, -margin)
        else:
            c.set_skip_model(skipnet.wide_weights())
            c.set_pass_skip(2, margin)
        cs = c.compress(data)
        st = c.skip_stats()
    finally:
        c.close()
    return cs, st


@pytest.mark.parametrize("margin", [-3, 0, 2, 6])
@pytest.mark.parametrize("which", ["grf3d_rate2", "grf2d_tiles_layers", "grf4d_two_vars"])
def test_skip_check_on_device_matches_host(which, margin, monkeypatch):
    if which == "grf3d_rate2":
        shape, kw = (64, 64, 64), dict(rate="2", decomp=(5, 5, 5, 0), cb=(5, 5, 3, 0))
        data = grf(shape, seed=5)
    elif which == "grf2d_tiles_layers":
        shape, kw = (300, 220), dict(rate="0.75,2.5", decomp=(4, 4, 0, 0), cb=(5, 5, 0, 0),
                                     tiles=(160, 128, 1, 1), erres=True)
        data = grf(shape, seed=6)
    else:
        shape = (40, 36, 24, 8)
        kw = dict(npar=2, rate="1.5", decomp=(3, 3, 2, 1), cb=(4, 4, 3, 2))
        data = np.concatenate([grf(shape, seed=7).ravel(), grf(shape, slope=-3.0, seed=8).ravel()])
    _say("skip check", which, margin)
    a, sa = _learned(shape, data, kw, margin, "device", monkeypatch)
    b, sb = _learned(shape, data, kw, margin, "host", monkeypatch)
    assert a == b
    assert sa == sb
    ref = ol.oracle_compress(data, shape, **kw)
    assert a == ref


def _stuffed_bytes(bits):
    """Bytes the packet-header writer emits for a bit string: MSB-first
    bytes, a byte after an 0xFF one holding 7 bits, and a last byte for
    pending bits or for a 7-bit byte due after a final 0xFF."""
    n, pos, cap, N = 0, 0, 8, len(bits)
    while pos < N:
        n += 1
        nxt = 7 if cap == 8 and pos + 8 <= N and bits[pos:pos + 8].all() else 8
        pos += cap
        cap = nxt
    if pos == N and cap == 7:
        n += 1
    return n


def test_stuffed_byte_count_matches_writer():
    rng = np.random.default_rng(5)
    strings = []
    for k in range(160):
        n = int(rng.choice([1, 7, 8, 9, 15, 16, 17, 31, 32, 33, 63, 64, 65])) if k < 40 else \
            int(rng.integers(1, 150_000))
        kind = k % 4
        if kind == 0:     # mostly ones: 0xFF bytes everywhere
            b = rng.random(n) < 0.97
        elif kind == 1:   # runs of ones and zeros
            b = np.repeat(rng.integers(0, 2, n // 5 + 1).astype(bool), rng.integers(1, 40, n // 5 + 1))[:n]
            if len(b) < n:
                b = np.concatenate([b, np.ones(n - len(b), bool)])
        elif kind == 2:   # all ones
            b = np.ones(n, bool)
        else:             # mostly zeros with some 8-bit runs of ones
            b = rng.random(n) < 0.02
            for s in rng.integers(0, n, max(1, n // 200)):
                b[s:s + 8] = True
        strings.append(b.astype(np.uint8))
    woff = [0]
    words = []
    for b in strings:
        nw = (len(b) + 31) // 32
        pad = np.zeros(nw * 32, np.uint8)
        pad[:len(b)] = b
        w = np.packbits(pad).view(">u4").astype(np.uint32)
        words.append(w)
        woff.append(woff[-1] + nw + int(rng.integers(0, 3)))
        words.append(np.zeros(woff[-1] - woff[-2] - nw, np.uint32))
    wa = np.ascontiguousarray(np.concatenate(words), dtype=np.uint32)
    wo = np.array(woff, np.uint64)
    nb = np.array([len(b) for b in strings], np.uint64)
    out = np.zeros(len(strings), np.uint64)
    import ctypes as C
    rc = bwc.lib().bwc_amd_stuffed_bytes(wa.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         wo.ctypes.data_as(C.POINTER(C.c_uint64)),
                                         nb.ctypes.data_as(C.POINTER(C.c_uint64)), len(strings),
                                         out.ctypes.data_as(C.POINTER(C.c_uint64)))
    assert rc == 0, bwc.last_error()
    want = [_stuffed_bytes(b.astype(bool)) for b in strings]
    assert out.tolist() == want
