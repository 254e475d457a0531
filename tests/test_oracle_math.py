"""Pins of the oracle that do not rest on the reference build: closed forms,
invariants and textbook cases (DWT, MQ coder, gain LUT, slopes)."""
import numpy as np
import pytest

import oracle_lib as ol


def _work_after_dwt(o, field):
    w, _ = o.fill_normalize(field)
    return o.dwt_forward(w)


def test_dwt_perfect_reconstruction():
    rng = np.random.default_rng(0)
    for shape, dec in (((37, 29, 19), (3, 3, 2, 0)), ((64,), (5, 0, 0, 0)), ((9, 7, 5, 6), (2, 2, 2, 2))):
        f = rng.standard_normal(int(np.prod(shape)))
        o = ol.Oracle(shape, rate="2", decomp=dec)
        w, _ = o.fill_normalize(f)
        w0 = w.copy()
        o.dwt_forward(w)
        assert not np.allclose(w, w0)
        o.dwt_inverse(w)
        # exact up to K_H*K_L = 1 - 1.2e-10 (float-rounded kappas) per level and axis
        np.testing.assert_allclose(w, w0, rtol=0, atol=1e-8 * np.abs(w0).max())


def test_dwt_dc_gain_and_vanishing_moments():
    # 1-D, 1 level: constant -> lowpass == constant, highpass == 0 (up to the
    # float-rounded lifting constants); a linear ramp is annihilated by the
    # highpass away from the boundary
    n = 64
    o = ol.Oracle((n,), rate="2", decomp=(1, 0, 0, 0))
    field = np.concatenate([[0.0, 2.0], np.full(n - 2, 1.0)])  # min 0, max 2 -> normalised 1 -> 2^62
    w, _ = o.fill_normalize(field)
    w[:] = 1.0
    o.dwt_forward(w)
    np.testing.assert_allclose(w[: n // 2], 1.0, rtol=1e-6)
    assert np.abs(w[n // 2:]).max() < 1e-6
    w[:] = np.arange(n, dtype=np.float64)
    o.dwt_forward(w)
    assert np.abs(w[n // 2 + 4: n - 4]).max() < 1e-5 * n


def test_mq_round_trip_and_state_table():
    rng = np.random.default_rng(1)
    for n in (1, 7, 1000, 50000):
        ctx = rng.integers(0, 19, n).astype(np.uint8)
        p = rng.uniform(0.02, 0.98, 19)[ctx]
        bits = (rng.random(n) < p).astype(np.uint8)
        out = np.zeros(2 * n + 64, np.uint8)
        nb = ol.lib().oc_mq_encode_symbols(ol.ptr(bits, ol.C.c_uint8), ol.ptr(ctx, ol.C.c_uint8), n,
                                           ol.ptr(out, ol.C.c_uint8), out.size)
        dec = np.zeros(n, np.uint8)
        ol.lib().oc_mq_decode_symbols(ol.ptr(out, ol.C.c_uint8), nb, ol.ptr(ctx, ol.C.c_uint8), n,
                                      ol.ptr(dec, ol.C.c_uint8))
        np.testing.assert_array_equal(dec, bits)
        # a skewed source compresses below 1 bit/symbol
        if n >= 1000:
            assert nb * 8 < n


def test_zero_codeblock():
    buf = np.zeros(16 * 16 * 4)
    r = ol.t1_encode_block(buf, (16, 16, 4), (0, 0, 0, 0, 16, 16, 4, 1), 1.0, 1.0, 1.0, 0, 64)
    assert (r["K"], r["Z"]) == (0, 0)
    assert r["S"][0] == 0xFFFF and r["data"] == b""


def test_single_coefficient_codeblock():
    # one coefficient of magnitude 5 (K = 3) in an otherwise empty block: the
    # plane order is CUP(2) then {SPP,MRP,CUP}(1), {SPP,MRP,CUP}(0) -> Z = 7
    buf = np.zeros(8 * 8)
    buf[3 * 8 + 2] = -5.0
    r = ol.t1_encode_block(buf, (8, 8, 1), (0, 0, 0, 0, 8, 8, 1, 1), 1.0, 1.0, 1.0, 0, 64)
    assert (r["K"], r["Kmsbs"], r["Z"]) == (3, 61, 7)
    assert np.all(np.diff(r["L"].astype(np.int64)) >= 0)
    assert r["S"][0] == 0xFFFF


def test_slope2log_closed_form():
    for lam in (1.0, 2.0 ** -10, 3.5, 1e5, 2.0 ** 30):
        s = ol.lib().oc_slope2log(lam)
        expect = 256 * (np.log2(lam) - 32) + 65536
        # the mantissa is truncated to 8 bits before the log table: one bucket
        # spans 256*log2(1+1/256) = 1.44 slope units, always rounding down
        assert -1.5 <= int(s) - expect <= 0.5, (lam, s, expect)
    assert ol.lib().oc_slope2log(2.0 ** 40) == 0xFFFF
    assert ol.lib().oc_slope2log(2.0 ** -250) == 0xFFFF  # exponent underflow wraps (tier1.c:2114-2121)


def test_gain_lut_accumulates():
    one = np.zeros(30)
    two = np.zeros(30)
    ol.lib().oc_gain_lut(1, ol.ptr(one, ol.C.c_double))
    ol.lib().oc_gain_lut(2, ol.ptr(two, ol.C.c_double))
    one = one.reshape(3, 10)
    two = two.reshape(3, 10)
    # 9/7 level-1 synthesis energies are the squared norms of the synthesis filters
    h = np.array([-0.091271763114250, -0.057543526228500, 0.591271763114250, 1.115087052457000], np.float32)
    g = np.array([0.053497514821620, 0.033728236885750, -0.156446533057980, -0.533728236885750,
                  1.205898036472721], np.float32)
    eh = float((h.astype(np.float64) ** 2).sum() * 2 - float(h[-1]) ** 2)
    eg = float((g.astype(np.float64) ** 2).sum() * 2 - float(g[-1]) ** 2)
    assert one[0, 1] == pytest.approx(eh, rel=1e-12)
    assert one[0, 6] == pytest.approx(eg, rel=1e-12)
    assert one[0, 0] == 1.0 and one[0, 5] == 1.0 and two[0, 0] == 1.0
    np.testing.assert_allclose(two[:, [1, 2, 3, 4, 6, 7, 8, 9]], 2 * one[:, [1, 2, 3, 4, 6, 7, 8, 9]], rtol=1e-14)
