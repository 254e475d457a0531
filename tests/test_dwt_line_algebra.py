"""Index algebra of the fast forward 9/7 kernels, on the CPU.

dwt97_x_kernel (even/odd split in LDS) and dwt97_stream_kernel (register
pipeline along a line, optionally in warmed-up segments) restated in Python
with the same operand order, checked bit for bit against the oracle's
multi-level transform of 1-D tiles: line lengths down to 2, both parities of
the resolution box origin (tiles at odd offsets), odd and even lengths."""
import struct

import numpy as np
import pytest

import oracle_lib as ol
from bigwhoop_amd.fields import grf


def _f32(x):
    return struct.unpack("f", struct.pack("f", x))[0]


A, B, G, D, KH, KL = map(_f32, [-1.586134342059924, -0.052980118572961, 0.882911075530934, 0.360523644801462,
                                1.230174104914001, 0.812893066115961])


def mirror(p, n):
    if n == 1:
        return 0
    while p < 0 or p > n - 1:
        p = -p if p < 0 else 2 * (n - 1) - p
    return p


def split_line(x, par0):
    """dwt97_x_kernel: Ev[j] holds working index 2j - 8, Od[j] holds 2j - 7."""
    n = len(x)
    half = n // 2 if par0 else (n + 1) // 2
    hp = (n + 16) // 2 + 1
    Ev, Od = [0.0] * hp, [0.0] * hp
    for e in range(n + 16):
        v = x[mirror(e - 8 + par0, n)]
        if e & 1:
            Od[e >> 1] = v
        else:
            Ev[e >> 1] = v
    for m in range((n + 7) // 2):
        Od[2 + m] += A * (Ev[2 + m] + Ev[3 + m])
    for m in range((n + 5) // 2):
        Ev[3 + m] += B * (Od[2 + m] + Od[3 + m])
    for m in range((n + 3) // 2):
        Od[3 + m] = KH * (Od[3 + m] + G * (Ev[3 + m] + Ev[4 + m]))
    for m in range((n + 1) // 2):
        Ev[4 + m] = KL * (Ev[4 + m] + D * (Od[3 + m] + Od[4 + m]))
    return [Ev[o + 4] if o < half else Od[o - half + 4 - par0] for o in range(n)]


def stream_line(x, par0, nseg=1):
    """dwt97_stream_kernel with the line's output iterations in nseg segments."""
    n = len(x)
    half = n // 2 if par0 else (n + 1) // 2
    out = [None] * n
    last = [-3 + 2 * ((n + 7) // 2 - 1), -2 + 2 * ((n + 5) // 2 - 1), -1 + 2 * ((n + 3) // 2 - 1),
            2 * ((n + 1) // 2 - 1)]
    ilo, ihi = -par0, n - par0
    kend = (ihi + 5) // 2
    ksz = -(-(kend - 1) // nseg)

    def ld(i):
        return x[mirror(i + par0, n)]

    for seg in range(nseg):
        kA = 1 + seg * ksz
        kB = min(kend, kA + ksz)
        if kA >= kB:
            continue
        k0 = kA - 4
        # carried values start as garbage on purpose: only exact ones may be stored
        xm2, xm1, am3, bm4, cm5 = ld(2 * k0 - 2), ld(2 * k0 - 1), 1e300, -7.0, float("nan")
        for k in range(k0, kB):
            x0, x1 = ld(2 * k), ld(2 * k + 1)
            a1 = xm1 + A * (xm2 + x0) if -3 <= 2 * k - 1 <= last[0] else xm1
            b2 = xm2 + B * (am3 + a1) if -2 <= 2 * k - 2 <= last[1] else xm2
            c3 = KH * (am3 + G * (bm4 + b2)) if -1 <= 2 * k - 3 <= last[2] else am3
            d4 = KL * (bm4 + D * (cm5 + c3)) if 0 <= 2 * k - 4 <= last[3] else bm4
            i = 2 * k - 4
            if k >= kA and ilo <= i < ihi:
                assert out[i >> 1] is None
                out[i >> 1] = d4
            i = 2 * k - 3
            if k >= kA and ilo <= i < ihi:
                j = half + ((i + 2 * par0 - 1) >> 1)
                assert out[j] is None
                out[j] = c3
            xm2, xm1, am3, bm4, cm5 = x0, x1, a1, b2, c3
    assert all(v is not None for v in out)
    return out


CASES = [(130, None, 5), (257, None, 5), (45, 24, 4), (100, 24, 4), (77, 13, 3), (9, None, 3), (600, 40, 5),
         (5, None, 2), (3, None, 1), (2, None, 1)]


@pytest.mark.parametrize("n,tile,levels", CASES)
def test_line_kernels_match_oracle(n, tile, levels):
    kw = dict(decomp=(levels, 0, 0, 0))
    if tile:
        kw["tiles"] = (tile, 1, 1, 1)
    o = ol.Oracle((n, 1, 1), rate="2", **kw)
    data = grf((n, 1, 1), seed=3)
    try:
        for t in range(o.ntiles):
            w, _ = o.fill_normalize(data, t, 0)
            ref = w.copy()
            o.dwt_forward(ref, t, 0)
            lo = t * (tile or n)
            hi = min(n, lo + (tile or n))
            for fn in (split_line, stream_line, lambda v, p: stream_line(v, p, 3),
                       lambda v, p: stream_line(v, p, 7)):
                ww = list(w)
                for lv in range(levels):
                    a, b = -(-lo // 2 ** lv), -(-hi // 2 ** lv)
                    if b - a <= 1:
                        continue
                    ww[:b - a] = fn(ww[:b - a], a & 1)
                assert np.array_equal(np.array(ww).view(np.uint64), ref.view(np.uint64)), (n, tile, t)
    finally:
        o.close()


def test_segments_equal_whole_lines():
    rng = np.random.default_rng(5)
    for n in [2, 3, 5, 8, 17, 32, 33, 65, 128, 200]:
        for par0 in (0, 1):
            x = list(rng.uniform(-3, 3, n))
            ref = np.array(stream_line(x, par0))
            for nseg in (2, 4, 16):
                assert np.array(stream_line(x, par0, nseg)).tobytes() == ref.tobytes()
