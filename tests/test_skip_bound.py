"""The two facts the pass-skip proof in rate control rests on, pinned on the
CPU against complete codings by the oracle (no GPU).

A codeblock coded to a prefix of Z passes is terminated early; rate control
then needs, per codeblock:

* the first pass U whose truncation length may differ from a complete
  coding's. The MQ kernel takes U as the first pass whose truncation-length
  search reads a byte at or past the bytes emitted before termination
  (t1v2.inc finish_block). Every pass before U must have the complete
  coding's length.
* a bound on the distortion reduction the uncoded passes Z, Z+1, ... can
  add. The MQ kernel computes it from the bit-length histogram of the
  quantised magnitudes: a value with nb bits becomes significant in plane
  ps = K - nb (planes counted from the top, pass z codes plane (z + 2) / 3)
  and adds at most 256000 * 4^-ps from there on; a value significant before
  the first uncoded plane p0 adds at most 65536 * 4^-p0, both in units of
  the pass distortion sums times the codeblock's scale (the reference's
  distortion tables, tier1.c:139-184). The sum of the positive pass
  distortions of the complete coding from pass Z on must not exceed it.

Both are checked for every prefix length of codeblocks from 1-D, 2-D and
3-D fields, with and without error resilience; a bound that leaves out the
significance term (what a dropped term would give) must fail somewhere.
"""
import numpy as np
import pytest

import oracle_lib as ol
from bigwhoop_amd.fields import grf

FIELDS = [
    dict(shape=(48, 40, 24), kw=dict(rate="2", decomp=(3, 3, 2, 0), cb=(4, 4, 3, 0)), slope=-5 / 3),
    dict(shape=(96, 80), kw=dict(rate="2", decomp=(3, 3, 0, 0), cb=(5, 4, 0, 0)), slope=-2.5),
    dict(shape=(4099,), kw=dict(rate="3", decomp=(4, 0, 0, 0), cb=(8, 0, 0, 0)), slope=-1.5),
    dict(shape=(40, 36, 28), kw=dict(rate="1", decomp=(2, 2, 2, 0), cb=(4, 4, 3, 0), erres=True), slope=-2.0),
]


def _plane(z):
    return (z + 2) // 3


def bound_scaled(hist, K, Z):
    """The kernel's bound on the uncoded passes' distortion, times 4^K (an
    exact integer)."""
    p0 = _plane(Z)
    acc = 0
    for nb in range(1, K + 1):
        ps = K - nb
        if hist[nb]:
            acc += int(hist[nb]) * (256000 * 4 ** (K - ps) if ps >= p0 else 65536 * 4 ** (K - p0))
    return acc


def weak_bound_scaled(hist, K, Z):
    """The same without the significance term (refinement only)."""
    p0 = _plane(Z)
    acc = 0
    for nb in range(1, K + 1):
        ps = K - nb
        if hist[nb]:
            acc += int(hist[nb]) * 65536 * 4 ** (K - max(ps, p0))
    return acc


def actual_scaled(dm, K, Z):
    return sum(max(int(dm[z]), 0) * 4 ** (K - _plane(z)) for z in range(Z, len(dm)))


def _blocks(f):
    x = grf(f["shape"], slope=f["slope"], seed=11)
    o = ol.Oracle(f["shape"], **f["kw"])
    try:
        work, _ = o.fill_normalize(x, 0, 0)
        o.dwt_forward(work, 0, 0)
        dims = o.param_dims(0)
        box, sb, meta = o.geometry(0, 0)
    finally:
        o.close()
    return work, dims, box, sb, meta


@pytest.mark.parametrize("f", FIELDS, ids=["x".join(map(str, f["shape"])) + ("_erres" if f["kw"].get("erres") else "")
                                          for f in FIELDS])
def test_prefix_lengths_and_uncoded_bound(f):
    work, dims, box, sb, meta = _blocks(f)
    erres = bool(f["kw"].get("erres"))
    rng = np.random.default_rng(5)
    idx = [i for i in range(len(box))]
    rng.shuffle(idx)
    checked = weak_fail = prefixes = 0
    for i in idx[:10]:
        args = (work, dims, box[i], sb[i][2], sb[i][1], sb[i][0], int(meta[i][0]), int(meta[i][1]))
        full = ol.t1_probe_block(*args, zlim=0, erres=erres)
        K, Zf = int(full["K"]), int(full["Z"])
        if Zf < 4:
            continue
        checked += 1
        for Z in range(1, Zf):
            pre = ol.t1_probe_block(*args, zlim=Z, erres=erres)
            prefixes += 1
            assert int(pre["K"]) == K and int(pre["Z"]) == Z
            np.testing.assert_array_equal(pre["dm"], full["dm"][:Z])
            np.testing.assert_array_equal(pre["hist"], full["hist"])
            # U: the first pass whose length search read past the final bytes
            U = next((z for z in range(Z) if pre["reach"][z] > pre["nfinal"]), Z)
            np.testing.assert_array_equal(pre["L"][:U], full["L"][:U], err_msg=f"cb {i} Z {Z} U {U}")
            # the uncoded passes' distortion bound
            act = actual_scaled(full["dm"], K, Z)
            assert act <= bound_scaled(full["hist"], K, Z), (i, Z)
            weak_fail += act > weak_bound_scaled(full["hist"], K, Z)
    assert checked >= 3 and prefixes >= 30
    # the significance term is needed: without it the bound fails somewhere
    assert weak_fail > 0
