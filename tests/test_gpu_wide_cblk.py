"""Codeblocks wider than a wavefront (64 stripe columns) take the fast tier-1
path in 64-column chunks, coded left to right within each stripe row: the
per-codeblock records are bit-identical to the oracle's and the codestreams
byte-identical, plain, with fixed pass limits and with learned ones."""
import numpy as np
import pytest

import oracle_lib as ol
from bigwhoop_amd import bwc, skipnet
from bigwhoop_amd.fields import grf

pytestmark = pytest.mark.gpu

CASES = [
    # 1-D codeblocks of 2^10 values, as the uneven batch uses
    dict(shape=(10000,), kw=dict(rate="2", decomp=(3, 0, 0, 0), cb=(10, 0, 0, 0))),
    dict(shape=(4099,), kw=dict(rate="1,4", decomp=(2, 0, 0, 0), cb=(10, 0, 0, 0))),
    # one stripe of 256 columns; two stripes of 128 columns with a ragged edge
    dict(shape=(700, 90), kw=dict(rate="3", decomp=(2, 2, 0, 0), cb=(8, 2, 0, 0))),
    dict(shape=(300, 50), kw=dict(rate="1.5", decomp=(2, 2, 0, 0), cb=(7, 3, 0, 0))),
    # slices of 128-column rows; 4-D
    dict(shape=(200, 20, 6), kw=dict(rate="2", decomp=(2, 1, 1, 0), cb=(7, 2, 1, 0))),
    dict(shape=(130, 9, 5, 4), kw=dict(rate="2", decomp=(1, 1, 1, 1), cb=(7, 2, 1, 1))),
]
IDS = ["x".join(map(str, c["shape"])) + "_cb" + "".join(map(str, c["kw"]["cb"])) for c in CASES]


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_wide_codeblock_records_match_oracle(case, monkeypatch):
    monkeypatch.setenv("BWC_AMD_T1", "0")
    shape, kw = case["shape"], case["kw"]
    print("wide records", shape, kw, flush=True)
    x = grf(shape, slope=-1.8, seed=11)
    c = bwc.Codec(shape, **kw)
    o = ol.Oracle(shape, **kw)
    try:
        w, _ = o.fill_normalize(x, 0, 0)
        o.dwt_forward(w, 0, 0)
        g = c.stage_t1(w.copy(), 0, 0)
        o.t1_encode(w, 0, 0)
        for i in range(o.ncblk(0, 0)):
            r = o.cblk_record(i, 0, 0)
            assert (g["K"][i], g["Kmsbs"][i], g["Z"][i]) == (r["K"], r["Kmsbs"], r["Z"]), i
            Z = r["Z"]
            np.testing.assert_array_equal(g["L"][i, :Z], r["L"].astype(np.uint32))
            np.testing.assert_array_equal(g["S"][i, : Z + 1], r["S"])
            assert g["data"][int(g["off"][i]): int(g["off"][i + 1])].tobytes() == r["data"], i
    finally:
        c.close()
        o.close()


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("skip", ["plain", "fixed4", "learned0", "learned3"])
def test_wide_codeblock_codestreams_match_oracle(case, skip):
    shape, kw = case["shape"], case["kw"]
    print("wide codestream", shape, kw, skip, flush=True)
    x = grf(shape, slope=-1.8, seed=12)
    ref = ol.oracle_compress(x, shape, **kw)
    c = bwc.Codec(shape, **kw)
    try:
        if skip == "fixed4":
            c.set_pass_skip(1, 4)
        elif skip.startswith("learned"):
            c.set_skip_model(skipnet.default_weights())
            c.set_pass_skip(2, int(skip[-1]))
        got = c.compress(x)
        assert got == ref
        assert c.compress(x) == ref
    finally:
        c.close()
