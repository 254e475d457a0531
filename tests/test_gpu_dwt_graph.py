"""With BWC_AMD_GRAPH=1 the forward DWT of a repeated compression is
replayed from a HIP graph (recorded on the second call with the same
geometry, buffers and launch knobs): every replay gives the codestream of
direct launches and of the oracle, with another codec's direct work and
synchronisations between the replays, across tiles and variables, and after
a launch knob changes."""
import numpy as np
import pytest

import oracle_lib as ol
from bigwhoop_amd import bwc
from bigwhoop_amd.fields import grf

pytestmark = pytest.mark.gpu

CASES = [
    dict(shape=(96, 80, 40), npar=1, kw=dict(rate="2", decomp=(3, 3, 2, 0), cb=(5, 5, 3, 0))),
    dict(shape=(48, 40, 24), npar=2, kw=dict(rate="1,4", decomp=(3, 3, 2, 0), cb=(4, 4, 3, 0),
                                             tiles=(32, 24, 16, 0))),
    dict(shape=(30000,), npar=1, kw=dict(rate="3", decomp=(4, 0, 0, 0), cb=(8, 0, 0, 0))),
]


def _field(shape, npar, seed):
    return np.concatenate([grf(shape, slope=-1.8, seed=seed + p).ravel() for p in range(npar)])


@pytest.mark.parametrize("case", CASES, ids=["x".join(map(str, c["shape"])) for c in CASES])
def test_graph_replays_match_direct_launches(case, monkeypatch):
    shape, npar, kw = case["shape"], case["npar"], case["kw"]
    print("dwt graph", shape, npar, kw, flush=True)
    x = _field(shape, npar, 5)
    ref = ol.oracle_compress(x, shape, npar=npar, **kw)
    monkeypatch.setenv("BWC_AMD_GRAPH", "0")
    c = bwc.Codec(shape, npar=npar, **kw)
    try:
        assert c.compress(x) == ref
    finally:
        c.close()
    other_shape = (40, 36, 20)
    y = grf(other_shape, slope=-2.0, seed=9)
    o = bwc.Codec(other_shape, rate="2", decomp=(2, 2, 2, 0))
    o_ref = o.compress(y)
    c = bwc.Codec(shape, npar=npar, **kw)
    unroll = bwc.get_tuning()["dwt_unroll"]
    try:
        for k in range(5):
            if k == 3:   # a new launch shape: a new graph, the same codestream
                bwc.set_tuning("dwt_unroll", 2 if unroll != 2 else 4)
            monkeypatch.setenv("BWC_AMD_GRAPH", "1")
            assert c.compress(x) == ref, k
            monkeypatch.setenv("BWC_AMD_GRAPH", "0")
            assert o.compress(y) == o_ref, k
    finally:
        bwc.set_tuning("dwt_unroll", unroll)
        c.close()
        o.close()
