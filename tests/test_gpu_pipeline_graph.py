"""Pipeline graphs (on by default): the forward DWT with tier-1 of each
(tile, parameter) and the device rate control of each tile are recorded as
HIP graphs the second time a codec repeats them and replayed after that.
Every replay gives the oracle's codestream, with other codecs' work (graphs
on and off) and synchronisations between the replays, across tiles,
variables, layers, the learned pass skip with re-coding rounds and a 4-D
field; the counters show the graphs were recorded and replayed."""
import sys

import numpy as np
import pytest

import oracle_lib as ol
from bigwhoop_amd import bwc, skipnet
from bigwhoop_amd.fields import grf

pytestmark = pytest.mark.gpu

CASES = [
    dict(name="grf3d_r2", shape=(96, 80, 64), npar=1, kw=dict(rate="2", decomp=(5, 5, 5, 0), cb=(5, 5, 3, 0))),
    dict(name="tiles_2var_2layers", shape=(48, 40, 24), npar=2,
         kw=dict(rate="1,4", decomp=(3, 3, 2, 0), cb=(4, 4, 3, 0), tiles=(32, 24, 16, 0))),
    dict(name="grf4d", shape=(32, 24, 16, 8), npar=1, kw=dict(rate="1.5", decomp=(3, 3, 2, 1), cb=(4, 4, 3, 2))),
    dict(name="line1d", shape=(30000,), npar=1, kw=dict(rate="3", decomp=(4, 0, 0, 0), cb=(8, 0, 0, 0))),
]


def _field(shape, npar, seed):
    return np.concatenate([grf(shape, slope=-1.8, seed=seed + p).ravel() for p in range(npar)])


@pytest.mark.parametrize("skip", ["plain", "learned", "fixed3"])
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_pipeline_graph_replays(case, skip, monkeypatch):
    shape, npar, kw = case["shape"], case["npar"], case["kw"]
    print("pipeline graph", case["name"], skip, file=sys.stderr, flush=True)
    monkeypatch.delenv("BWC_AMD_GRAPH", raising=False)
    monkeypatch.delenv("BWC_AMD_RC", raising=False)
    x = _field(shape, npar, 5)
    ref = ol.oracle_compress(x, shape, npar=npar, **kw)
    y = grf((40, 36, 20), slope=-2.0, seed=9)
    o = bwc.Codec((40, 36, 20), rate="2", decomp=(2, 2, 2, 0))
    o_ref = ol.oracle_compress(y, (40, 36, 20), rate="2", decomp=(2, 2, 2, 0))
    c = bwc.Codec(shape, npar=npar, **kw)
    try:
        if skip == "learned":
            c.set_skip_model(skipnet.wide_weights())
            c.set_pass_skip(2, 2)
        elif skip == "fixed3":
            c.set_pass_skip(1, 3)
        for k in range(6):
            assert c.compress(x) == ref, k
            if k % 2:
                monkeypatch.setenv("BWC_AMD_GRAPH", "0")
            assert o.compress(y) == o_ref, k
            monkeypatch.delenv("BWC_AMD_GRAPH", raising=False)
        rec, rep = c.graph_counts()
        # one chain per (tile, parameter) and one rate control per tile
        assert rec >= 2, (rec, rep)
        assert rep >= 3 * rec, (rec, rep)
    finally:
        c.close()
        o.close()


def test_graphs_off_gives_the_same_bytes(monkeypatch):
    shape, kw = (64, 64, 48), dict(rate="0.5,2", decomp=(4, 4, 3, 0), cb=(5, 5, 3, 0))
    x = grf(shape, seed=12)
    outs = []
    for g in ("0", "1"):
        monkeypatch.setenv("BWC_AMD_GRAPH", g)
        c = bwc.Codec(shape, **kw)
        try:
            outs.append([c.compress(x) for _ in range(4)])
            counts = c.graph_counts()
        finally:
            c.close()
        assert (counts[1] > 0) == (g == "1"), counts
    assert outs[0] == outs[1]
    assert outs[0][0] == ol.oracle_compress(x, shape, **kw)


@pytest.mark.parametrize("skip", ["plain", "learned"])
def test_graph_replays_with_changing_fields(skip, monkeypatch):
    """One codec, graphs recorded on one field and replayed on others: each
    call's codestream equals the oracle's for the field it was given."""
    monkeypatch.delenv("BWC_AMD_GRAPH", raising=False)
    monkeypatch.delenv("BWC_AMD_RC", raising=False)
    shape, kw = (80, 72, 48), dict(rate="1,3", decomp=(4, 4, 3, 0), cb=(5, 5, 3, 0))
    xs = [grf(shape, slope=s, seed=20 + k) for k, s in enumerate((-1.2, -5.0 / 3.0, -2.6))]
    refs = [ol.oracle_compress(x, shape, **kw) for x in xs]
    c = bwc.Codec(shape, **kw)
    try:
        if skip == "learned":
            c.set_skip_model(skipnet.wide_weights())
            c.set_pass_skip(2, 3)
        for k in range(9):
            i = (k * 2) % 3 if k > 2 else 1
            print("changing fields", skip, k, i, file=sys.stderr, flush=True)
            assert c.compress(xs[i]) == refs[i], (k, i)
        assert c.graph_counts()[1] > 0
    finally:
        c.close()
