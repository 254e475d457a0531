"""Pass skipping on the GPU: tier-1 codes each codeblock only to a pass limit,
rate control proves its thresholds and cuts equal those of complete codings
(coding flagged codeblocks further until they do), and the codestream is
byte-identical to the plain encoder's and to the reference's golden
codestreams."""
import sys

import numpy as np
import pytest

from bigwhoop_amd import bwc
from bigwhoop_amd.fields import grf, modes_field
from golden_cases import cases, codec_kw, make_input, sha

pytestmark = pytest.mark.gpu
CASES = cases()


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if bwc.lib().bwc_amd_device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need the MI355X")


@pytest.mark.parametrize("limit", [1, 5, 12])
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_fixed_limit_matches_reference(case, limit):
    _say("skip fixed", limit, case["name"])
    data = make_input(case)
    c = bwc.Codec(case["gen"]["shape"], npar=case["npar"], prec=case["prec"], rate=case["rate"], **codec_kw(case))
    try:
        c.set_pass_skip(1, limit)
        cs = c.compress(data)
        st = c.skip_stats()
        assert len(cs) == case["codestream_size"]
        assert sha(cs) == case["codestream_sha256"], st
    finally:
        c.close()


FIELDS = [
    dict(name="grf96_r2", shape=(96, 96, 96), rate="2", gen="grf"),
    dict(name="grf96_r1248", shape=(96, 96, 96), rate="1,2,4,8", gen="grf"),
    dict(name="grf96_r05", shape=(96, 96, 96), rate="0.5", gen="grf"),
    dict(name="modes64_r4", shape=(64, 64, 64), rate="4", gen="modes"),
    dict(name="grf128_steep_r1", shape=(128, 128, 64), rate="1", gen="grf3"),
]


def _field(spec):
    if spec["gen"] == "modes":
        return modes_field(spec["shape"], seed=11)
    if spec["gen"] == "grf3":
        return grf(spec["shape"], slope=-3.0, seed=12)
    return grf(spec["shape"], seed=13)


@pytest.mark.parametrize("limit", [3, 9, 24])
@pytest.mark.parametrize("spec", FIELDS, ids=[f["name"] for f in FIELDS])
def test_fixed_limit_matches_plain(spec, limit):
    _say("skip vs plain", limit, spec["name"])
    x = _field(spec)
    kw = dict(rate=spec["rate"], decomp=(4, 4, 4, 0), cb=(5, 5, 3, 0))
    plain = bwc.Codec(spec["shape"], **kw)
    skip = bwc.Codec(spec["shape"], **kw)
    try:
        ref = plain.compress(x)
        skip.set_pass_skip(1, limit)
        got = skip.compress(x)
        st = skip.skip_stats()
        assert got == ref, st
        assert np.array_equal(skip.cuts(), plain.cuts())
        assert st["passes_complete"] > 0
    finally:
        plain.close()
        skip.close()


@pytest.mark.parametrize("margin", [0, 3, 6])
@pytest.mark.parametrize("spec", FIELDS, ids=[f["name"] for f in FIELDS])
def test_learned_limit_matches_plain(spec, margin):
    from bigwhoop_amd import skipnet
    _say("skip learned", margin, spec["name"])
    x = _field(spec)
    kw = dict(rate=spec["rate"], decomp=(4, 4, 4, 0), cb=(5, 5, 3, 0))
    plain = bwc.Codec(spec["shape"], **kw)
    skip = bwc.Codec(spec["shape"], **kw)
    try:
        ref = plain.compress(x)
        skip.set_skip_model(skipnet.default_weights())
        skip.set_pass_skip(2, margin)
        got = skip.compress(x)
        st = skip.skip_stats()
        assert got == ref, st
        assert st["passes_coded"] < st["passes_complete"], st
    finally:
        plain.close()
        skip.close()


@pytest.mark.parametrize("model", ["headline", "wide"])
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_learned_limit_matches_reference(case, model):
    from bigwhoop_amd import skipnet
    _say("skip learned ref", model, case["name"])
    data = make_input(case)
    c = bwc.Codec(case["gen"]["shape"], npar=case["npar"], prec=case["prec"], rate=case["rate"], **codec_kw(case))
    try:
        c.set_skip_model(skipnet.default_weights() if model == "headline" else skipnet.wide_weights())
        c.set_pass_skip(2, 2)
        cs = c.compress(data)
        assert sha(cs) == case["codestream_sha256"], c.skip_stats()
    finally:
        c.close()
