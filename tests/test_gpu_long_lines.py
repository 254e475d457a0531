"""X lines longer than the LDS line buffer of the X kernel (about 20K
samples) stream through the strided-axis kernel, segment by segment: 1-D and
flat 2-D fields give the oracle's codestream byte for byte, plain and with
learned pass skipping."""
import numpy as np
import pytest

import oracle_lib as ol
from bigwhoop_amd import bwc, skipnet
from bigwhoop_amd.fields import grf

pytestmark = pytest.mark.gpu

CASES = [
    dict(shape=(100000,), kw=dict(rate="2", decomp=(5, 0, 0, 0), cb=(6, 0, 0, 0))),
    dict(shape=(65536 + 37,), kw=dict(rate="4", decomp=(4, 0, 0, 0), cb=(6, 0, 0, 0))),
    dict(shape=(30011, 12), kw=dict(rate="1.5", decomp=(3, 1, 0, 0), cb=(5, 2, 0, 0))),
]


@pytest.mark.parametrize("case", CASES, ids=["x".join(map(str, c["shape"])) for c in CASES])
@pytest.mark.parametrize("skip", [False, True])
def test_long_lines_match_oracle(case, skip):
    shape, kw = case["shape"], case["kw"]
    x = grf(shape, slope=-1.7, seed=3)
    print("long lines", shape, kw, "skip" if skip else "plain", flush=True)
    ref = ol.oracle_compress(x, shape, **kw)
    c = bwc.Codec(shape, **kw)
    try:
        if skip:
            c.set_skip_model(skipnet.default_weights())
            c.set_pass_skip(2, 3)
        got = c.compress(x)
    finally:
        c.close()
    assert got == ref
