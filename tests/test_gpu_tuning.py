"""Launch-shape knobs change how the work is laid out, never the result:
the same codestream for every DWT shape and MQ lane count (plain and
learned encoders)."""
import pytest

from bigwhoop_amd import bwc, skipnet
from bigwhoop_amd.fields import grf

pytestmark = pytest.mark.gpu

SHAPE = (96, 80, 64)
KW = dict(rate="2", decomp=(4, 4, 4, 0), cb=(5, 5, 3, 0))


def encode(x, skip):
    c = bwc.Codec(SHAPE, **KW)
    try:
        if skip:
            c.set_skip_model(skipnet.default_weights())
            c.set_pass_skip(2, 3)
        return c.compress(x)
    finally:
        c.close()


@pytest.mark.parametrize("skip", [False, True])
def test_knobs_do_not_change_codestream(skip):
    x = grf(SHAPE, seed=11)
    extra = ("dwt_nt", "dwt_xy", "dwt_xy_min_log2")
    before = dict(bwc.get_tuning(), **{k: bwc.get_tuning(k) for k in extra})
    try:
        ref = encode(x, skip)
        for knobs in [dict(dwt_unroll=2, dwt_seg_log2=15, dwt_seg_min=4),
                      dict(dwt_unroll=8, dwt_seg_log2=19, dwt_seg_min=16),
                      dict(mq_lanes_full=1, mq_lanes_limited=1),
                      dict(mq_lanes_full=32, mq_lanes_limited=32),
                      dict(mq_lanes_limited=3), dict(mq_lanes_limited=6),
                      dict(dwt_nt=0, dwt_xy=0), dict(dwt_xy=1, dwt_xy_min_log2=0)]:
            for k, v in knobs.items():
                bwc.set_tuning(k, v)
            print("knobs", knobs, flush=True)
            assert encode(x, skip) == ref
            for k, v in before.items():
                bwc.set_tuning(k, v)
    finally:
        for k, v in before.items():
            bwc.set_tuning(k, v)
