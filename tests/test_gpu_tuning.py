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
    before = bwc.get_tuning()
    try:
        ref = encode(x, skip)
        for knobs in [dict(dwt_unroll=2, dwt_seg_log2=15, dwt_seg_min=4),
                      dict(dwt_unroll=8, dwt_seg_log2=19, dwt_seg_min=16),
                      dict(mq_lanes_full=1, mq_lanes_limited=1),
                      dict(mq_lanes_full=32, mq_lanes_limited=32)]:
            for k, v in knobs.items():
                bwc.set_tuning(k, v)
            print("knobs", knobs, flush=True)
            assert encode(x, skip) == ref
            for k, v in before.items():
                bwc.set_tuning(k, v)
    finally:
        for k, v in before.items():
            bwc.set_tuning(k, v)
