"""Concurrent compression of many fields (bwc_amd_compress_batch): every
codestream equals the one bwc_compress writes for that codec alone, for
fields of 1 to 4 axes, float32 and float64, plain and with learned pass
skipping, with several host threads and twice in a row."""
import numpy as np
import pytest

from bigwhoop_amd import bwc, skipnet
from bigwhoop_amd.fields import grf

pytestmark = pytest.mark.gpu

FIELDS = [
    dict(shape=(70000,), dtype=np.float64, rate="2", decomp=(4, 0, 0, 0), cb=(6, 0, 0, 0)),
    dict(shape=(300, 170), dtype=np.float32, rate="1.25", decomp=(4, 4, 0, 0), cb=(5, 5, 0, 0)),
    dict(shape=(64, 48, 40), dtype=np.float64, rate="3", decomp=(3, 3, 3, 0), cb=(5, 5, 3, 0)),
    dict(shape=(33, 40, 20, 12), dtype=np.float32, rate="0.75", decomp=(2, 2, 2, 1), cb=(4, 4, 3, 2)),
    dict(shape=(96, 80, 64), dtype=np.float32, rate="6", decomp=(4, 4, 4, 0), cb=(5, 5, 3, 0)),
    dict(shape=(5000,), dtype=np.float32, rate="4", decomp=(3, 0, 0, 0), cb=(10, 0, 0, 0)),
]


def make(skip):
    out = []
    for i, f in enumerate(FIELDS):
        x = grf(f["shape"], slope=-1.5 - 0.2 * i, seed=20 + i).astype(f["dtype"])
        c = bwc.Codec(f["shape"], prec=x.itemsize, rate=f["rate"], decomp=f["decomp"], cb=f["cb"])
        if skip:
            c.set_skip_model(skipnet.default_weights())
            c.set_pass_skip(2, 3)
        out.append((c, x))
    return out


@pytest.mark.parametrize("skip", [False, True])
def test_batch_matches_single(skip):
    items = make(skip)
    try:
        single = [c.compress(x) for c, x in items]
        for rep in range(2):
            print("batch", "skip" if skip else "plain", "round", rep, flush=True)
            sizes = bwc.compress_batch([c for c, _ in items], [x.ctypes.data for _, x in items], workers=4)
            for (c, _), n, ref in zip(items, sizes, single):
                assert c.out[:n].tobytes() == ref
    finally:
        for c, _ in items:
            c.close()
