"""The uneven batch workload at reduced element counts, against the CPU
oracle: 16 fields drawn by the batch workload's own generator rule
(dimensionality 1-4, random aspect ratios, float32/float64, spectral slope
U(-3, -1), bitrate log-uniform 0.5-8 bpv, DWT levels
min(5, floor(log2(shortest axis)) - 2), the workload's codeblock shapes),
with element counts log-uniform in [1e4, 1e6] instead of [1e5, 1e8].
Every codestream of one concurrent bwc_amd_compress_batch call, plain and
with the learned pass limits (network and margin from skipnet.choose), must
equal the oracle's for that field, with the fields' tier-1 coded through the
batch's global codeblock queue (one generation and one MQ launch over all
fields) and with each field coded on its own (BWC_AMD_BATCH=threads)."""
import numpy as np
import pytest

import oracle_lib as ol
from bigwhoop_amd import bwc, skipnet
from bigwhoop_amd.fields import grf
from bigwhoop_amd.workloads import batch_spec

pytestmark = pytest.mark.gpu

SPEC = batch_spec(16, seed=4, nmin=1e4, nmax=1e6)


def _field(f):
    d = f["ndim"]
    x = grf(f["shape"][:d], slope=f["slope"], seed=f["seed"])
    return np.ascontiguousarray(x.astype(np.float32 if f["dtype"] == "float32" else np.float64)).reshape(-1)


@pytest.fixture(scope="module")
def fields_and_oracle():
    fields = [_field(f) for f in SPEC]
    ref = []
    for f, x in zip(SPEC, fields):
        ref.append(ol.oracle_compress(x, f["shape"], prec=x.itemsize, rate=f["rate"], decomp=f["decomp"],
                                      cb=f["cb"]))
    return fields, ref


@pytest.mark.parametrize("mode", ["queue", "threads"])
@pytest.mark.parametrize("learned", [False, True])
def test_uneven_batch_matches_oracle(fields_and_oracle, learned, mode, monkeypatch):
    fields, ref = fields_and_oracle
    if mode == "threads":
        monkeypatch.setenv("BWC_AMD_BATCH", "threads")
    else:
        monkeypatch.delenv("BWC_AMD_BATCH", raising=False)
    codecs = []
    try:
        for f, x in zip(SPEC, fields):
            c = bwc.Codec(f["shape"], prec=x.itemsize, rate=f["rate"], decomp=f["decomp"], cb=f["cb"])
            if learned:
                w, margin = skipnet.choose(f["ndim"], float(f["rate"]))
                c.set_skip_model(w)
                c.set_pass_skip(2, margin)
            codecs.append(c)
        print("uneven batch", "learned" if learned else "plain", len(codecs), "fields", flush=True)
        sizes = bwc.compress_batch(codecs, [x.ctypes.data for x in fields], workers=8)
        for i, (c, n) in enumerate(zip(codecs, sizes)):
            assert c.out[:n].tobytes() == ref[i], (i, SPEC[i]["shape"], SPEC[i]["dtype"], SPEC[i]["rate"])
            assert c.batch_queued() == (1 if mode == "queue" else 0), i
        # a second batch call on the same codecs
        sizes = bwc.compress_batch(codecs, [x.ctypes.data for x in fields], workers=8)
        for i, (c, n) in enumerate(zip(codecs, sizes)):
            assert c.out[:n].tobytes() == ref[i], i
    finally:
        for c in codecs:
            c.close()
