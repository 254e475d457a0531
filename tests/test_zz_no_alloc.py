"""bwc_amd_stage_dwt_noalloc, the allocation-free fill + forward DWT, against
bwc_amd_stage_dwt (the same stage through the allocating entry point):
bit-identical coefficients and {min, max, alpha, beta} for every (tile,
parameter) of two codecs called one after the other, the second on a
smaller field (float32, two variables, tiles), so the second call runs on
scratch the first one left behind. A parameter above the scratch limit is
refused with an error before any GPU work."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle_lib as ol
from bigwhoop_amd import bwc
from bigwhoop_amd.fields import grf

pytestmark = pytest.mark.gpu

CASES = [
    dict(shape=(96, 80, 64), npar=1, prec=8, kw=dict(rate="2", decomp=(4, 4, 3, 0), cb=(5, 5, 3, 0))),
    dict(shape=(40, 36, 28), npar=2, prec=4, kw=dict(rate="1", decomp=(3, 3, 2, 0), cb=(4, 4, 3, 0),
                                                     tiles=(32, 24, 16, 0))),
]


def _field(case, seed):
    x = np.concatenate([grf(case["shape"], slope=-5 / 3, seed=seed + p).reshape(-1) for p in range(case["npar"])])
    return x.astype(np.float32 if case["prec"] == 4 else np.float64)


def _noalloc(codec, x_dev, tile, param):
    n = int(np.prod(codec.param_dims(tile)))
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    norm = torch.empty(4, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rc = bwc.lib().bwc_amd_stage_dwt_noalloc(codec.h, tile, param, C.c_void_p(x_dev.data_ptr()),
                                            C.c_void_p(out.data_ptr()), C.c_void_p(norm.data_ptr()),
                                            C.c_void_p(stream))
    assert rc == 0, bwc.last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy(), norm.cpu().numpy()


def test_noalloc_dwt_matches_stage_dwt():
    for k, case in enumerate(CASES):
        x = _field(case, 100 + 10 * k)
        x_dev = torch.from_numpy(x).to("cuda")
        c = bwc.Codec(case["shape"], npar=case["npar"], prec=case["prec"], **case["kw"])
        o = ol.Oracle(case["shape"], npar=case["npar"], prec=case["prec"], **case["kw"])
        ntiles = o.ntiles
        o.close()
        try:
            for t in range(ntiles):
                for p in range(case["npar"]):
                    print("noalloc", case["shape"], "tile", t, "param", p, flush=True)
                    ref, ref_mm = c.stage_dwt(x, t, p)
                    got, got_mm = _noalloc(c, x_dev, t, p)
                    assert got.tobytes() == ref.tobytes(), (k, t, p)
                    assert got_mm.tobytes() == ref_mm.tobytes(), (k, t, p)
        finally:
            c.close()


def test_noalloc_refuses_a_parameter_above_the_limit():
    shape = (256, 256, 129)   # 8,454,144 values > 8,388,608
    c = bwc.Codec(shape, rate="2", decomp=(2, 2, 2, 0))
    try:
        x_dev = torch.zeros(int(np.prod(shape)), dtype=torch.float64, device="cuda")
        out = torch.empty_like(x_dev)
        norm = torch.empty(4, dtype=torch.float64, device="cuda")
        rc = bwc.lib().bwc_amd_stage_dwt_noalloc(c.h, 0, 0, C.c_void_p(x_dev.data_ptr()),
                                                C.c_void_p(out.data_ptr()), C.c_void_p(norm.data_ptr()), None)
        assert rc == 1
        assert "allocation-free limit" in bwc.last_error()
    finally:
        c.close()
