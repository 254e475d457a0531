"""GPU decompression through the C-ABI (bwc_create_decompression +
bwc_decompress) against the reference decoder's output and the oracle.

- every golden codestream decodes to exactly the bytes of the reference's
  bwc_decompress (SHA-256 in tests/golden/cases.json);
- codestreams this library writes (plain and with learned pass skipping,
  several layouts) decode bit for bit like the oracle's decoder;
- decoding into device memory gives the same field;
- at the headline size (512^3, 2 bpv) the decoded field matches the input to
  the PSNR a 2 bpv codestream reaches (a size-independent round trip)."""
import hashlib

import numpy as np
import pytest

import oracle_lib as ol
from golden_cases import GOLDEN, cases, codec_kw, make_input
from bigwhoop_amd import bwc
from bigwhoop_amd.fields import grf

pytestmark = pytest.mark.gpu
CASES = cases()


def golden_codestream(case):
    f = GOLDEN / f"{case['name']}.bwc"
    if f.exists():
        return f.read_bytes()
    data = make_input(case)
    return ol.oracle_compress(data, case["gen"]["shape"], npar=case["npar"], prec=case["prec"], rate=case["rate"],
                              **codec_kw(case))


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_decoder_matches_reference(case):
    cs = golden_codestream(case)
    print("decode", case["name"], len(cs), "bytes", flush=True)
    vals, info = bwc.decompress_bytes(cs)
    assert vals.nbytes == case["decode_size"]
    assert hashlib.sha256(vals.tobytes()).hexdigest() == case["decode_sha256"]


LAYOUTS = [
    dict(shape=(64, 48, 40), rate="2", decomp=(4, 4, 3, 0), cb=(5, 5, 3, 0)),
    dict(shape=(57, 33, 21), rate="0.75,3", decomp=(3, 2, 2, 0), cb=(4, 4, 4, 0)),
    dict(shape=(40, 40, 16, 6), rate="4", decomp=(3, 3, 2, 1), cb=(4, 4, 3, 1)),
    dict(shape=(48, 40, 24), rate="1.5,6", decomp=(3, 3, 2, 0), cb=(4, 4, 3, 0), tiles=(32, 24, 16, 0), erres=True),
    # odd line lengths down to 2, tile origins that put the lowpass on odd
    # positions (lines [45, 69), [23, 35) and [3, 5) of the second tile),
    # segments of the streaming inverse passes. An odd origin with an odd
    # length is left out: the reference decoder then reads one sample of its
    # working buffer that it never filled (dwt.c:169-272), so its output
    # there depends on the lines its thread decoded before.
    dict(shape=(37, 29, 9), rate="3", decomp=(3, 2, 3, 0), cb=(4, 4, 3, 0)),
    dict(shape=(21, 69, 11), rate="2", decomp=(1, 4, 2, 0), cb=(4, 4, 2, 0), tiles=(21, 45, 11, 0)),
    dict(shape=(9, 130, 67), rate="2", decomp=(2, 5, 5, 0), cb=(3, 5, 4, 0)),
]


@pytest.mark.parametrize("lay", LAYOUTS, ids=[f"{'x'.join(map(str, l['shape']))}_r{l['rate']}" for l in LAYOUTS])
@pytest.mark.parametrize("skip", [False, True])
def test_gpu_roundtrip_matches_oracle_decoder(lay, skip):
    from bigwhoop_amd import skipnet
    lay = dict(lay)
    shape = lay.pop("shape")
    x = grf(shape, slope=-2.0, seed=5)
    c = bwc.Codec(shape, **lay)
    try:
        if skip:
            c.set_skip_model(skipnet.default_weights())
            c.set_pass_skip(2, 3)
        cs = c.compress(x)
    finally:
        c.close()
    print("roundtrip", shape, lay, "skip" if skip else "plain", len(cs), flush=True)
    got, info = bwc.decompress_bytes(cs)
    ref, dims = ol.oracle_decompress(cs)
    assert got.dtype == ref.dtype and got.size == ref.size
    assert got.tobytes() == ref.tobytes()


def test_gpu_decode_into_device_memory():
    import torch
    shape = (40, 36, 28)
    x = grf(shape, slope=-1.8, seed=9)
    cs = bwc.compress_array(x, shape, rate="3", decomp=(3, 3, 3, 0), cb=(4, 4, 3, 0))
    host, _ = bwc.decompress_bytes(cs)
    dev = torch.zeros(host.size, dtype=torch.float64, device="cuda")
    dev, _ = bwc.decompress_bytes(cs, out=dev)
    torch.cuda.synchronize()
    assert dev.cpu().numpy().tobytes() == host.tobytes()


def psnr(x, y):
    mse = float(np.mean((x - y) ** 2))
    return 20 * np.log10(float(x.max() - x.min()) / (2 * np.sqrt(mse)))


def test_gpu_roundtrip_headline_size():
    import torch
    from bigwhoop_amd.fields import grf_device
    n = 512
    f = grf_device(n)
    c = bwc.Codec((n, n, n), rate="2", decomp=(5, 5, 5, 0), cb=(5, 5, 3, 0))
    try:
        cs = c.compress_ptr(f.data_ptr())
    finally:
        c.close()
    print("512^3 codestream", len(cs), flush=True)
    dev = torch.empty(n ** 3, dtype=torch.float64, device="cuda")
    dev, info = bwc.decompress_bytes(cs, out=dev)
    x = f.reshape(-1)
    mse = torch.mean((x - dev) ** 2).item()
    q = 20 * np.log10((x.max() - x.min()).item() / (2 * np.sqrt(mse)))
    print("512^3 PSNR", q, flush=True)
    assert info["shape"] == (n, n, n, 1)
    assert q > 40
