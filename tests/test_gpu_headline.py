"""The headline workload at full size against the reference's own bytes:
the 512^3 float64 k^-5/3 Gaussian random field (seed 1), 5 DWT levels per
axis, codeblocks 32x32x8, one layer at 2 bpv and four layers at 1, 2, 4 and
8 bpv. The field is regenerated on the host from its seed and checked
against the input digest the reference was fed; it is compressed from host
memory (the copy to HBM inside bwc_compress), by the plain encoder and by
the learned pass-skip encoder, and each codestream must have the SHA-256 and
size of the reference's bwc_compress output. Decoding it must give the
SHA-256 of the reference's bwc_decompress output."""
import numpy as np
import pytest

from bigwhoop_amd import bwc, skipnet
from golden_cases import cases, codec_kw, make_input, sha

pytestmark = pytest.mark.gpu

CASES = [c for c in cases(large=True) if c.get("large")]


@pytest.fixture(scope="module")
def field():
    print("headline: generating the 512^3 field", flush=True)
    return make_input(CASES[0])


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
@pytest.mark.parametrize("learned", [False, True], ids=["plain", "learned"])
def test_headline_codestream_matches_reference(field, case, learned):
    assert case["input_sha256"] == CASES[0]["input_sha256"]
    c = bwc.Codec(case["gen"]["shape"], npar=case["npar"], prec=case["prec"], rate=case["rate"], **codec_kw(case))
    try:
        if learned:
            c.set_skip_model(skipnet.default_weights())
            c.set_pass_skip(2, 6)
        print("headline:", case["name"], "learned" if learned else "plain", flush=True)
        cs = c.compress(field)
        assert len(cs) == case["codestream_size"]
        assert sha(cs) == case["codestream_sha256"]
        if learned:
            print("headline: decoding", case["name"], flush=True)
            dec, _ = bwc.decompress_bytes(cs)
            assert dec.nbytes == case["decode_size"]
            assert sha(dec.tobytes()) == case["decode_sha256"]
            del dec
    finally:
        c.close()
