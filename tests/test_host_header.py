"""bwc_open_header on the host: the golden codestreams' main headers give
back the settings they were written with (codestream.c:601-1053,
libbwc.c:2840-2935)."""
import numpy as np
import pytest

from golden_cases import GOLDEN, cases
from bigwhoop_amd import bwc

CASES = [c for c in cases() if (GOLDEN / f"{c['name']}.bwc").exists()]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_open_header(case):
    info = bwc.header_info((GOLDEN / f"{case['name']}.bwc").read_bytes())
    shape = tuple(case["gen"]["shape"]) + (1,) * (4 - len(case["gen"]["shape"]))
    assert info["shape"] == shape
    assert info["npar"] == case["npar"]
    assert info["data_prec"] == case["prec"]
    assert info["codec_prec"] == 8
    want = sorted({float(np.float32(r)) for r in case["rate"].split(",")})
    assert info["rates"] == want


def test_open_header_rejects_garbage():
    with pytest.raises(RuntimeError):
        bwc.header_info(bytes(64))
