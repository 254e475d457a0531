"""The CPU oracle against the reference's golden vectors (tests/golden/,
written by tests/golden/make_golden.py from the reference built from source)."""
import hashlib

import numpy as np
import pytest

import oracle_lib as ol
from golden_cases import GOLDEN, cases, codec_kw, load, make_input, sha

CASES = cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_codestream_bytes(case):
    data = make_input(case)
    cs = ol.oracle_compress(data, case["gen"]["shape"], npar=case["npar"], prec=case["prec"], rate=case["rate"],
                            **codec_kw(case))
    assert len(cs) == case["codestream_size"]
    assert sha(cs) == case["codestream_sha256"]
    f = GOLDEN / f"{case['name']}.bwc"
    if f.exists():
        assert cs == f.read_bytes()


STAGE_CASES = [c for c in CASES if any(f.endswith(".t1.npz") for f in c["files"])]


@pytest.mark.parametrize("case", STAGE_CASES, ids=[c["name"] for c in STAGE_CASES])
def test_stages(case):
    data = make_input(case)
    o = ol.Oracle(case["gen"]["shape"], npar=case["npar"], prec=case["prec"], rate=case["rate"], **codec_kw(case))
    geo = load(case, "geo")
    t1 = load(case, "t1")
    dwt = load(case, "dwt")
    boxes, sbs, metas, coefs = [], [], [], []
    recs = []
    slopes = []
    for t in range(o.ntiles):
        for p in range(case["npar"]):
            b, s, m = o.geometry(t, p)
            boxes.append(b), sbs.append(s), metas.append(m)
            work, _ = o.fill_normalize(data, t, p)
            o.dwt_forward(work, t, p)
            coefs.append(work.copy())
            o.t1_encode(work, t, p)
            recs += [o.cblk_record(c, t, p) for c in range(o.ncblk(t, p))]
            slopes.append(o.tile_slopes(t))
    np.testing.assert_array_equal(np.concatenate(boxes), geo["box"])
    np.testing.assert_array_equal(np.concatenate(sbs), geo["sb"])
    np.testing.assert_array_equal(np.concatenate(metas), geo["meta"])
    allc = np.concatenate(coefs)
    assert allc.size == int(dwt["n"][0])
    np.testing.assert_array_equal(allc[::97], dwt["sample"])
    assert hashlib.sha256(allc.tobytes()).digest() == dwt["sha256"].tobytes()
    assert len(recs) == len(t1["K"])
    for i, r in enumerate(recs):
        assert (r["K"], r["Kmsbs"], r["Z"]) == (t1["K"][i], t1["Kmsbs"][i], t1["Z"][i]), i
        np.testing.assert_array_equal(r["S"], t1["S"][i, : r["Z"] + 1])
        np.testing.assert_array_equal(r["L"], t1["L"][i, : r["Z"]])
        assert hashlib.sha256(r["data"]).digest() == t1["data_sha256"][i].tobytes()
    np.testing.assert_array_equal(np.array(slopes, np.uint16), t1["slopes"])
    o.close()
