"""Loader for the committed golden fixtures (tests/golden/)."""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np

from bigwhoop_amd.fields import grf, modes_field

GOLDEN = Path(__file__).resolve().parent / "golden"


def cases():
    return json.loads((GOLDEN / "cases.json").read_text())


def make_input(case):
    gen = case["gen"]
    parts = []
    for p in range(gen.get("npar", 1)):
        if gen["kind"] == "modes":
            parts.append(modes_field(gen["shape"], seed=gen["seed"] + 1000 * p))
        else:
            parts.append(grf(gen["shape"], slope=gen["slope"], seed=gen["seed"] + 1000 * p))
    data = np.concatenate([a.reshape(-1) for a in parts])
    if case["prec"] == 4:
        data = data.astype(np.float32)
    assert hashlib.sha256(data.tobytes()).hexdigest() == case["input_sha256"], "input generator drifted"
    return data


def codec_kw(case):
    return {k: case[k] for k in ("decomp", "cb", "precincts", "qm", "tiles", "erres") if k in case}


def load(case, kind):
    return np.load(GOLDEN / f"{case['name']}.{kind}.npz")


def sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()
