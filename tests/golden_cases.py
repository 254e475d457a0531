"""Loader for the committed golden fixtures (tests/golden/)."""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np

from bigwhoop_amd.fields import grf, modes_field, turbulence_4d

GOLDEN = Path(__file__).resolve().parent / "golden"

# settings a case may carry, passed to the codec / oracle / reference driver
# under the same names
CODEC_KEYS = ("decomp", "cb", "precincts", "qm", "tiles", "erres", "kernels", "prog", "qstyle", "qstep", "com",
              "aux")


def cases(large: bool = False):
    """The golden cases; ``large`` also returns the full-size ones (the
    512^3 headline field), which only GPU tests run."""
    allc = json.loads((GOLDEN / "cases.json").read_text())
    return [c for c in allc if large or not c.get("large")]


def _one_var(v, shape, seed):
    kind = v["kind"]
    if kind == "modes":
        return modes_field(shape, seed=seed)
    if kind == "grf":
        return grf(shape, slope=v["slope"], seed=seed)
    if kind == "const":
        return np.full(tuple(shape)[::-1], float(v["value"]), np.float64)
    raise ValueError(f"unknown variable kind {kind}")


def generate(gen, prec=8):
    """The input of a case: variables [param][t][z][y][x] flattened, as
    float64 or float32 (prec 4)."""
    kind = gen["kind"]
    shape = gen["shape"]
    if kind == "turb4d":
        data = turbulence_4d(shape, nvar=gen.get("npar", 1), seed=gen["seed"]).reshape(-1)
    elif kind == "mixed":
        data = np.concatenate([_one_var(v, shape, gen["seed"] + 1000 * p).reshape(-1)
                               for p, v in enumerate(gen["vars"])])
    else:
        data = np.concatenate([_one_var(gen, shape, gen["seed"] + 1000 * p).reshape(-1)
                               for p in range(gen.get("npar", 1))])
    if prec == 4:
        data = data.astype(np.float32)
    return data


def make_input(case):
    data = generate(case["gen"], case["prec"])
    assert hashlib.sha256(data.tobytes()).hexdigest() == case["input_sha256"], "input generator drifted"
    return data


def codec_kw(case):
    return {k: case[k] for k in CODEC_KEYS if k in case}


def load(case, kind):
    return np.load(GOLDEN / f"{case['name']}.{kind}.npz")


def sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()
