"""Kernel launch-shape knobs (bwc_amd_set_tuning / bwc_amd_get_tuning): the
C API accepts exactly the documented values, the tuning file and the
environment variables set them at load time. No GPU needed."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

from bigwhoop_amd import bwc

ROOT = Path(__file__).resolve().parents[1]
DEFAULTS = {"dwt_unroll": 4, "dwt_seg_log2": 17, "dwt_seg_min": 8, "mq_lanes_full": 16, "mq_lanes_limited": 5}


def run_py(code, env):
    e = dict(os.environ)
    for k in list(e):
        if k.startswith("BWC_AMD_"):
            del e[k]
    e.update(env)
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=e, cwd=str(ROOT),
                       timeout=120)
    assert p.returncode == 0, p.stderr
    return p.stdout.strip().splitlines()[-1]


def test_set_get_and_ranges():
    before = bwc.get_tuning()
    try:
        for name, good, bad in [("dwt_unroll", [2, 4, 8], [1, 3, 16]), ("dwt_seg_log2", [8, 17, 24], [7, 25]),
                                ("dwt_seg_min", [2, 100], [0, 1]), ("mq_lanes_full", [1, 2, 32], [0, 3, 64]),
                                ("mq_lanes_limited", [3, 5, 6, 16], [7, 12, 33]), ("dwt_nt", [0, 3], [-1, 4]),
                                ("dwt_xy", [0, 1], [2]), ("dwt_xy_min_log2", [0, 24, 40], [-1, 41]),
                                ("dwt_inv_stream", [0, 1], [2])]:
            for v in good:
                bwc.set_tuning(name, v)
                assert bwc.get_tuning(name) == v
            for v in bad:
                with pytest.raises(ValueError):
                    bwc.set_tuning(name, v)
                assert bwc.get_tuning(name) == good[-1]
        with pytest.raises(ValueError):
            bwc.set_tuning("no_such_knob", 1)
        with pytest.raises(ValueError):
            bwc.get_tuning("no_such_knob")
    finally:
        for k, v in before.items():
            bwc.set_tuning(k, v)


def test_defaults_without_file():
    out = run_py("from bigwhoop_amd import bwc; print(bwc.get_tuning(), repr(bwc.tuning_source()))",
                 {"BWC_AMD_TUNING": "/nonexistent/tuning.txt"})
    assert out == f"{DEFAULTS} ''"


def test_file_then_environment(tmp_path):
    f = tmp_path / "t.txt"
    f.write_text("# comment\nmq_lanes_limited 8\ndwt_unroll 3   # not a valid value: ignored\n"
                 "dwt_seg_log2 18\nunknown_knob 5\nmq_lanes_full 32\n")
    out = run_py("from bigwhoop_amd import bwc; print(bwc.get_tuning(), bwc.tuning_source())",
                 {"BWC_AMD_TUNING": str(f), "BWC_AMD_MQ_LANES_FULL": "8"})
    want = dict(DEFAULTS, mq_lanes_limited=8, dwt_seg_log2=18, mq_lanes_full=8)
    assert out == f"{want} {f}"
