"""Autotuner for the kernels' launch-shape knobs (include/bwc_amd.h).

    python -m bigwhoop_amd.autotune [--size 512] [--out PATH]

On the GPU it is run on, it times every candidate of

* the strided DWT passes' shape: pair iterations unrolled per load batch
  (dwt_unroll), thread target 2^dwt_seg_log2 per pass, minimum pair
  iterations per line segment (dwt_seg_min), by the DWT stage time;
* the codeblock-to-wavefront mapping of the MQ coder: codeblocks per
  wavefront for codings to a pass limit (mq_lanes_limited, learned encoder)
  and for complete codings (mq_lanes_full, plain encoder), by the tier-1
  stage time;

on an n^3 float64 k^-5/3 Gaussian random field with the headline settings
(5 levels, codeblocks 32x32x8, 2 bits per value), stage times measured with
HIP events on the library's stream (median of the repeats). Every candidate
must give the same codestream. The winners go to a tuning file ("name value"
lines), by default bigwhoop_amd/data/tuning_gfx950.txt, which the library
reads when it loads (BWC_AMD_TUNING names another file).
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import statistics
import sys
import time
from pathlib import Path

DEFAULT_OUT = Path(__file__).resolve().parent / "data" / "tuning_gfx950.txt"
SETTINGS = dict(rate="2", decomp=(5, 5, 5, 0), cb=(5, 5, 3, 0))
DWT_CANDIDATES = dict(dwt_unroll=(2, 4, 8), dwt_seg_log2=(15, 16, 17, 18, 19), dwt_seg_min=(4, 8, 16))
MQ_LIMITED = (2, 3, 4, 5, 6, 8)
MQ_FULL = (8, 16, 32)
STAGE_DWT, STAGE_T1 = 2, 4   # indices into bwc_amd_last_timings


def _stage_ms(codec, ptr, stage, reps, ref):
    """Median stage time over `reps` calls; every codestream must equal ref."""
    ts = []
    for _ in range(reps):
        n = codec.compress_into(ptr)
        if ref is not None and codec.out[:n].tobytes() != ref:
            raise SystemExit("a tuning candidate changed the codestream")
        ts.append(codec.timings()[0][stage])
    return statistics.median(ts)


def autotune(n=512, out=DEFAULT_OUT, reps=3, log=print):
    import torch
    from bigwhoop_amd import bwc, skipnet
    from bigwhoop_amd.fields import grf_device

    # stage times of direct launches: with pipeline graphs the DWT runs inside
    # tier-1's graph and has no stage time of its own
    os.environ["BWC_AMD_GRAPH"] = "0"

    field = grf_device(n)
    ptr = field.data_ptr()
    learned = bwc.Codec((n, n, n), **SETTINGS)
    learned.set_skip_model(skipnet.default_weights())
    learned.set_pass_skip(2, 6)
    size = learned.compress_into(ptr)
    ref = learned.out[:size].tobytes()
    result = {"device": torch.cuda.get_device_name(0), "workload": f"{n}^3 float64 k^-5/3 GRF, 2 bpv",
              "dwt": [], "mq_limited": [], "mq_full": []}

    # DWT launch shape
    best = None
    for u, s, m in itertools.product(*DWT_CANDIDATES.values()):
        bwc.set_tuning("dwt_unroll", u)
        bwc.set_tuning("dwt_seg_log2", s)
        bwc.set_tuning("dwt_seg_min", m)
        learned.compress_into(ptr)
        ms = _stage_ms(learned, ptr, STAGE_DWT, reps, ref)
        result["dwt"].append([u, s, m, round(ms, 4)])
        if best is None or ms < best[0]:
            best = (ms, u, s, m)
    _, u, s, m = best
    for k, v in zip(DWT_CANDIDATES, (u, s, m)):
        bwc.set_tuning(k, v)
    log(f"autotune: dwt unroll {u} seg_log2 {s} seg_min {m}: {best[0]:.3f} ms")

    # MQ lanes, codings to a pass limit
    best = None
    for lanes in MQ_LIMITED:
        bwc.set_tuning("mq_lanes_limited", lanes)
        learned.compress_into(ptr)
        ms = _stage_ms(learned, ptr, STAGE_T1, reps, ref)
        result["mq_limited"].append([lanes, round(ms, 3)])
        if best is None or ms < best[0]:
            best = (ms, lanes)
    bwc.set_tuning("mq_lanes_limited", best[1])
    log(f"autotune: mq lanes (pass limits) {best[1]}: tier-1 {best[0]:.2f} ms")
    learned.close()

    # MQ lanes, complete codings
    plain = bwc.Codec((n, n, n), **SETTINGS)
    best = None
    for lanes in MQ_FULL:
        bwc.set_tuning("mq_lanes_full", lanes)
        ms = _stage_ms(plain, ptr, STAGE_T1, max(1, reps - 1), ref)
        result["mq_full"].append([lanes, round(ms, 3)])
        if best is None or ms < best[0]:
            best = (ms, lanes)
    bwc.set_tuning("mq_lanes_full", best[1])
    log(f"autotune: mq lanes (complete codings) {best[1]}: tier-1 {best[0]:.2f} ms")
    plain.close()

    chosen = bwc.get_tuning()
    result["chosen"] = chosen
    lines = ["# kernel launch-shape knobs, written by python -m bigwhoop_amd.autotune",
             f"# device: {result['device']}; workload: {result['workload']}; {time.strftime('%Y-%m-%d')}",
             "# DWT stage ms per (unroll, seg_log2, seg_min), best five:"]
    for row in sorted(result["dwt"], key=lambda r: r[3])[:5]:
        lines.append(f"#   {row[0]} {row[1]} {row[2]}: {row[3]}")
    lines.append("# tier-1 stage ms per MQ lanes, pass limits: " +
                 ", ".join(f"{a}: {b}" for a, b in result["mq_limited"]))
    lines.append("# tier-1 stage ms per MQ lanes, complete codings: " +
                 ", ".join(f"{a}: {b}" for a, b in result["mq_full"]))
    lines += [f"{k} {v}" for k, v in chosen.items()]
    out = Path(out)
    out.parent.mkdir(parents=True, exist_ok=True)
    out.write_text("\n".join(lines) + "\n")
    log(f"autotune: wrote {out}")
    return result


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=str(DEFAULT_OUT))
    a = ap.parse_args(argv)
    res = autotune(a.size, a.out, a.reps, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
