"""Benchmark: compression throughput of one MI355X on the headline workload.

Workload (one step = one bwc_compress call): a 512^3 float64 Gaussian random
field with isotropic energy spectrum E(k) ~ k^-5/3 (seeded FFT synthesis,
unit variance, mean 10), 5 CDF 9/7 levels per axis, codeblocks 32x32x8, one
quality layer at 2 bits/value.

`value` is input bytes over the wall time of bwc_compress with the field in
page-locked host memory, as the C API takes it: the copy to HBM (PCIe) is
inside the timed call, and the codestream is written to page-locked host
memory. The encoder is the learned pass-skip encoder (tier-1 codes each
codeblock to the pass-limit network's prediction plus a margin; rate control
proves the cuts are those of complete codings). Beside it: the same field in
ordinary (pageable) host memory, the same field already resident in HBM (a
device pointer), the plain encoder (every pass coded) and decompression.
Every codestream must be byte-identical to the others, and at 512^3 to the
SHA-256 of the codestream the reference's own bwc_compress wrote for this
field (tests/golden), or the run fails.

Prints one JSON line. Also reports the DWT's achieved HBM bandwidth against
the 8 TB/s peak (algorithmic bytes: 16 B per value per axis pass, summed over
levels) measured with HIP events on the library's stream, and the reference
CPU path (built from source into oracle/_ref when present, else the C oracle)
timed on this host on a bounded 256^3 sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
SETTINGS = dict(rate="2", decomp=(5, 5, 5, 0), cb=(5, 5, 3, 0))
SKIP_MARGIN = 6
TRAFFIC_PROFILE = ROOT / "profiles" / "dwt_512cube_hbm_traffic.md"


def dwt_traffic():
    """HBM bytes of the DWT stage measured with rocprofv3 counters (FETCH_SIZE
    and WRITE_SIZE, each scaled by its reading for a 1 GiB device copy), as
    recorded in the committed profile summary; None when that is absent."""
    try:
        for line in TRAFFIC_PROFILE.read_text().splitlines():
            if line.startswith("dwt_stage_hbm_bytes:"):
                return int(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def make_field_host(n: int):
    """The headline field on the host: the seeded numpy synthesis the
    reference's golden codestream was written from (bigwhoop_amd.fields.grf,
    seed 1), with its committed input digest when one exists for n."""
    import hashlib

    from bigwhoop_amd.fields import grf

    x = grf((n, n, n), slope=-5.0 / 3.0, seed=1)
    golden = None
    try:
        for c in json.loads((ROOT / "tests" / "golden" / "cases.json").read_text()):
            if c.get("large") and c["gen"]["shape"] == [n, n, n] and c["rate"] == SETTINGS["rate"]:
                golden = c
    except OSError:
        pass
    if golden is not None and hashlib.sha256(x.tobytes()).hexdigest() != golden["input_sha256"]:
        print("bench: the field differs from the one the reference compressed", file=sys.stderr)
        sys.exit(1)
    return x, golden


def dwt_algorithmic_bytes(n: int, levels: int) -> int:
    """16 bytes (one f64 read + one write) per value of the current box per axis pass."""
    tot = 0
    for lv in range(levels):
        m = -(-n // (1 << lv))
        tot += 3 * 16 * m ** 3
    return tot


CPU_SNIPPET = r"""
import ctypes as C, sys, time, numpy as np
sys.path.insert(0, sys.argv[1])
from bigwhoop_amd.fields import grf
lib = C.CDLL(sys.argv[2])
n = int(sys.argv[3])
x = grf((n, n, n), seed=1)
vp, u8, u64 = C.c_void_p, C.c_uint8, C.c_uint64
lib.bwc_alloc_coder.restype = vp; lib.bwc_alloc_coder.argtypes = [u64, u64, u64, u64, u8, C.c_int]
lib.bwc_init_stream.restype = vp; lib.bwc_init_stream.argtypes = [vp, vp, C.c_int]
lib.bwc_set_decomp.argtypes = [vp, u8, u8, u8, u8]; lib.bwc_set_codeblocks.argtypes = [vp, u8, u8, u8, u8]
lib.bwc_create_compression.restype = C.c_ubyte; lib.bwc_create_compression.argtypes = [vp, vp, C.c_char_p]
lib.bwc_compress.restype = C.c_size_t; lib.bwc_compress.argtypes = [vp, vp]
out = np.zeros(x.nbytes * 2, np.uint8)
codec = lib.bwc_alloc_coder(n, n, n, 1, 1, 8)
lib.bwc_set_decomp(codec, 5, 5, 5, 0); lib.bwc_set_codeblocks(codec, 5, 5, 3, 0)
st = lib.bwc_init_stream(x.ctypes.data_as(vp), out.ctypes.data_as(vp), 0)
rate = C.create_string_buffer(b"2")
assert lib.bwc_create_compression(codec, st, rate) == 0
t = time.perf_counter(); sz = lib.bwc_compress(codec, st); dt = time.perf_counter() - t
print(dt, sz)
"""


def cpu_baseline(n: int = 256):
    ref = ROOT / "oracle" / "_ref" / "libbwc_ref.so"
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    if ref.exists():
        kind, lib = "reference", ref
        p = subprocess.run([sys.executable, "-c", CPU_SNIPPET, str(ROOT), str(lib), str(n)], capture_output=True,
                           text=True, env=env, timeout=600)
        if p.returncode != 0:
            return {"value": None, "unit": "GB/s", "cores": cores, "kind": kind, "sample": "failed: " + p.stderr[-300:]}
        dt, sz = p.stdout.split()[-2:]
        dt = float(dt)
    else:
        kind = "port"
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib as ol
        from bigwhoop_amd.fields import grf

        x = grf((n, n, n), seed=1)
        o = ol.Oracle((n, n, n), **SETTINGS)
        t = time.perf_counter()
        o.compress(x)
        dt = time.perf_counter() - t
        o.close()
    return {"value": round(n ** 3 * 8 / dt / 1e9, 5), "unit": "GB/s", "cores": cores, "kind": kind,
            "sample": f"{n}^3 float64 k^-5/3 GRF, same settings, one bwc_compress call: {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one plain device copy of the field first (counter calibration runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import hashlib

    import torch
    from bigwhoop_amd import bwc, skipnet

    n = args.size
    field_np, golden = make_field_host(n)
    print(f"bench: field made ({n}^3 float64)", file=sys.stderr, flush=True)
    # page-locked host copy (the headline input) and a device copy
    host = torch.empty(n ** 3, dtype=torch.float64, pin_memory=True)
    host.copy_(torch.from_numpy(field_np.reshape(-1)))
    field = torch.from_numpy(field_np.reshape(-1)).to("cuda")
    torch.cuda.synchronize()
    if args.copy_calibration:
        # one device copy of exactly n^3 * 8 bytes: the reference reading for
        # the FETCH_SIZE / WRITE_SIZE counters of the DWT kernels
        scratch = torch.empty_like(field)
        scratch.copy_(field)
        torch.cuda.synchronize()
        del scratch

    def run(codec, steps, src):
        """`steps` timed bwc_compress calls into the codec's own output buffer;
        wall time, the last codestream, the DWT stage ms of each call."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dwt_ms, size = [], 0
        for _ in range(steps):
            size = codec.compress_into(src)
            dwt_ms.append(codec.timings()[0][2])
        torch.cuda.synchronize()
        return time.perf_counter() - t0, codec.out[:size].tobytes(), dwt_ms

    def learned_codec():
        c = bwc.Codec((n, n, n), pinned=True, **SETTINGS)
        c.set_skip_model(skipnet.default_weights())
        c.set_pass_skip(2, SKIP_MARGIN)
        return c

    def leg(name, codec, src):
        for _ in range(args.warmup):
            codec.compress_into(src)
        dt, cs, dwt = run(codec, args.steps, src)
        print(f"bench: {name} {dt / args.steps * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
        return dt, cs, dwt

    # headline: learned pass skip, field in page-locked host memory
    codec = learned_codec()
    dt, cs, dwt_ms = leg("learned pass skip, page-locked host input", codec, host.data_ptr())
    ms_last, _ = codec.timings()
    st = codec.skip_stats()
    codec.close()
    # the same field in ordinary (pageable) host memory
    pg = learned_codec()
    dt_pg, cs_pg, _ = leg("learned pass skip, pageable host input", pg, field_np.ctypes.data)
    pg.close()
    # the same field resident in HBM (device pointer)
    dv = learned_codec()
    dt_dv, cs_dv, dwt_dv = leg("learned pass skip, device input", dv, field.data_ptr())
    ms_dev, _ = dv.timings()
    dv.close()
    # The DWT stage timed with HIP events around its launches on the
    # library's stream: the same device-input compression with pipeline
    # graphs off (with them on, the DWT is recorded into tier-1's graph and
    # has no events of its own); these calls also give the per-stage times
    os.environ["BWC_AMD_GRAPH"] = "0"
    tm = learned_codec()
    _, cs_tm, dwt_tm = leg("learned pass skip, device input, no graphs (stage timing)", tm, field.data_ptr())
    ms_stage, _ = tm.timings()
    tm.close()
    del os.environ["BWC_AMD_GRAPH"]
    # plain encoder (every pass of every codeblock coded), page-locked host input
    plain = bwc.Codec((n, n, n), pinned=True, **SETTINGS)
    dt_plain, cs_plain, dwt_plain = leg("plain encoder, page-locked host input", plain, host.data_ptr())
    plain.close()
    for name, other in (("pageable-input", cs_pg), ("device-input", cs_dv), ("plain-encoder", cs_plain),
                        ("no-graph", cs_tm)):
        if other != cs:
            print(f"bench: {name} codestream differs from the headline one", file=sys.stderr)
            sys.exit(1)
    same_as_ref = None
    if golden is not None:
        same_as_ref = (len(cs) == golden["codestream_size"]
                       and hashlib.sha256(cs).hexdigest() == golden["codestream_sha256"])
        if not same_as_ref:
            print("bench: codestream differs from the reference's", file=sys.stderr)
            sys.exit(1)
    del host
    # decompression of the headline codestream into device memory
    dec = torch.empty(n ** 3, dtype=torch.float64, device="cuda")
    for _ in range(args.warmup):
        bwc.decompress_bytes(cs, out=dec)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bwc.decompress_bytes(cs, out=dec)
    torch.cuda.synchronize()
    dt_dec = time.perf_counter() - t0
    print(f"bench: decompress {dt_dec / args.steps * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
    dec_ok = None
    if golden is not None:
        dec_ok = hashlib.sha256(dec.cpu().numpy().tobytes()).hexdigest() == golden["decode_sha256"]
    err = (dec - field).double()
    psnr_db = 20 * math.log10(float(field.max() - field.min()) / (2 * math.sqrt(float((err * err).mean()))))
    del dec, err

    bytes_in = n ** 3 * 8
    value = bytes_in * args.steps / dt / 1e9
    plain_gbs = bytes_in * args.steps / dt_plain / 1e9
    alg = dwt_algorithmic_bytes(n, 5)
    dwt_avg = sum(dwt_tm) / len(dwt_tm)
    achieved = alg / (dwt_avg / 1e3) / 1e9
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    stage_names = ["fill_minmax", "normalize", "dwt", "quantize", "tier1", "records_d2h", "rate_control",
                   "headers_gather", "assembly_d2h", "total"]
    line = {
        "metric": "compress GB/s at 2 bpv, bit-exact codestream; learned-skip vs plain GPU",
        "value": round(value, 4),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 float64 GRF k^-5/3, 5 levels, cb 32x32x8, 1 layer at 2 bpv",
                   "input": "page-locked host memory, copied to HBM inside bwc_compress (PCIe)",
                   "encoder": f"learned pass skip (margin {SKIP_MARGIN} passes)", "codestream_bytes": len(cs),
                   "codestream_identical_to_plain": True, "codestream_identical_to_reference": same_as_ref},
        "pageable_input": {"value": round(bytes_in * args.steps / dt_pg / 1e9, 4), "unit": "GB/s",
                           "ms_per_step": round(dt_pg / args.steps * 1e3, 3)},
        "device_input": {"value": round(bytes_in * args.steps / dt_dv / 1e9, 4), "unit": "GB/s",
                         "ms_per_step": round(dt_dv / args.steps * 1e3, 3),
                         "stages_ms": {k: round(v, 3) for k, v in zip(stage_names, ms_dev)},
                         "stages_ms_without_graphs": {k: round(v, 3) for k, v in zip(stage_names, ms_stage)}},
        "decompress": {"value": round(bytes_in * args.steps / dt_dec / 1e9, 4), "unit": "GB/s",
                       "ms_per_step": round(dt_dec / args.steps * 1e3, 3), "psnr_db": round(psnr_db, 3),
                       "output": "device memory", "identical_to_reference": dec_ok},
        "plain_encoder": {"value": round(plain_gbs, 4), "unit": "GB/s",
                          "ms_per_step": round(dt_plain / args.steps * 1e3, 3),
                          "speedup_learned_vs_plain": round(dt_plain / dt, 3)},
        "pass_skip": {"passes_coded": st["passes_coded"], "passes_complete": st["passes_complete"],
                      "extra_rate_control_rounds": st["extra_rounds"], "codeblocks_coded_further": st["extended"],
                      "fallbacks": st["fallbacks"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": dwt_traffic(),
                     "kernel": "DWT stage (5 levels x 3 axes in 13 launches: X and Y fused at the two finest levels, normalisation fused)", "dwt_ms": round(dwt_avg, 4),
                     "algorithmic_bytes": alg},
        "stages_ms": {k: round(v, 3) for k, v in zip(stage_names, ms_last)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
