"""Benchmark: compression throughput of one MI355X on the headline workload.

Workload (one step = one bwc_compress call): a 512^3 float64 Gaussian random
field with isotropic energy spectrum E(k) ~ k^-5/3 (seeded FFT synthesis,
unit variance, mean 10), 5 CDF 9/7 levels per axis, codeblocks 32x32x8, one
quality layer at 2 bits/value. The field is resident in HBM before timing
starts (the C API reads a device pointer in place); the codestream is written
to host memory inside the timed region.

`value` is the learned pass-skip encoder (tier-1 codes each codeblock to the
pass-limit network's prediction plus a margin; rate control proves the cuts
are those of complete codings). The plain encoder (every pass coded) is timed
on the same field, and the two codestreams must be byte-identical or the run
fails.

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


def make_field_gpu(n: int, seed: int = 1, slope: float = -5.0 / 3.0):
    """Seeded FFT synthesis of the GRF directly in HBM (torch on the GPU)."""
    import torch

    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    white = torch.randn((n, n, n), generator=g, device="cuda", dtype=torch.float64)
    spec = torch.fft.rfftn(white)
    del white
    f = torch.fft.fftfreq(n, d=1.0 / n, device="cuda", dtype=torch.float64)
    fr = torch.fft.rfftfreq(n, d=1.0 / n, device="cuda", dtype=torch.float64)
    k2 = f[:, None, None] ** 2 + f[None, :, None] ** 2 + fr[None, None, :] ** 2
    kmag = torch.sqrt(k2)
    del k2
    filt = torch.where(kmag > 0, kmag.clamp_min(1e-30) ** ((slope - 2.0) / 2.0), torch.zeros_like(kmag))
    del kmag
    spec *= filt
    del filt
    out = torch.fft.irfftn(spec, s=(n, n, n))
    del spec
    out -= out.mean()
    out *= 1.0 / out.std()
    out += 10.0
    torch.cuda.synchronize()
    return out.contiguous()


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

    import torch
    from bigwhoop_amd import bwc

    n = args.size
    field = make_field_gpu(n)
    if args.copy_calibration:
        # one device copy of exactly n^3 * 8 bytes: the reference reading for
        # the FETCH_SIZE / WRITE_SIZE counters of the DWT kernels
        scratch = torch.empty_like(field)
        scratch.copy_(field)
        torch.cuda.synchronize()
        del scratch
    from bigwhoop_amd import skipnet

    ptr = field.data_ptr()

    def run(codec, steps, src=ptr):
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

    # plain encoder (every pass of every codeblock coded)
    plain = bwc.Codec((n, n, n), pinned=True, **SETTINGS)
    for _ in range(args.warmup):
        plain.compress_into(ptr)
    dt_plain, cs_plain, dwt_plain = run(plain, args.steps)
    plain.close()
    print(f"bench: plain encoder {dt_plain / args.steps * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
    # learned pass skip (the headline encoder); the codestream goes to a
    # page-locked host buffer, as a caller that wants the DMA path allocates it
    codec = bwc.Codec((n, n, n), pinned=True, **SETTINGS)
    codec.set_skip_model(skipnet.default_weights())
    codec.set_pass_skip(2, SKIP_MARGIN)
    for _ in range(args.warmup):
        codec.compress_into(ptr)
    dt, cs, dwt_ms = run(codec, args.steps)
    print(f"bench: learned pass skip {dt / args.steps * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
    ms_last, _ = codec.timings()
    st = codec.skip_stats()
    codec.close()
    # the same into ordinary (pageable) host memory
    pg = bwc.Codec((n, n, n), **SETTINGS)
    pg.set_skip_model(skipnet.default_weights())
    pg.set_pass_skip(2, SKIP_MARGIN)
    for _ in range(args.warmup):
        pg.compress_into(ptr)
    dt_pg, cs_pg, _ = run(pg, args.steps)
    pg.close()
    print(f"bench: pageable output {dt_pg / args.steps * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
    if cs_pg != cs:
        print("pageable-output codestream differs", file=sys.stderr)
        sys.exit(1)
    # the field handed over in page-locked host memory: bwc_compress copies it
    # to HBM over PCIe inside the call (not the headline value)
    host = torch.empty(n ** 3, dtype=torch.float64, pin_memory=True)
    host.copy_(field.reshape(-1))
    hi = bwc.Codec((n, n, n), pinned=True, **SETTINGS)
    hi.set_skip_model(skipnet.default_weights())
    hi.set_pass_skip(2, SKIP_MARGIN)
    for _ in range(args.warmup):
        hi.compress_into(host.data_ptr())
    dt_hi, cs_hi, _ = run(hi, args.steps, host.data_ptr())
    hi.close()
    del host
    print(f"bench: host input {dt_hi / args.steps * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
    if cs_hi != cs:
        print("host-input codestream differs", file=sys.stderr)
        sys.exit(1)
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
    err = (dec - field.reshape(-1)).double()
    psnr_db = 20 * math.log10(float(field.max() - field.min()) / (2 * math.sqrt(float((err * err).mean()))))
    del dec, err
    if cs != cs_plain:
        print("learned pass skip and plain encoder codestreams differ", file=sys.stderr)
        sys.exit(1)

    bytes_in = n ** 3 * 8
    value = bytes_in * args.steps / dt / 1e9
    plain_gbs = bytes_in * args.steps / dt_plain / 1e9
    alg = dwt_algorithmic_bytes(n, 5)
    dwt_all = dwt_ms + dwt_plain
    dwt_avg = sum(dwt_all) / len(dwt_all)
    achieved = alg / (dwt_avg / 1e3) / 1e9
    cpu = None if args.no_cpu_baseline else cpu_baseline()
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
                   "encoder": f"learned pass skip (margin {SKIP_MARGIN} passes)", "codestream_bytes": len(cs),
                   "codestream_identical_to_plain": True},
        "pageable_output": {"value": round(bytes_in * args.steps / dt_pg / 1e9, 4), "unit": "GB/s",
                            "ms_per_step": round(dt_pg / args.steps * 1e3, 3)},
        "host_input": {"value": round(bytes_in * args.steps / dt_hi / 1e9, 4), "unit": "GB/s",
                       "ms_per_step": round(dt_hi / args.steps * 1e3, 3),
                       "input": "page-locked host memory, copied to HBM inside bwc_compress (PCIe)"},
        "decompress": {"value": round(bytes_in * args.steps / dt_dec / 1e9, 4), "unit": "GB/s",
                       "ms_per_step": round(dt_dec / args.steps * 1e3, 3), "psnr_db": round(psnr_db, 3),
                       "output": "device memory"},
        "plain_encoder": {"value": round(plain_gbs, 4), "unit": "GB/s",
                          "ms_per_step": round(dt_plain / args.steps * 1e3, 3),
                          "speedup_learned_vs_plain": round(dt_plain / dt, 3)},
        "pass_skip": {"passes_coded": st["passes_coded"], "passes_complete": st["passes_complete"],
                      "extra_rate_control_rounds": st["extra_rounds"], "codeblocks_coded_further": st["extended"],
                      "fallbacks": st["fallbacks"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": dwt_traffic(),
                     "kernel": "DWT stage (15 axis passes, normalisation fused)", "dwt_ms": round(dwt_avg, 4),
                     "algorithmic_bytes": alg},
        "stages_ms": {k: round(v, 3) for k, v in zip(
            ["fill_minmax", "normalize", "dwt", "quantize", "tier1", "records_d2h", "rate_control", "headers_gather",
             "assembly_d2h", "total"], ms_last)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
