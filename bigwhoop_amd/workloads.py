"""The larger measured workloads, next to the headline one in bench.py.

    python -m bigwhoop_amd.workloads learned1024   # 1024^3 f64 k^-5/3, 1 and 2 bpv
    python -m bigwhoop_amd.workloads layers512     # 512^3 f64, 4 layers at 1, 2, 4, 8 bpv
    python -m bigwhoop_amd.workloads turb4d        # 256^3 x 16 x 5 variables, 2 bpv
    python -m bigwhoop_amd.workloads batch64       # 64 uneven fields (1-4 axes, f32/f64)
    python -m bigwhoop_amd.workloads modes64       # 64^3 f64, four modes + noise, 4 bpv

Each prints one JSON line per case: input GB/s of the plain encoder (every
pass coded) and of the learned pass-skip encoder on the same field, the
speedup, and whether the two codestreams are byte-identical (they must be;
the run fails otherwise). Fields are synthesised in HBM with torch's FFT
(seeded) and compressed in place through the C API (a device pointer).

* learned1024: 1024^3 float64 Gaussian random field, E(k) ~ k^-5/3, unit
  variance, mean 10; 5 levels per axis, codeblocks 32x32x8, one layer.
* layers512: the 512^3 field of the headline workload with four quality
  layers at 1, 2, 4 and 8 bits per value.
* turb4d: 256^3 x 16 time steps, five variables: u, v, w from independent
  k^-5/3 fields, p and rho from k^-7/3 fields plus a 5% Gaussian pulse
  background; in time the complex spectral coefficients follow a
  first-order AR(0.95) process, so the phases evolve smoothly. 5 spatial and
  2 temporal DWT levels, codeblocks 16x16x8x4, 2 bits per value.
* modes64: the reference's CPU-runnable case: 64^3 float64, the sum of four
  seeded sinusoidal modes (|k| <= 8) plus N(0, 1e-3) noise, 3 levels per
  axis, codeblocks 16x16x16, one layer at 4 bits per value (made on the host
  and copied to HBM).
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

def _top_rate(rate: str) -> float:
    return max(float(r) for r in rate.replace("-", ",").replace(" ", ",").split(",") if r)


def _spectral_filter(n, slope, device):
    import torch

    f = torch.fft.fftfreq(n, d=1.0 / n, device=device, dtype=torch.float64)
    fr = torch.fft.rfftfreq(n, d=1.0 / n, device=device, dtype=torch.float64)
    kmag = torch.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2 + fr[None, None, :] ** 2)
    return torch.where(kmag > 0, kmag.clamp_min(1e-30) ** ((slope - 2.0) / 2.0), torch.zeros_like(kmag))


def turbulence_4d(n=256, nt=16, seed=7, rho=0.95):
    """Five variables (u, v, w, p, rho) of n^3 x nt float64 values, layout
    [variable][t][z][y][x], as one contiguous device tensor."""
    import torch

    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    out = torch.empty((5, nt, n, n, n), dtype=torch.float64, device="cuda")
    for v in range(5):
        slope = -5.0 / 3.0 if v < 3 else -7.0 / 3.0
        filt = _spectral_filter(n, slope, "cuda")
        shape = (n, n, n // 2 + 1)
        w = torch.complex(torch.randn(shape, generator=g, device="cuda", dtype=torch.float64),
                          torch.randn(shape, generator=g, device="cuda", dtype=torch.float64))
        s = (1.0 - rho * rho) ** 0.5
        for t in range(nt):
            if t:
                w = rho * w + s * torch.complex(torch.randn(shape, generator=g, device="cuda", dtype=torch.float64),
                                                torch.randn(shape, generator=g, device="cuda", dtype=torch.float64))
            x = torch.fft.irfftn(w * filt, s=(n, n, n))
            x -= x.mean()
            x /= x.std()
            if v >= 3:
                # 5% Gaussian pulse background, centred, width n/6
                ax = torch.arange(n, device="cuda", dtype=torch.float64) - n / 2
                r2 = ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2
                x += 0.05 * torch.exp(-r2 / (2 * (n / 6.0) ** 2)) * x.abs().max()
            out[v, t] = x + 10.0
        del w, filt
    torch.cuda.synchronize()
    return out


def grf_nd_device(shape, slope, seed, dtype):
    """Gaussian random field on the GPU over the axes of `shape` (x fastest;
    1 for absent axes): E(k) ~ k^slope in d dimensions (amplitude filter
    k^((slope - (d - 1)) / 2)), unit variance, mean 10, cast to dtype. The
    field is synthesised on the next power-of-two grid and cropped, so the
    FFTs use few distinct lengths."""
    import torch

    dims = [n for n in shape if n > 1] or [shape[0]]
    d = len(dims)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    want = tuple(dims[::-1])     # numpy order: slowest axis first
    tshape = tuple(1 << max(3, (n - 1).bit_length()) for n in want)
    white = torch.randn(tshape, generator=g, device="cuda", dtype=torch.float64)
    spec = torch.fft.rfftn(white)
    del white
    k2 = None
    for a, n in enumerate(tshape):
        f = (torch.fft.rfftfreq(n, d=1.0 / n, device="cuda", dtype=torch.float64) if a == d - 1
             else torch.fft.fftfreq(n, d=1.0 / n, device="cuda", dtype=torch.float64))
        sh = [1] * d
        sh[a] = -1
        t = (f ** 2).reshape(sh)
        k2 = t if k2 is None else k2 + t
    kmag = torch.sqrt(k2)
    del k2
    spec *= torch.where(kmag > 0, kmag.clamp_min(1e-30) ** ((slope - (d - 1)) / 2.0), torch.zeros_like(kmag))
    del kmag
    out = torch.fft.irfftn(spec, s=tshape)
    del spec
    out = out[tuple(slice(0, n) for n in want)].contiguous()
    out -= out.mean()
    out /= out.std()
    out += 10.0
    return out.to(dtype).contiguous()


def batch_spec(count=64, seed=4, nmin=1e5, nmax=1e8):
    """The uneven batch: `count` fields with dimensionality uniform in 1-4,
    element count log-uniform in [nmin, nmax] split over the axes with random
    aspect ratios, float32 or float64 (50/50), spectral slope U(-3, -1),
    bitrate log-uniform in [0.5, 8] bits per value, DWT levels
    min(5, floor(log2(shortest axis)) - 2) on every axis (at least 0), and
    codeblocks of 2^10, 2^5 x 2^5, 2^5 x 2^5 x 2^3 or 2^4 x 2^4 x 2^3 x 2^2 values
    for 1-4 axes. An extent above 2^24 that would round up to a float is
    lowered to the float below it (the codec refuses the former)."""
    import numpy as np

    rng = np.random.Generator(np.random.PCG64(seed))
    cbs = {1: (10, 0, 0, 0), 2: (5, 5, 0, 0), 3: (5, 5, 3, 0), 4: (4, 4, 3, 2)}
    out = []
    for i in range(count):
        d = int(rng.integers(1, 5))
        n = float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        w = rng.uniform(-1.0, 1.0, size=d)
        w -= w.mean()
        dims = [max(8, int(round(n ** (1.0 / d) * np.exp(x)))) for x in w]
        # extents above 2^24 that round up to a float are refused by the codec
        # (single-precision resolution boxes): take the float just below
        dims = [m if int(np.float32(m)) <= m else int(np.nextafter(np.float32(m), np.float32(0))) for m in dims]
        shape = tuple(dims + [1] * (4 - d))
        lv = max(0, min(5, int(np.floor(np.log2(min(dims)))) - 2))
        out.append(dict(shape=shape, ndim=d, dtype="float32" if rng.uniform() < 0.5 else "float64",
                        slope=float(rng.uniform(-3.0, -1.0)),
                        rate=f"{float(np.exp(rng.uniform(np.log(0.5), np.log(8.0)))):.3f}",
                        decomp=tuple(lv if a < d else 0 for a in range(4)), cb=cbs[d], seed=1000 + i))
    return out


def batch_fields(spec):
    """The fields of batch_spec on the GPU (one tensor each)."""
    import torch

    fields = []
    for i, f in enumerate(spec):
        dt = torch.float32 if f["dtype"] == "float32" else torch.float64
        fields.append(grf_nd_device(f["shape"], f["slope"], f["seed"], dt))
        torch.cuda.synchronize()
        print(f"batch: field {i} {f['shape']} {f['dtype']} made", file=sys.stderr, flush=True)
    return fields


def _time(codec, ptr, steps):
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    size = 0
    for _ in range(steps):
        size = codec.compress_into(ptr)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, codec.out[:size].tobytes()


def compare(name, shape, npar, ptr, nbytes, settings, steps=2, warmup=1):
    """Plain and learned encoders on the same device field; one JSON line."""
    from bigwhoop_amd import bwc, skipnet

    res = {"workload": name, "shape": list(shape), "variables": npar, "settings": {k: (list(v) if isinstance(v, tuple)
                                                                                       else v)
                                                                                   for k, v in settings.items()}}
    plain = bwc.Codec(shape, npar=npar, **settings)
    for _ in range(warmup):
        plain.compress_into(ptr)
    dt_p, cs_p = _time(plain, ptr, steps)
    plain.close()
    learned = bwc.Codec(shape, npar=npar, **settings)
    w, margin = skipnet.choose(sum(1 for n in shape if n > 1), _top_rate(settings.get("rate", "2")))
    learned.set_skip_model(w)
    learned.set_pass_skip(2, margin)
    res["margin"] = margin
    for _ in range(warmup):
        learned.compress_into(ptr)
    dt_l, cs_l = _time(learned, ptr, steps)
    st = learned.skip_stats()
    learned.close()
    res.update({
        "plain_gbs": round(nbytes / dt_p / 1e9, 3), "plain_ms": round(dt_p * 1e3, 2),
        "learned_gbs": round(nbytes / dt_l / 1e9, 3), "learned_ms": round(dt_l * 1e3, 2),
        "speedup_learned_vs_plain": round(dt_p / dt_l, 3), "codestream_bytes": len(cs_l),
        "identical": cs_l == cs_p, "pass_skip": st,
    })
    print(json.dumps(res), flush=True)
    if cs_l != cs_p:
        raise SystemExit(f"{name}: learned and plain codestreams differ")
    return res


def batch64(count=64, seed=4, steps=2):
    """The uneven batch compressed field after field (one bwc_compress per
    field) by the plain and by the learned encoder (network and margin from
    skipnet.choose), then all fields concurrently; aggregate
    input GB/s over the whole batch; every field's codestreams must be
    identical across the four runs."""
    import torch
    from bigwhoop_amd import bwc, skipnet

    spec = batch_spec(count, seed)
    fields = batch_fields(spec)
    total = sum(f.numel() * f.element_size() for f in fields)

    def codecs(learned):
        cs = []
        for f, t in zip(spec, fields):
            c = bwc.Codec(f["shape"], prec=t.element_size(), rate=f["rate"], decomp=f["decomp"], cb=f["cb"])
            if learned:
                w, margin = skipnet.choose(f["ndim"], _top_rate(f["rate"]))
                c.set_skip_model(w)
                c.set_pass_skip(2, margin)
            cs.append(c)
        return cs

    def run(cs):
        per = []
        for i, (c, t) in enumerate(zip(cs, fields)):
            t0 = time.perf_counter()
            c.compress_into(t.data_ptr())
            per.append(time.perf_counter() - t0)
            print(f"batch: field {i} compressed in {per[-1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sizes = []
        for _ in range(steps):
            sizes = [c.compress_into(t.data_ptr()) for c, t in zip(cs, fields)]
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        return dt, [c.out[:n].tobytes() for c, n in zip(cs, sizes)], per

    def run_batch(cs, ref, mode):
        # "queue": tier-1 of every field through the batch's global codeblock
        # queue; "threads": each field compressed on its own on host threads
        if mode == "threads":
            os.environ["BWC_AMD_BATCH"] = "threads"
        else:
            os.environ.pop("BWC_AMD_BATCH", None)
        ptrs = [t.data_ptr() for t in fields]
        bwc.compress_batch(cs, ptrs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            sizes = bwc.compress_batch(cs, ptrs)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        os.environ.pop("BWC_AMD_BATCH", None)
        if any(c.out[:n].tobytes() != r for c, n, r in zip(cs, sizes, ref)):
            raise SystemExit("batch: concurrent compression differs from one field at a time")
        return dt

    pl = codecs(False)
    dt_p, cs_p, per_p = run(pl)
    print(f"batch: plain {dt_p * 1e3:.1f} ms per pass", file=sys.stderr, flush=True)
    dtc_p = run_batch(pl, cs_p, "queue")
    dtt_p = run_batch(pl, cs_p, "threads")
    print(f"batch: plain, one queue {dtc_p * 1e3:.1f} ms, threads {dtt_p * 1e3:.1f} ms per pass", file=sys.stderr,
          flush=True)
    for c in pl:
        c.close()
    le = codecs(True)
    dt_l, cs_l, per_l = run(le)
    print(f"batch: learned {dt_l * 1e3:.1f} ms per pass", file=sys.stderr, flush=True)
    dtc_l = run_batch(le, cs_l, "queue")
    dtt_l = run_batch(le, cs_l, "threads")
    print(f"batch: learned, one queue {dtc_l * 1e3:.1f} ms, threads {dtt_l * 1e3:.1f} ms per pass", file=sys.stderr,
          flush=True)
    stats = [c.skip_stats() for c in le]
    for c in le:
        c.close()
    same = all(a == b for a, b in zip(cs_p, cs_l))
    res = {"workload": f"batch of {count} uneven fields (seed {seed})",
           "fields": count, "input_bytes": total,
           "dims": [sum(1 for f in spec if f["ndim"] == d) for d in (1, 2, 3, 4)],
           "float32_fields": sum(1 for f in spec if f["dtype"] == "float32"),
           "plain_gbs": round(total / dt_p / 1e9, 3), "plain_ms": round(dt_p * 1e3, 2),
           "learned_gbs": round(total / dt_l / 1e9, 3), "learned_ms": round(dt_l * 1e3, 2),
           "speedup_learned_vs_plain": round(dt_p / dt_l, 3), "codestream_bytes": sum(len(c) for c in cs_l),
           "concurrent_queue": {"plain_gbs": round(total / dtc_p / 1e9, 3), "plain_ms": round(dtc_p * 1e3, 2),
                                "learned_gbs": round(total / dtc_l / 1e9, 3), "learned_ms": round(dtc_l * 1e3, 2),
                                "identical_to_one_at_a_time": True},
           "concurrent_threads": {"plain_gbs": round(total / dtt_p / 1e9, 3), "plain_ms": round(dtt_p * 1e3, 2),
                                  "learned_gbs": round(total / dtt_l / 1e9, 3),
                                  "learned_ms": round(dtt_l * 1e3, 2), "workers": 8,
                                  "identical_to_one_at_a_time": True},
           "identical": same,
           "extra_rounds": sum(s["extra_rounds"] for s in stats), "fallbacks": sum(s["fallbacks"] for s in stats),
           "slowest_fields": [dict(shape=spec[i]["shape"], dtype=spec[i]["dtype"], rate=spec[i]["rate"],
                                   plain_ms=round(per_p[i] * 1e3, 1), learned_ms=round(per_l[i] * 1e3, 1),
                                   extra_rounds=stats[i]["extra_rounds"])
                              for i in sorted(range(count), key=lambda k: -per_l[k])[:12]]}
    print(json.dumps(res), flush=True)
    if not same:
        raise SystemExit("batch: learned and plain codestreams differ")
    return res


def main(argv):
    import torch
    from bigwhoop_amd.fields import grf_device

    which = argv[1:] or ["learned1024", "layers512", "turb4d", "batch64"]
    for w in which:
        if w == "learned1024":
            n = 1024
            f = grf_device(n)
            for rate in ("1", "2"):
                compare(f"1024^3 k^-5/3 at {rate} bpv", (n, n, n), 1, f.data_ptr(), n ** 3 * 8,
                        dict(rate=rate, decomp=(5, 5, 5, 0), cb=(5, 5, 3, 0)))
            del f
        elif w == "layers512":
            n = 512
            f = grf_device(n)
            compare("512^3 k^-5/3, layers 1,2,4,8 bpv", (n, n, n), 1, f.data_ptr(), n ** 3 * 8,
                    dict(rate="1,2,4,8", decomp=(5, 5, 5, 0), cb=(5, 5, 3, 0)))
            del f
        elif w == "modes64":
            from bigwhoop_amd.fields import modes_field

            f = torch.from_numpy(np.ascontiguousarray(modes_field((64, 64, 64), seed=1)).reshape(-1)).to("cuda")
            compare("64^3 four modes + noise at 4 bpv", (64, 64, 64), 1, f.data_ptr(), f.numel() * 8,
                    dict(rate="4", decomp=(3, 3, 3, 0), cb=(4, 4, 4, 0)), steps=10)
            del f
        elif w == "batch64":
            batch64()
        elif w == "turb4d":
            f = turbulence_4d()
            compare("256^3 x 16, 5 variables, 2 bpv", (256, 256, 256, 16), 5, f.data_ptr(), f.numel() * 8,
                    dict(rate="2", decomp=(5, 5, 5, 2), cb=(4, 4, 3, 2)), steps=1)
            del f
        else:
            raise SystemExit(f"unknown workload {w}")
        torch.cuda.empty_cache()


if __name__ == "__main__":
    main(sys.argv)
