"""The larger measured workloads, next to the headline one in bench.py.

    python -m bigwhoop_amd.workloads learned1024   # 1024^3 f64 k^-5/3, 1 and 2 bpv
    python -m bigwhoop_amd.workloads layers512     # 512^3 f64, 4 layers at 1, 2, 4, 8 bpv
    python -m bigwhoop_amd.workloads turb4d        # 256^3 x 16 x 5 variables, 2 bpv

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
"""
from __future__ import annotations

import json
import sys
import time

SKIP_MARGIN = 6


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
    learned.set_skip_model(skipnet.default_weights())
    learned.set_pass_skip(2, SKIP_MARGIN)
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


def main(argv):
    import torch
    from bigwhoop_amd.fields import grf_device

    which = argv[1:] or ["learned1024", "layers512", "turb4d"]
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
