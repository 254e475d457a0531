"""Pass-limit network: training data from the plain encoder, training and
inference on the GPU through libbwc (bwc_amd_skipnet_* / bwc_amd_set_skip_model).

The network predicts, per codeblock, how many coding passes PCRD keeps at the
requested bitrate (in units of 8 passes). Tier-1 then codes each codeblock
only to that prediction plus a margin; rate control proves the result equals
the complete coding's (coding codeblocks further where it cannot), so the
codestream is the plain encoder's, byte for byte.

    python -m bigwhoop_amd.skipnet --fields 200 --size 256 --rates 0.5,1,2,4,8 --nd-fields 400

collects (features, cut) pairs from the plain encoder on seeded 3-D float64
Gaussian random fields with spectral slopes U(-3, -1) (5 levels, codeblocks
32x32x8), and on fields of 1 to 4 axes drawn like the uneven batch workload
(random extents and aspect ratios, float32 or float64, its level and
codeblock rules, three bitrates each log-uniform in 0.5-8 bpv), trains on the
GPU and writes bigwhoop_amd/data/skipnet_w.npy (+ .json with the training
record). A codeblock's target is its cut in the last quality layer, so a
rate string's top rate is what the network predicts for.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import bwc

PKG = Path(__file__).resolve().parent
# two trained instances of the network: one on the headline population (3-D
# float64 fields, 5 levels, codeblocks 32x32x8, 1 and 2 bpv), one on a wide
# population (that plus fields of 1 to 4 axes drawn like the uneven batch, and
# rates 0.5 to 8 bpv)
WEIGHTS = PKG / "data" / "skipnet_w.npy"
WEIGHTS_WIDE = PKG / "data" / "skipnet_w_wide.npy"
NPARAM = 5345
MU, SC = 5313, 5329
CODEC = dict(decomp=(5, 5, 5, 0), cb=(5, 5, 3, 0))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _load(path) -> np.ndarray:
    w = np.load(path, allow_pickle=False).astype(np.float32)
    if w.size != NPARAM:
        raise ValueError("weight file has the wrong size")
    return w


def default_weights() -> np.ndarray:
    """The weights trained on the headline population (float32, 5345 values)."""
    return _load(WEIGHTS)


def wide_weights() -> np.ndarray:
    """The weights trained on fields of 1 to 4 axes and rates 0.5 to 8 bpv."""
    return _load(WEIGHTS_WIDE)


def choose(ndim: int, top_rate: float):
    """(weights, margin) for a field of `ndim` non-unit axes whose highest
    layer rate is `top_rate` bpv: the headline population's network with 6
    passes of margin for 3-D fields up to 2 bpv, else the wide one with 8.
    Measured on 512^3 float64 k^-5/3 fields: at 2 bpv the first needs no
    re-coding round at margin 6 (40 ms) and the wide one needs margin 7
    (46 ms); at 4 bpv and for layers 1,2,4,8 the wide one at margin 8 needs
    none (58 ms and 90 ms, against a round or two with the first)."""
    if ndim == 3 and top_rate <= 2.0:
        return default_weights(), 6
    return wide_weights(), 8


def init_weights(seed: int, mu: np.ndarray, sc: np.ndarray) -> np.ndarray:
    rng = np.random.default_rng(seed)
    w = np.zeros(NPARAM, np.float32)
    w[0:1024] = rng.normal(0, np.sqrt(2.0 / 16), 1024)
    w[1088:5184] = rng.normal(0, np.sqrt(2.0 / 64), 4096)
    w[5248:5312] = rng.normal(0, np.sqrt(1.0 / 64), 64)
    w[MU:SC] = mu
    w[SC:SC + 16] = sc
    return w


def gpu_train(X: np.ndarray, t: np.ndarray, w: np.ndarray, epochs: int, lr: float, seed: int = 0):
    X = np.ascontiguousarray(X, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    w = np.ascontiguousarray(w, np.float32).copy()
    loss = np.zeros(max(1, epochs), np.float32)
    if bwc.lib().bwc_amd_skipnet_train(_fp(X), _fp(t), len(t), _fp(w), epochs, lr, seed, _fp(loss)):
        raise RuntimeError("bwc_amd_skipnet_train failed: " + bwc.last_error())
    return w, loss


def gpu_grad(X: np.ndarray, t: np.ndarray, w: np.ndarray):
    X = np.ascontiguousarray(X, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    g = np.zeros(5313, np.float32)
    loss = np.zeros(1, np.float32)
    if bwc.lib().bwc_amd_skipnet_grad(_fp(X), _fp(t), len(t), _fp(w), _fp(g), _fp(loss)):
        raise RuntimeError("bwc_amd_skipnet_grad failed: " + bwc.last_error())
    return float(loss[0]), g


def gpu_predict(X: np.ndarray, w: np.ndarray) -> np.ndarray:
    X = np.ascontiguousarray(X, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    y = np.zeros(len(X), np.float32)
    if bwc.lib().bwc_amd_skipnet_predict(_fp(X), len(X), _fp(w), _fp(y)):
        raise RuntimeError("bwc_amd_skipnet_predict failed: " + bwc.last_error())
    return y


def collect(n: int, seed: int, slope: float, rates, log=None):
    """Features and cuts of every codeblock (K > 0) of one n^3 field, from the
    plain encoder, at each rate."""
    from .fields import grf_device

    f = grf_device(n, slope=slope, seed=seed)
    Xs, ts = [], []
    for rate in rates:
        c = bwc.Codec((n, n, n), rate=str(rate), **CODEC)
        try:
            c.collect_features(True)
            c.compress_ptr(f.data_ptr())
            X = c.features()
            cut = c.cuts().astype(np.float32)
        finally:
            c.close()
        keep = X[:, 0] > 0
        Xs.append(X[keep])
        ts.append(cut[keep] / 8.0)
    del f
    return np.concatenate(Xs), np.concatenate(ts)


def collect_nd(f, rates):
    """Features and cuts of every codeblock (K > 0) of one field described
    like a batch_spec entry, from the plain encoder, at each rate."""
    import torch

    from .workloads import grf_nd_device

    dt = torch.float32 if f["dtype"] == "float32" else torch.float64
    x = grf_nd_device(f["shape"], f["slope"], f["seed"], dt)
    Xs, ts = [], []
    for rate in rates:
        c = bwc.Codec(f["shape"], prec=4 if f["dtype"] == "float32" else 8, rate=f"{rate:.3f}",
                      decomp=f["decomp"], cb=f["cb"])
        try:
            c.collect_features(True)
            c.compress_ptr(x.data_ptr())
            X = c.features()
            cut = c.cuts().astype(np.float32)
        finally:
            c.close()
        keep = X[:, 0] > 0
        Xs.append(X[keep])
        ts.append(cut[keep] / 8.0)
    del x
    return np.concatenate(Xs), np.concatenate(ts)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--fields", type=int, default=200)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--rates", default="1,2")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=20240601)
    ap.add_argument("--nd-fields", type=int, default=0,
                    help="fields of 1 to 4 axes drawn like the uneven batch (1e5 to 4e6 values)")
    ap.add_argument("--out", default=str(WEIGHTS))
    ap.add_argument("--data-out", default="", help="also save the (features, targets) set as .npz")
    args = ap.parse_args(argv)
    rates = [float(r) for r in args.rates.split(",")]
    rng = np.random.default_rng(args.seed)
    Xs, ts = [], []
    t0 = time.time()
    for i in range(args.fields):
        slope = float(rng.uniform(-3.0, -1.0))
        X, t = collect(args.size, args.seed + i, slope, rates)
        Xs.append(X)
        ts.append(t)
        print(f"field {i + 1}/{args.fields} slope {slope:.2f}: {len(t)} samples, "
              f"{time.time() - t0:.1f} s", flush=True)
    if args.nd_fields:
        from .workloads import batch_spec

        for i, f in enumerate(batch_spec(args.nd_fields, seed=args.seed % 100000 + 7, nmin=1e5, nmax=4e6)):
            rs = [float(np.exp(rng.uniform(np.log(0.5), np.log(8.0)))) for _ in range(3)]
            X, t = collect_nd(f, rs)
            Xs.append(X)
            ts.append(t)
            print(f"nd field {i + 1}/{args.nd_fields} {f['shape']} {f['dtype']}: {len(t)} samples, "
                  f"{time.time() - t0:.1f} s", flush=True)
    X = np.concatenate(Xs)
    t = np.concatenate(ts)
    if args.data_out:
        np.savez_compressed(args.data_out, X=X, t=t)
    mu = X.mean(0)
    sd = X.std(0)
    sc = np.where(sd > 1e-6, 1.0 / np.maximum(sd, 1e-6), 1.0).astype(np.float32)
    w0 = init_weights(args.seed, mu, sc)
    w, loss = gpu_train(X, t, w0, args.epochs, args.lr, args.seed)
    y = gpu_predict(X, w)
    err = 8 * (y - t)
    rec = {"samples": int(len(t)), "fields": args.fields, "size": args.size, "rates": rates,
           "nd_fields": args.nd_fields,
           "epochs": args.epochs, "lr": args.lr, "loss": [float(v) for v in loss],
           "train_error_passes": {"mean": float(err.mean()), "p05": float(np.percentile(err, 5)),
                                  "p50": float(np.percentile(err, 50)), "p95": float(np.percentile(err, 95))},
           "under_by_more_than_2": float(np.mean(err < -2))}
    out = Path(args.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    np.save(out, w.astype(np.float32))
    out.with_suffix(".json").write_text(json.dumps(rec, indent=1))
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    sys.exit(main())
