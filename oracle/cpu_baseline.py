"""CPU baseline of the reference's compression path (a baseline only, never a
target): the reference library built from source (oracle/_ref/libbwc_ref.so,
`make -C oracle ref`: -O3, FP64 build, OpenMP) and the project's C oracle
(oracle/liboracle.so), one bwc_compress per process (the reference's gain
table accumulates per process), at 1 and N OpenMP threads, on the baseline
configurations:

  C1   64^3 float64, four sinusoidal modes + noise, 3 levels, cb 16^3, "4"
  C2   512^3 float64 k^-5/3 GRF, 5 levels, cb 32x32x8, "2"
  C2'  the same field at "1,2,4,8"
  C3s  the 4-D five-variable configuration (decomp 5,5,5,2, cb 16x16x8x4, "2")
       on a 128^3 x 16 sample of the field generator (1/8 of C3)

    python oracle/cpu_baseline.py [--threads 1,8] [--configs C1,C2,C2',C3s]

Prints one markdown table row per configuration: seconds and GB/s of the
reference at each thread count and of the oracle at the largest one, and
their ratio.
"""
import argparse
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent

CONFIGS = {
    "C1": dict(gen=("modes", (64, 64, 64), 1, 1), rate="4", decomp=(3, 3, 3, 0), cb=(4, 4, 4, 0)),
    "C2": dict(gen=("grf", (512, 512, 512), 1, 1), rate="2", decomp=(5, 5, 5, 0), cb=(5, 5, 3, 0)),
    "C2'": dict(gen=("grf", (512, 512, 512), 1, 1), rate="1,2,4,8", decomp=(5, 5, 5, 0), cb=(5, 5, 3, 0)),
    "C3s": dict(gen=("turb4d", (128, 128, 128, 16), 7, 5), rate="2", decomp=(5, 5, 5, 2), cb=(4, 4, 3, 2)),
}

# one compression in a fresh process: the field, the library, the settings
RUN = r"""
import ctypes as C, sys, time, numpy as np
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[1] + "/tests")
from bigwhoop_amd.fields import grf, modes_field, turbulence_4d
kind, shape, seed, npar = sys.argv[4], tuple(int(v) for v in sys.argv[5].split(",")), int(sys.argv[6]), int(sys.argv[7])
rate, decomp, cb = sys.argv[8], [int(v) for v in sys.argv[9].split(",")], [int(v) for v in sys.argv[10].split(",")]
if kind == "modes": x = modes_field(shape, seed=seed)
elif kind == "grf": x = grf(shape, seed=seed)
else: x = turbulence_4d(shape, nvar=npar, seed=seed)
x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
s4 = list(shape) + [1] * (4 - len(shape))
if sys.argv[3] == "reference":
    lib = C.CDLL(sys.argv[2])
    vp, u8, u64 = C.c_void_p, C.c_uint8, C.c_uint64
    lib.bwc_alloc_coder.restype = vp; lib.bwc_alloc_coder.argtypes = [u64, u64, u64, u64, u8, C.c_int]
    lib.bwc_init_stream.restype = vp; lib.bwc_init_stream.argtypes = [vp, vp, C.c_int]
    lib.bwc_set_decomp.argtypes = [vp, u8, u8, u8, u8]; lib.bwc_set_codeblocks.argtypes = [vp, u8, u8, u8, u8]
    lib.bwc_create_compression.restype = C.c_ubyte; lib.bwc_create_compression.argtypes = [vp, vp, C.c_char_p]
    lib.bwc_compress.restype = C.c_size_t; lib.bwc_compress.argtypes = [vp, vp]
    out = np.zeros(x.nbytes + (1 << 24), np.uint8)
    codec = lib.bwc_alloc_coder(s4[0], s4[1], s4[2], s4[3], npar, 8)
    lib.bwc_set_decomp(codec, *decomp); lib.bwc_set_codeblocks(codec, *cb)
    st = lib.bwc_init_stream(x.ctypes.data_as(vp), out.ctypes.data_as(vp), 0)
    r = C.create_string_buffer(rate.encode())
    assert lib.bwc_create_compression(codec, st, r) == 0
    t = time.perf_counter(); sz = lib.bwc_compress(codec, st); dt = time.perf_counter() - t
else:
    import oracle_lib as ol
    o = ol.Oracle(s4, npar=npar, rate=rate, decomp=tuple(decomp), cb=tuple(cb))
    t = time.perf_counter(); sz = len(o.compress(x)); dt = time.perf_counter() - t
print(dt, sz, x.nbytes)
"""


def run(kind, cfg, threads):
    lib = ROOT / "oracle" / ("_ref/libbwc_ref.so" if kind == "reference" else "liboracle.so")
    g = cfg["gen"]
    args = [sys.executable, "-c", RUN, str(ROOT), str(lib), kind, g[0], ",".join(map(str, g[1])), str(g[2]),
            str(g[3]), cfg["rate"], ",".join(map(str, cfg["decomp"])), ",".join(map(str, cfg["cb"]))]
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    p = subprocess.run(args, capture_output=True, text=True, env=env, timeout=3600)
    if p.returncode != 0:
        raise SystemExit(f"{kind} {threads} threads failed: {p.stderr[-500:]}")
    dt, sz, nb = p.stdout.split()[-3:]
    return float(dt), int(sz), int(nb)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", default=f"1,{os.cpu_count() or 1}")
    ap.add_argument("--configs", default=",".join(CONFIGS))
    a = ap.parse_args()
    ts = [int(t) for t in a.threads.split(",")]
    print("| config | input GB | " + " | ".join(f"reference {t} thr s (GB/s)" for t in ts) +
          f" | oracle {ts[-1]} thr s (GB/s) | ratio oracle/reference |")
    print("|---|---|" + "---|" * (len(ts) + 2))
    for name in a.configs.split(","):
        cfg = CONFIGS[name]
        cells, last = [], None
        for t in ts:
            dt, sz, nb = run("reference", cfg, t)
            cells.append(f"{dt:.2f} ({nb / dt / 1e9:.3f})")
            last = dt
        do, szo, nb = run("oracle", cfg, ts[-1])
        if szo != sz:
            raise SystemExit(f"{name}: oracle codestream size {szo} differs from the reference's {sz}")
        print(f"| {name} | {nb / 1e9:.3f} | " + " | ".join(cells) + f" | {do:.2f} ({nb / do / 1e9:.3f}) | {do / last:.2f} |",
              flush=True)


if __name__ == "__main__":
    main()
