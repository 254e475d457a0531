"""Write the HDF5-filter replay's golden digests from the reference library.

Run in a container that has the reference checkout:

    make -C oracle ref && python tests/golden/make_h5z_golden.py

1. tests/golden/tgv128_trig.bin: the five 128-entry sine / cosine tables the
   Taylor-Green generator of the reference's HDF5 test evaluates
   (test_hdf5.c:121-128: PI and the grid step rounded as there), computed
   with this machine's C library (math.sin / math.cos call it).
2. oracle/_ref/h5z_replay_ref (tests/h5z_replay.c built against
   include/bwc.h and linked with the reference library) compresses and
   decompresses every chunk; the SHA-256 and size of each codestream and
   decoded chunk, and the five variables' PSNR, go to
   tests/golden/h5z_cases.json.
3. tests/golden/abi_layout.txt: tests/abi_layout.c built against the
   reference's own headers prints the offsets and sizes of the public
   structs and the enum values.
"""
from __future__ import annotations

import hashlib
import json
import math
import struct
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
REPLAY = ROOT / "oracle" / "_ref" / "h5z_replay_ref"


def trig_tables(n=128):
    pi_f = struct.unpack("f", struct.pack("f", 3.14159265358979323846))[0]
    dx = (2.0 * pi_f) / n
    rows = [[math.sin(i * dx) for i in range(n)], [math.cos(i * dx) for i in range(n)],
            [math.sin(j * dx) for j in range(n)], [math.cos(j * dx) for j in range(n)],
            [math.cos(k * dx) for k in range(n)]]
    return np.array(rows, np.float64)


def main():
    if not REPLAY.exists():
        raise SystemExit("oracle/_ref/h5z_replay_ref missing: run `make -C oracle ref` where the reference exists")
    trig = HERE / "tgv128_trig.bin"
    trig.write_bytes(trig_tables().tobytes())
    with tempfile.TemporaryDirectory() as td:
        p = subprocess.run([str(REPLAY), td, str(trig)], capture_output=True, text=True, check=True)
        out = {"chunks": [], "tgv_psnr": None}
        for line in p.stdout.splitlines():
            f = line.split()
            if f[0] == "tgv_psnr":
                out["tgv_psnr"] = float(f[1])
                continue
            name = f[0]
            cs = Path(td, name + ".bwc").read_bytes()
            dec = Path(td, name + ".dec").read_bytes()
            out["chunks"].append(dict(name=name, input_bytes=int(f[1]), codestream_size=len(cs),
                                      codestream_sha256=hashlib.sha256(cs).hexdigest(), decode_size=len(dec),
                                      decode_sha256=hashlib.sha256(dec).hexdigest()))
            print(name, len(cs))
    out["trig_sha256"] = hashlib.sha256(trig.read_bytes()).hexdigest()
    with tempfile.TemporaryDirectory() as td:
        exe = str(Path(td, "abi_ref"))
        subprocess.run(["gcc", "-std=gnu11", "-DREF_HEADERS", "-DBWC_DOUBLE_PRECISION",
                        "-I/root/reference/include/library/private", "-o", exe, str(ROOT / "tests" / "abi_layout.c")],
                       check=True)
        (HERE / "abi_layout.txt").write_text(subprocess.run([exe], capture_output=True, text=True,
                                                            check=True).stdout)
    (HERE / "h5z_cases.json").write_text(json.dumps(out, indent=1))
    print("tgv_psnr", out["tgv_psnr"])
    return 0


if __name__ == "__main__":
    sys.exit(main())
