"""The BigWhoop HDF5 filter's calls to libbwc (src/interfaces/HDF5/H5Zbwc.c:
cd_values -> setters -> bwc_create_compression("%05.3f") -> bwc_compress,
and bwc_open_header -> bwc_create_decompression -> bwc_decompress), replayed
by tests/h5z_replay.c, a C program compiled against include/bwc.h. The
golden digests (tests/golden/h5z_cases.json) come from the same program
linked with the reference's own library, so the codestreams, the decoded
chunks and the PSNR of the reference's HDF5 test scenario (five 128^3
Taylor-Green variables at 0.01 bpv, Qm 32, error resilience,
test_hdf5.c:290-443) must match byte for byte through this library."""
import hashlib
import json
import subprocess
import tempfile
from pathlib import Path

import pytest

from bigwhoop_amd import bwc

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
REPLAY = ROOT / "tests" / "_bin" / "h5z_replay"
CASES = json.loads((GOLDEN / "h5z_cases.json").read_text())


def test_trig_fixture_digest():
    assert hashlib.sha256((GOLDEN / "tgv128_trig.bin").read_bytes()).hexdigest() == CASES["trig_sha256"]


def test_replay_links_against_the_library():
    """The replay, built against include/bwc.h, resolves every call in this
    library (the build links it with -lbwc and no other bwc library)."""
    assert REPLAY.exists(), "tests/_bin/h5z_replay is missing: run __graft_entry__.build()"
    out = subprocess.run(["ldd", str(REPLAY)], capture_output=True, text=True).stdout
    assert str(bwc.LIB_PATH.resolve()) in out or "libbwc.so" in out


@pytest.mark.gpu
def test_replay_matches_reference_bytes():
    with tempfile.TemporaryDirectory() as td:
        print("h5z replay", flush=True)
        p = subprocess.run([str(REPLAY), td, str(GOLDEN / "tgv128_trig.bin")], capture_output=True, text=True,
                           timeout=300)
        assert p.returncode == 0, p.stderr[-2000:]
        psnr = None
        for line in p.stdout.splitlines():
            if line.startswith("tgv_psnr"):
                psnr = float(line.split()[1])
        for ch in CASES["chunks"]:
            cs = Path(td, ch["name"] + ".bwc").read_bytes()
            dec = Path(td, ch["name"] + ".dec").read_bytes()
            assert len(cs) == ch["codestream_size"], ch["name"]
            assert hashlib.sha256(cs).hexdigest() == ch["codestream_sha256"], ch["name"]
            assert hashlib.sha256(dec).hexdigest() == ch["decode_sha256"], ch["name"]
    # the reference test's known answer (test_hdf5.c:441-442: fails when
    # PSNR - 75.16 > 0.01), and the value the reference library gives
    assert psnr is not None and psnr - 75.16 <= 0.01
    assert abs(psnr - CASES["tgv_psnr"]) < 1e-6


def test_public_struct_layout_matches_reference(tmp_path):
    """Offsets and sizes of bwc_span, bwc_gl_inf, bwc_stream, bwc_gl_ctrl,
    bwc_header and bwc_codec, and the enum values, through include/bwc.h,
    equal those the reference's headers give (types.h:89-600)."""
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-std=gnu11", "-I" + str(ROOT / "include"), "-o", str(exe),
                    str(ROOT / "tests" / "abi_layout.c")], check=True)
    got = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    assert got == (GOLDEN / "abi_layout.txt").read_text()
