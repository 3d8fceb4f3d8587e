"""The `prism` command line: flags, file-name rules and error behaviour of the
reference's src/utils/parameters.cu:9-234, and (on the GPU) streams and
decompressed files equal to the oracle's for the same field and bound."""
import os
import subprocess

import numpy as np
import pytest
import torch

from prism_amd import fields

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "prism_amd", "_build", "prism")


def run(*args, cwd=None, timeout=120):
    if not os.path.exists(CLI):
        pytest.fail("prism executable not built (run __graft_entry__.build())")
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, cwd=cwd, timeout=timeout)


# ---------------------------------------------------------------- CPU only --
def test_help_prints_usage():
    r = run("-h")
    assert r.returncode == 0
    for flag in ("-i <path>", "-z <path>", "-x <path>", "-f", "-d", "-1 <nx>", "-3 <nx> <ny> <nz>", "-A", "-R",
                 "--report", "-prog", "-errors"):
        assert flag in r.stdout


@pytest.mark.parametrize("args, message", [
    (["-f", "-3", 4, 4, 4, "-z"], "input file path must be specified."),
    (["-i", "x.f32", "-f", "-z"], "dim must be specified."),
    (["-i", "x", "-f", "-3", 8, 8, 8, "-x", "-errors", "-3", 1e-1, 1e-2], "Not enough error bounds provided."),
    (["-f", "-3", 8, 8, 8, "--report", "cr", "-x", "c.prisma"], "input file path must be specified."),
])
def test_argument_errors(args, message):
    r = run(*args)
    assert r.returncode != 0
    assert message in r.stderr


def test_unknown_option_exits_with_failure():
    r = run("-i", "x", "-4", 1, 1, 1, 1)  # -4 is listed by the reference but never parsed
    assert r.returncode != 0
    assert "Unrecognized option '-4'" in r.stdout


# -------------------------------------------------------------------- GPU --
def _field(shape, dtype, seed):
    nz, ny, nx = shape
    f = fields.smooth_mix(nx, ny, nz, seed=seed, dtype=torch.float64, device="cpu").numpy()
    return np.ascontiguousarray(f.astype(dtype))


@pytest.mark.gpu
@pytest.mark.parametrize("dtype, flag", [(np.float32, "-f"), (np.float64, "-d")])
def test_plain_files_equal_oracle(oracle, tmp_path, dtype, flag):
    shape = (33, 36, 40)
    data = _field(shape, dtype, 7)
    src = tmp_path / "field.bin"
    data.tofile(src)
    nz, ny, nx = shape
    r = run("-i", src, flag, "-3", nx, ny, nz, "-R", 1e-3, "-z", tmp_path / "c.prisma", "-x", tmp_path / "d.out",
            "--report", "time,cr")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "compressed file written" in r.stdout and "PSNR" in r.stdout
    ref, _ = oracle.compress(data, 1e-3, mode=1)
    got = np.fromfile(tmp_path / "c.prisma", np.uint8)
    assert np.array_equal(got, ref)
    dec = np.fromfile(tmp_path / "d.out", dtype).reshape(shape)
    dec_o = oracle.decompress(ref, shape, dtype, 1e-3)
    assert np.array_equal(dec.view(np.uint8), dec_o.view(np.uint8))


@pytest.mark.gpu
def test_default_file_names(oracle, tmp_path):
    shape = (20, 24, 28)
    data = _field(shape, np.float32, 3)
    data.tofile(tmp_path / "f.bin")
    nz, ny, nx = shape
    r = run("-i", "f.bin", "-f", "-3", nx, ny, nz, "-A", 1e-2, "-z", "-x", cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    ref, _ = oracle.compress(data, 1e-2, abs_mode=True, mode=1)
    assert np.array_equal(np.fromfile(tmp_path / "f.bin.prisma", np.uint8), ref)
    dec_o = oracle.decompress(ref, shape, np.float32, 1e-2, abs_mode=True)
    dec = np.fromfile(tmp_path / "f.bin.prisma.out", np.float32).reshape(shape)
    assert np.array_equal(dec.view(np.uint32), dec_o.view(np.uint32))


@pytest.mark.gpu
def test_progressive_files_equal_oracle(oracle, tmp_path):
    shape = (40, 33, 47)
    data = _field(shape, np.float64, 11)
    src = tmp_path / "field.d64"
    data.tofile(src)
    nz, ny, nx = shape
    targets = [1e-2, 1e-4, 1e-6]
    r = run("-i", src, "-d", "-3", nx, ny, nz, "-R", 1e-6, "-z", tmp_path / "c.prisma", "-x", tmp_path / "d.out",
            "-prog", "-errors", f"-{len(targets)}", *targets, "--report", "cr")
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("Total Retrieved size") == len(targets)
    ref, _ = oracle.compress(data, 1e-6, mode=0)
    assert np.array_equal(np.fromfile(tmp_path / "c.prisma", np.uint8), ref)
    p = oracle.Progressive(ref, shape, np.float64, 1e-6)
    for t in targets:
        last, _, _ = p.retrieve(t)
    dec = np.fromfile(tmp_path / "d.out", np.float64).reshape(shape)
    assert np.array_equal(dec.view(np.uint64), last.view(np.uint64))


@pytest.mark.gpu
def test_malformed_stream_is_rejected(oracle, tmp_path):
    """Decompression alone finds the stream at <input>.prisma (the reference's
    naming rule); size-table entries pointing past the stream are reported
    and the run fails instead of reading out of bounds."""
    shape = (20, 24, 28)
    data = _field(shape, np.float32, 5)
    nz, ny, nx = shape
    ref, _ = oracle.compress(data, 1e-3, mode=1)
    ref.tofile(tmp_path / "f.bin.prisma")
    r = run("-i", "f.bin", "-f", "-3", nx, ny, nz, "-R", 1e-3, "-x", cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    dec_o = oracle.decompress(ref, shape, np.float32, 1e-3)
    dec = np.fromfile(tmp_path / "f.bin.prisma.out", np.float32).reshape(shape)
    assert np.array_equal(dec.view(np.uint32), dec_o.view(np.uint32))
    bad = ref.copy()
    bad[8:8 + 2 * 4] = 0xFF  # four size-table entries far beyond the stream
    bad.tofile(tmp_path / "f.bin.prisma")
    r = run("-i", "f.bin", "-f", "-3", nx, ny, nz, "-R", 1e-3, "-x", cwd=tmp_path)
    assert r.returncode != 0
    assert "malformed compressed stream" in r.stderr


DRIVER = os.path.join(ROOT, "prism_amd", "_build", "reference_driver")


@pytest.mark.gpu
@pytest.mark.parametrize("dtype, flag", [(np.float32, "-f"), (np.float64, "-d")])
def test_reference_api_driver_files_equal_oracle(oracle, tmp_path, dtype, flag):
    """A driver written against the reference's C++ API names only (one
    Compressor, one stream reload and one output buffer per progressive
    round, as src/prism.cu does) writes the oracle's stream and, after the
    last round, the oracle's reconstruction; the free findStrategy_h finds
    nothing left to read for the last target."""
    if not os.path.exists(DRIVER):
        pytest.fail("reference_driver not built (run __graft_entry__.build())")
    shape = (29, 34, 41)
    data = _field(shape, dtype, 13)
    src = tmp_path / "field.bin"
    data.tofile(src)
    nz, ny, nx = shape
    targets = [1e-2, 1e-3, 1e-5]
    print("CASE reference driver", dtype.__name__, flush=True)
    r = subprocess.run([DRIVER, "-i", str(src), flag, "-3", str(nx), str(ny), str(nz), "-R", "1e-5", "-z",
                        str(tmp_path / "c.prisma"), "-x", str(tmp_path / "d.out"), "-prog", "-errors",
                        f"-{len(targets)}", *map(str, targets), "--report", "time,cr"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("Total Retrieved size") == len(targets)
    assert r.stdout.count("PSNR") == len(targets) and "extrema disagree" not in r.stdout
    assert "findStrategy_h after the last round: no new planes" in r.stdout
    ref, _ = oracle.compress(data, 1e-5, mode=0)
    assert np.array_equal(np.fromfile(tmp_path / "c.prisma", np.uint8), ref)
    p = oracle.Progressive(ref, shape, dtype, 1e-5)
    for t in targets:
        last, _, _ = p.retrieve(t)
    dec = np.fromfile(tmp_path / "d.out", dtype).reshape(shape)
    assert np.array_equal(dec.view(np.uint8), last.view(np.uint8))
