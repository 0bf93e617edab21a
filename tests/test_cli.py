"""Command-line drop-ins: argument and file errors keep the reference's exit
codes (checked before any GPU work, so these run without a GPU)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO

BIN = os.path.join(REPO, "porelb_amd", "bin")


def run(args, cwd, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run(args, cwd=cwd, capture_output=True, text=True, env=e, timeout=60)


def test_porelb_argc_error(tmp_path):
    r = run([os.path.join(BIN, "porelb.x")], tmp_path)
    assert r.returncode == (-10) & 0xFF  # porelb.c:406-410
    assert "usage ./porelb.x it_max" in r.stderr


def test_porelb_missing_vbin(tmp_path):
    r = run([os.path.join(BIN, "porelb.x"), "10"], tmp_path)
    assert r.returncode == (-1) & 0xFF  # porelb.c:18-22
    assert "Node map file openning error!" in r.stderr


def test_porelb_rank_count_mismatch(tmp_path, plb):
    flag = np.ones((4, 4, 4), dtype=np.uint8)
    import oracle as O

    V, start = O.build_V(flag, 2)
    plb.write_vbin(str(tmp_path / "V.bin"), V, start, flag.shape)
    r = run([os.path.join(BIN, "porelb.x"), "10"], tmp_path, env={"PMI_SIZE": "3", "PMI_RANK": "0"})
    assert r.returncode == (-2) & 0xFF  # porelb.c:29-34
    assert "differs from num_p 2" in r.stderr


def test_preprocess_errors(tmp_path):
    r = run([os.path.join(BIN, "preprocess.x")], tmp_path)
    assert r.returncode == (-1) & 0xFF
    r = run([os.path.join(BIN, "preprocess.x"), "1", "4", "4", "4"], tmp_path)
    assert r.returncode == (-1) & 0xFF and "flag data file opening error!" in r.stderr
    (tmp_path / "flag.bin").write_bytes(b"\x01" * 10)
    r = run([os.path.join(BIN, "preprocess.x"), "1", "4", "4", "4"], tmp_path)
    assert r.returncode == (-2) & 0xFF and "reading error" in r.stderr


def test_mkflag_writes_mask(tmp_path, plb):
    r = run([os.path.join(BIN, "mkflag.x"), "noise", "16", "12", "8", "2.0", "0.4", "5", "m.bin"], tmp_path)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(tmp_path / "m.bin", dtype=np.uint8).reshape(8, 12, 16)
    assert np.array_equal(got, plb.mask_noise((8, 12, 16), 2.0, 0.4, 5))
