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


def test_getK_matches_oracle_and_table(tmp_path):
    """getK.x prints the Sangani-Acrivos K1 and Kozeny-Carman K2 of getK.c:39-72
    (pinned through the oracle's BCC tests); the table block carries the extra
    porosity factor of getK.c:85."""
    import oracle as O

    for chi, nx in [(0.2, 33), (0.5, 65), (0.85, 129)]:
        r = run([os.path.join(BIN, "getK.x"), repr(chi), str(nx)], tmp_path)
        assert r.returncode == 0, r.stderr
        got = dict(line.split("=") for line in r.stdout.split())
        want = O.getK_bcc(chi, nx)
        assert float(got["K1"]) == pytest.approx(want["K1"], rel=1e-6)
        assert float(got["K2"]) == pytest.approx(want["K2"], rel=1e-6)
    r = run([os.path.join(BIN, "getK.x"), "0.5", "65", "table"], tmp_path)
    rows = [list(map(float, ln.split("\t"))) for ln in r.stdout.splitlines()[3:]]
    assert len(rows) == 31
    for c, eps, k1, k2 in rows[1:]:
        want = O.getK_bcc(c, 65)
        assert eps == pytest.approx(want["eps"], rel=1e-6)
        assert k1 == pytest.approx(want["K1"] * want["eps"], rel=1e-6)
        assert k2 == pytest.approx(want["K2"], rel=1e-6)
    r = run([os.path.join(BIN, "getK.x"), "0.5"], tmp_path)
    assert r.returncode == (-1) & 0xFF and "argc err" in r.stderr


def test_postprocess_permeability_report(tmp_path, plb):
    """postprocess.x K sums the part files in V.bin order in FP64 and reports
    <u_x>, q = sum u_x / (nx ny nz) and K = (tau - 1/2)/3 * q / g_x."""
    import oracle as O

    flag = (np.random.default_rng(4).random((6, 5, 7)) < 0.55).astype(np.uint8)
    V, start = O.build_V(flag, 2)
    plb.write_vbin(str(tmp_path / "V.bin"), V, start, flag.shape)
    rng = np.random.default_rng(5)
    bounds = list(start) + [len(V)]
    parts = []
    for p in range(2):
        n = bounds[p + 1] - bounds[p]
        m = rng.standard_normal((4, n)) * 1e-3
        m[0] += 1.0
        m.tofile(tmp_path / f"part_{p:03d}_0000070.bin")
        parts.append(m)
    r = run([os.path.join(BIN, "postprocess.x"), "K", "70"], tmp_path, {"PLB_TAU": "0.9", "PLB_GX": "2e-5"})
    assert r.returncode == 0, r.stderr
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("t= 70")][0].split()
    got = {line[i].rstrip("="): line[i + 1] for i in range(0, len(line), 2)}
    s_rho = s_ux = 0.0
    for m in parts:             # sequential order, as the C loop adds
        for a, b in zip(m[0], m[1]):
            s_rho += float(a)
            s_ux += float(b)
    assert int(got["N"]) == len(V)
    assert float(got["sum_rho"]) == s_rho and float(got["sum_ux"]) == s_ux
    assert float(got["mean_ux"]) == s_ux / len(V)
    q = s_ux / flag.size
    assert float(got["q"]) == q
    assert float(got["K"]) == 1.0 / 3.0 * (0.9 - 0.5) * q / 2e-5


def test_bench_refuses_several_gpus(tmp_path):
    """The step runs on one GPU; asking bench.py for more exits before any GPU work."""
    import sys

    r = run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], tmp_path)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_postprocess_errors(tmp_path, plb):
    """postprocess.c:92-112: bad mode, missing V.bin, slice out of range."""
    r = run([os.path.join(BIN, "postprocess.x")], tmp_path)
    assert r.returncode == (-2) & 0xFF and "Usage" in r.stdout
    r = run([os.path.join(BIN, "postprocess.x"), "Q", "1"], tmp_path)
    assert r.returncode == (-3) & 0xFF and "Not implemented yet!" in r.stderr
    r = run([os.path.join(BIN, "postprocess.x"), "K", "10"], tmp_path)
    assert r.returncode == (-1) & 0xFF and "Node map file openning error!" in r.stderr
    import oracle as O

    flag = np.ones((3, 4, 5), dtype=np.uint8)
    V, start = O.build_V(flag, 1)
    plb.write_vbin(str(tmp_path / "V.bin"), V, start, flag.shape)
    flag.tofile(tmp_path / "flag.bin")
    r = run([os.path.join(BIN, "postprocess.x"), "X", "5", "10"], tmp_path)
    assert r.returncode == (-3) & 0xFF and "out of range" in r.stderr
    r = run([os.path.join(BIN, "postprocess.x"), "K", "10"], tmp_path)  # part file missing
    assert r.returncode == (-1) & 0xFF and "Part file openning error!" in r.stderr


@pytest.mark.skipif(not os.path.exists("/dev/full"), reason="needs /dev/full")
def test_postprocess_full_disk_is_an_error(tmp_path, plb):
    """An output that cannot be written (here /dev/full, which fails every
    write with ENOSPC) ends postprocess.x with -3 instead of a truncated file."""
    import oracle as O

    flag = np.ones((3, 4, 5), dtype=np.uint8)
    V, start = O.build_V(flag, 1)
    plb.write_vbin(str(tmp_path / "V.bin"), V, start, flag.shape)
    flag.tofile(tmp_path / "flag.bin")
    np.ones((4, len(V))).tofile(tmp_path / "part_000_0000010.bin")
    r = run([os.path.join(BIN, "postprocess.x"), "X", "1", "10"], tmp_path)
    assert r.returncode == 0, r.stderr
    for args, name, text in [(["X", "1", "10"], "Slice_X_0000010.dat", "Slice file write error!"),
                             (["A", "10"], "full_0000010.dat", "Full file write error!")]:
        out = tmp_path / name
        if out.exists():
            out.unlink()
        out.symlink_to("/dev/full")
        r = run([os.path.join(BIN, "postprocess.x")] + args, tmp_path)
        assert r.returncode == (-3) & 0xFF and text in r.stderr, (args, r.returncode, r.stderr)


def test_porelb_bad_environment_values(tmp_path, plb):
    """Unknown values of the configuration variables stop porelb.x with -10
    before any GPU work."""
    import oracle as O

    flag = np.ones((4, 4, 4), dtype=np.uint8)
    V, start = O.build_V(flag, 1)
    plb.write_vbin(str(tmp_path / "V.bin"), V, start, flag.shape)
    for env, text in [({"PLB_PRECISION": "fp8"}, "bad value PLB_PRECISION=fp8"),
                      ({"PLB_INDEX": "tile4"}, "bad value PLB_INDEX=tile4"),
                      ({"PLB_BRICK": "4,4"}, "bad value PLB_BRICK=4,4"),
                      ({"PLB_SAVE": "0"}, "PLB_SAVE must be >= 1"),
                      ({"PLB_TOL": "-1"}, "PLB_TOL must be >= 0"),
                      ({"PLB_TOL": "1e-6", "PLB_CHECK": "0"}, "PLB_CHECK >= 1")]:
        r = run([os.path.join(BIN, "porelb.x"), "10"], tmp_path, env=env)
        assert r.returncode != 0 and text in r.stderr, (env, r.stderr)
    # values the library's configuration check rejects (before any device call)
    for env in ({"PLB_WAVES": "9"}, {"PLB_WAVES": "-1"}, {"PLB_BLOCK": "100"}, {"PLB_GRAPH": "3"}):
        r = run([os.path.join(BIN, "porelb.x"), "10"], tmp_path, env=env)
        assert r.returncode == (-20) & 0xFF and "libplb: invalid argument" in r.stderr, (env, r.stderr)


def test_bench_workloads_are_complete():
    """Every bench workload carries the keys the benchmark line reads (checked
    without a GPU, so a missing key fails here rather than on the GPU)."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(REPO, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    need = {"kind", "dims", "tau", "gx", "precision", "bb", "index", "layout", "ordering", "brick", "block",
            "label", "seed", "porosity"}
    for name, w in bench.WORKLOADS.items():
        assert need <= set(w), (name, need - set(w))
        assert w["ordering"] != 2 or w["brick"][0] == w["dims"][2], name  # pencils span whole rows
    assert set(bench.FP64_PLAIN) >= {"precision", "bb", "index", "layout", "ordering", "brick", "block", "waves"}
