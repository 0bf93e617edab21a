"""The drop-in command lines on the GPU: preprocess.x + porelb.x reproduce
the reference's V.bin fields and part files bit for bit, keep its density
guard message and exit code, resume from a checkpoint, and postprocess.x
writes the reference's slice format."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, REPO, load_tiny

pytestmark = pytest.mark.gpu
BIN = os.path.join(REPO, "porelb_amd", "bin")


def run(args, cwd, env=None, timeout=120):
    e = dict(os.environ)
    e.update(env or {})
    print("[case]", " ".join(args), env, flush=True)
    return subprocess.run(args, cwd=cwd, capture_output=True, text=True, env=e, timeout=timeout)


def read_parts(cwd, num_p, it):
    parts = [np.fromfile(os.path.join(cwd, f"part_{p:03d}_{it:07d}.bin")).reshape(4, -1) for p in range(num_p)]
    return np.concatenate(parts, axis=1)


@pytest.mark.parametrize("name", ["tiny_a", "tiny_d"])
def test_preprocess_and_porelb_match_reference(tmp_path, name):
    d = load_tiny(name)
    flag = d["flag"]
    nz, ny, nx = flag.shape
    flag.tofile(tmp_path / "flag.bin")
    for np_ in d["ranks"]:
        r = run([os.path.join(BIN, "preprocess.x"), str(np_), str(nx), str(ny), str(nz)], tmp_path)
        assert r.returncode == 0, r.stderr
        vb = O.read_vbin(str(tmp_path / "V.bin"))
        x, y, z, t, pid, nid = O.semantic_fields(vb["V"])
        assert np.array_equal(vb["start"], d[f"np{np_}_start"])
        for k, v in dict(x=x, y=y, z=z, type=t, pid=pid, id=nid).items():
            assert np.array_equal(v, d[f"np{np_}_V_{k}"].astype(np.int64)), k
        steps = int(d["steps"])
        env = {"PLB_TAU": repr(float(d["tau"])), "PLB_GX": repr(float(d["gx"])), "PLB_PRINT": "0"}
        r = run([os.path.join(BIN, "porelb.x"), str(steps)], tmp_path, env)
        assert r.returncode == 0, r.stderr
        assert np.array_equal(read_parts(tmp_path, int(np_), steps), d[f"np{np_}_macros"])
        assert os.path.exists(tmp_path / f"part_{int(np_) - 1:03d}_0000000.bin")  # it % 1000 == 0 save


def test_porelb_under_mpirun(tmp_path):
    """`mpirun -np P ./porelb.x` with V.bin for P ranks: one GPU process does
    the work for every rank's records."""
    if not os.path.exists(O.MPIRUN):
        pytest.skip("no mpirun in this image")
    d = load_tiny("tiny_a")
    flag = d["flag"]
    V, start = O.build_V(flag, 2)
    import porelb_amd as P

    P.write_vbin(str(tmp_path / "V.bin"), V, start, flag.shape)
    env = {"PLB_TAU": repr(float(d["tau"])), "PLB_GX": repr(float(d["gx"])), "PLB_PRINT": "0"}
    r = run([O.MPIRUN, "-np", "2", os.path.join(BIN, "porelb.x"), str(int(d["steps"]))], tmp_path, env)
    assert r.returncode == 0, r.stderr
    assert np.array_equal(read_parts(tmp_path, 2, int(d["steps"])), d["np2_macros"])
    r = run([O.MPIRUN, "-np", "3", os.path.join(BIN, "porelb.x"), "5"], tmp_path, env)
    assert r.returncode != 0 and "differs from num_p 2" in r.stderr


def test_density_guard_cli(tmp_path):
    d = dict(np.load(os.path.join(GOLDEN, "fault.npz")))
    flag = d["flag"]
    flag.tofile(tmp_path / "flag.bin")
    assert run([os.path.join(BIN, "preprocess.x"), "1", "8", "8", "8"], tmp_path).returncode == 0
    r = run([os.path.join(BIN, "porelb.x"), "200"], tmp_path,
            {"PLB_TAU": repr(float(d["tau"])), "PLB_GX": repr(float(d["gx"])), "PLB_PRINT": "0"})
    assert r.returncode == int(d["returncode"]) == 250
    assert str(d["message"]) in r.stderr


def test_checkpoint_resume_and_postprocess(tmp_path):
    flag = (np.random.default_rng(9).random((10, 9, 12)) < 0.6).astype(np.uint8)
    flag.tofile(tmp_path / "flag.bin")
    assert run([os.path.join(BIN, "preprocess.x"), "1", "12", "9", "10"], tmp_path).returncode == 0
    env = {"PLB_TAU": "0.9", "PLB_GX": "1e-4", "PLB_PRINT": "0", "PLB_SAVE": "25"}
    straight = tmp_path / "straight"
    straight.mkdir()
    os.link(tmp_path / "V.bin", straight / "V.bin")
    assert run([os.path.join(BIN, "porelb.x"), "50"], straight, env).returncode == 0
    r = run([os.path.join(BIN, "porelb.x"), "30"], tmp_path, dict(env, PLB_CHECKPOINT="ck.bin"))
    assert r.returncode == 0, r.stderr
    r = run([os.path.join(BIN, "porelb.x"), "50"], tmp_path, dict(env, PLB_RESTART="ck.bin"))
    assert r.returncode == 0, r.stderr
    a = np.fromfile(straight / "part_000_0000050.bin")
    b = np.fromfile(tmp_path / "part_000_0000050.bin")
    assert np.array_equal(a, b)
    # postprocess: X slice at x = 3 of step 50 against the part file scattered by hand
    r = run([os.path.join(BIN, "postprocess.x"), "X", "3", "50"], tmp_path)
    assert r.returncode == 0, r.stderr
    lines = open(tmp_path / "Slice_X_0000050.dat").read().splitlines()
    assert lines[1].startswith("Variables") and "I =    9, J =   10" in lines[2]
    vals = np.array([[float(t) for t in ln.split()] for ln in lines[3:]])
    V = O.read_vbin(str(tmp_path / "V.bin"))["V"]
    mac = b.reshape(4, -1)
    grid = np.zeros((4, 10, 9, 12))
    grid[:, V["z"], V["y"], V["x"]] = mac
    want_rho = grid[0, :, :, 3].reshape(-1)
    assert np.allclose(vals[:, 2], want_rho, rtol=1e-13, atol=0)
    assert np.allclose(vals[:, 3], grid[1, :, :, 3].reshape(-1), rtol=1e-13, atol=1e-300)
    assert np.array_equal(vals[:, 6], flag[:, :, 3].reshape(-1).astype(float))


def test_porelb_autotune_keeps_results(tmp_path):
    """PLB_AUTOTUNE picks a launch shape (cached in the given file) and the
    FP64 part files stay bit-identical to the reference's."""
    d = load_tiny("tiny_d")
    flag = d["flag"]
    nz, ny, nx = flag.shape
    flag.tofile(tmp_path / "flag.bin")
    assert run([os.path.join(BIN, "preprocess.x"), "2", str(nx), str(ny), str(nz)], tmp_path).returncode == 0
    steps = int(d["steps"])
    env = {"PLB_TAU": repr(float(d["tau"])), "PLB_GX": repr(float(d["gx"])), "PLB_PRINT": "0",
           "PLB_AUTOTUNE": str(tmp_path / "tune.txt")}
    for _ in range(2):  # second run reads the cache
        r = run([os.path.join(BIN, "porelb.x"), str(steps)], tmp_path, env, timeout=300)
        assert r.returncode == 0, r.stderr
        assert ">>> autotune:" in r.stdout
        assert np.array_equal(read_parts(tmp_path, 2, steps), d["np2_macros"])
    assert len(open(tmp_path / "tune.txt").read().splitlines()) == 1


def test_postprocess_K_matches_device_reduction(tmp_path, plb):
    """postprocess.x K on porelb.x's part files agrees with the device FP64
    reduction of the same run (summation order differs: 1e-12)."""
    flag = plb.mask_spheres((24, 20, 28), 2.0, 4.0, 0.4, 3)
    nz, ny, nx = flag.shape
    flag.tofile(tmp_path / "flag.bin")
    assert run([os.path.join(BIN, "preprocess.x"), "2", str(nx), str(ny), str(nz)], tmp_path).returncode == 0
    env = {"PLB_TAU": "0.9", "PLB_GX": "1e-5", "PLB_PRINT": "0"}
    assert run([os.path.join(BIN, "porelb.x"), "300"], tmp_path, env).returncode == 0
    r = run([os.path.join(BIN, "postprocess.x"), "K", "300"], tmp_path, env)
    assert r.returncode == 0, r.stderr
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("t= 300")][0].split()
    got = {line[i].rstrip("="): float(line[i + 1]) for i in range(0, len(line), 2)}
    vb = O.read_vbin(str(tmp_path / "V.bin"))
    with plb.Lattice.from_V(vb["V"], vb["start"], vb["dims"], tau=0.9, gx=1e-5) as lat:
        lat.step(0, 300)
        red = lat.reduce()
    assert got["N"] == red["n_nodes"]
    for k in ("sum_ux", "mean_ux", "q", "K"):
        assert got[k] == pytest.approx(red[k], rel=1e-12), k
    assert got["K"] > 0


def test_porelb_fp32_pencil_tile8_close_to_fp64(tmp_path):
    """The chosen forms through the drop-in CLI: FP32 storage, half-way
    bounce-back, TILE8, AoSoA and pencils write part files within FP32 error
    of the FP64 half-way run."""
    flag = (np.random.default_rng(21).random((20, 18, 24)) < 0.62).astype(np.uint8)
    nz, ny, nx = flag.shape
    flag.tofile(tmp_path / "flag.bin")
    assert run([os.path.join(BIN, "preprocess.x"), "2", str(nx), str(ny), str(nz)], tmp_path).returncode == 0
    base = {"PLB_TAU": "0.9", "PLB_GX": "2e-5", "PLB_PRINT": "0", "PLB_BB": "halfway"}
    out = {}
    for name, env in [("fp64", {}), ("fp32", {"PLB_PRECISION": "fp32", "PLB_INDEX": "tile8", "PLB_LAYOUT": "aosoa",
                                              "PLB_ORDER": "pencil", "PLB_BRICK": "24,4,1", "PLB_GRAPH": "8"})]:
        wd = tmp_path / name
        wd.mkdir()
        os.link(tmp_path / "V.bin", wd / "V.bin")
        r = run([os.path.join(BIN, "porelb.x"), "400"], wd, dict(base, **env))
        assert r.returncode == 0, r.stderr
        out[name] = read_parts(wd, 2, 400)
    ux64, ux32 = out["fp64"][1], out["fp32"][1]
    assert np.abs(out["fp32"][0] - out["fp64"][0]).max() < 1e-6
    assert np.linalg.norm(ux32 - ux64) / np.linalg.norm(ux64) < 1e-4


def test_porelb_async_saves_identical(tmp_path):
    """Part files written by the save worker thread (PLB_ASYNC_SAVE=1, the
    default) equal those written inline, at every save step."""
    flag = (np.random.default_rng(31).random((16, 12, 14)) < 0.6).astype(np.uint8)
    nz, ny, nx = flag.shape
    flag.tofile(tmp_path / "flag.bin")
    assert run([os.path.join(BIN, "preprocess.x"), "2", str(nx), str(ny), str(nz)], tmp_path).returncode == 0
    env = {"PLB_TAU": "0.9", "PLB_GX": "1e-5", "PLB_PRINT": "0", "PLB_SAVE": "7"}
    for mode in ("0", "1"):
        wd = tmp_path / f"async{mode}"
        wd.mkdir()
        os.link(tmp_path / "V.bin", wd / "V.bin")
        r = run([os.path.join(BIN, "porelb.x"), "50"], wd, dict(env, PLB_ASYNC_SAVE=mode))
        assert r.returncode == 0, r.stderr
    for it in list(range(0, 50, 7)) + [50]:
        a = read_parts(tmp_path / "async0", 2, it)
        b = read_parts(tmp_path / "async1", 2, it)
        assert np.array_equal(a, b), it


def test_porelb_aa_stream_identical(tmp_path):
    """PLB_STREAM=aa (one population buffer) writes the part files and the
    checkpoint of PLB_STREAM=pull, half-way bounce-back, at every save step."""
    flag = (np.random.default_rng(37).random((16, 12, 14)) < 0.6).astype(np.uint8)
    nz, ny, nx = flag.shape
    flag.tofile(tmp_path / "flag.bin")
    assert run([os.path.join(BIN, "preprocess.x"), "2", str(nx), str(ny), str(nz)], tmp_path).returncode == 0
    env = {"PLB_TAU": "0.9", "PLB_GX": "1e-5", "PLB_PRINT": "0", "PLB_SAVE": "7", "PLB_BB": "halfway",
           "PLB_PRECISION": "fp32", "PLB_CHECKPOINT": "ck.bin"}
    for mode in ("pull", "aa"):
        wd = tmp_path / mode
        wd.mkdir()
        os.link(tmp_path / "V.bin", wd / "V.bin")
        r = run([os.path.join(BIN, "porelb.x"), "45"], wd, dict(env, PLB_STREAM=mode))
        assert r.returncode == 0, r.stderr
    for it in list(range(0, 45, 7)) + [45]:
        assert np.array_equal(read_parts(tmp_path / "pull", 2, it), read_parts(tmp_path / "aa", 2, it)), it
    assert (tmp_path / "pull" / "ck.bin").read_bytes() == (tmp_path / "aa" / "ck.bin").read_bytes()


@pytest.mark.skipif(not os.path.exists("/dev/full"), reason="needs /dev/full")
@pytest.mark.parametrize("mode", ["0", "1"])
def test_porelb_part_file_write_failure(tmp_path, mode):
    """A part file that cannot be written (here /dev/full, which fails every
    write) ends porelb.x with the reference's save_data code -3
    (porelb.c:375-391), from the inline writer and from the worker thread;
    so does a checkpoint that cannot be written."""
    flag = (np.random.default_rng(32).random((8, 6, 7)) < 0.6).astype(np.uint8)
    nz, ny, nx = flag.shape
    flag.tofile(tmp_path / "flag.bin")
    assert run([os.path.join(BIN, "preprocess.x"), "1", str(nx), str(ny), str(nz)], tmp_path).returncode == 0
    env = {"PLB_TAU": "0.9", "PLB_GX": "1e-5", "PLB_PRINT": "0", "PLB_SAVE": "5", "PLB_ASYNC_SAVE": mode}
    (tmp_path / "part_000_0000010.bin").symlink_to("/dev/full")
    r = run([os.path.join(BIN, "porelb.x"), "20"], tmp_path, env)
    assert r.returncode == (-3) & 0xFF and "Macro binary file open error!" in r.stderr, (r.returncode, r.stderr)
    (tmp_path / "part_000_0000010.bin").unlink()
    r = run([os.path.join(BIN, "porelb.x"), "20"], tmp_path, dict(env, PLB_CHECKPOINT="/dev/full"))
    assert r.returncode == (-3) & 0xFF and "cannot write checkpoint" in r.stderr, (r.returncode, r.stderr)


def test_porelb_stops_on_convergence(tmp_path):
    """PLB_TOL / PLB_CHECK: porelb.x stops at the first converged check and its
    final part files equal those of a plain run of that many steps."""
    flag = (np.random.default_rng(33).random((12, 10, 14)) < 0.65).astype(np.uint8)
    nz, ny, nx = flag.shape
    flag.tofile(tmp_path / "flag.bin")
    assert run([os.path.join(BIN, "preprocess.x"), "2", str(nx), str(ny), str(nz)], tmp_path).returncode == 0
    env = {"PLB_TAU": "0.9", "PLB_GX": "1e-5", "PLB_PRINT": "0", "PLB_BB": "halfway", "PLB_SAVE": "100000"}
    wd = tmp_path / "tol"
    wd.mkdir()
    os.link(tmp_path / "V.bin", wd / "V.bin")
    r = run([os.path.join(BIN, "porelb.x"), "50000"], wd, dict(env, PLB_TOL="1e-6", PLB_CHECK="40"))
    assert r.returncode == 0, r.stderr
    line = [ln for ln in r.stdout.splitlines() if ln.startswith(">>> Converged at step")]
    assert len(line) == 1, r.stdout[-500:]
    steps = int(line[0].split()[4].rstrip(":"))
    assert 0 < steps < 50000 and steps % 40 == 0
    wd2 = tmp_path / "plain"
    wd2.mkdir()
    os.link(tmp_path / "V.bin", wd2 / "V.bin")
    r = run([os.path.join(BIN, "porelb.x"), str(steps)], wd2, env)
    assert r.returncode == 0, r.stderr
    assert np.array_equal(read_parts(wd, 2, steps), read_parts(wd2, 2, steps))
