"""postprocess.x outputs against an independent scatter of the part files.

The part files and V.bin here come from the oracle (oracle build_V and the
push restatement of porelb.c, several ranks), so these tests need no GPU.
Every output line is rebuilt in Python from the scattered grid with the
reference's own format strings (postprocess.c:19-85: title, variable and
zone lines, then "%f %f % 21.14e  % 21.14E  % 21.14E  % 21.14E  % 8.1e" per
point) and compared as text; the full field (the `A` mode, an ASCII Tecplot
BLOCK zone in place of TECIO's binary file, postprocess.c:216-281) is
checked value by value at "%.14e"."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import REPO

BIN = os.path.join(REPO, "porelb_amd", "bin")
ROW = "%f %f % 21.14e  % 21.14E  % 21.14E  % 21.14E  % 8.1e\n"


def run(args, cwd):
    return subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=60)


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    """A 3-rank V.bin and part files of two time steps of the oracle's push
    run (porelb.c:426-437) on a random 11 x 9 x 12 mask."""
    wd = tmp_path_factory.mktemp("post")
    flag = (np.random.default_rng(41).random((12, 9, 11)) < 0.6).astype(np.uint8)
    nz, ny, nx = flag.shape
    V, start = O.build_V(flag, 3)
    flag.tofile(wd / "flag.bin")
    import porelb_amd as P

    P.write_vbin(str(wd / "V.bin"), V, start, flag.shape)
    f0, f1 = O.init_equilibrium(len(V))
    macros = {}
    for t, n in ((7, 7), (20, 13)):
        macros[t] = O.push_run(V, start, f0, f1, n, 0.9, 1e-4, it0=t - n)
        ends = list(start[1:]) + [len(V)]
        for p, (b, e) in enumerate(zip(start, ends)):
            np.ascontiguousarray(macros[t][:, b:e]).tofile(wd / f"part_{p:03d}_{t:07d}.bin")
    grids = {}
    for t, m in macros.items():
        g = np.zeros((4, nz, ny, nx))
        g[:, V["z"], V["y"], V["x"]] = m
        grids[t] = g
    return dict(wd=wd, flag=flag, V=V, grids=grids)


def expected_slice(case, axis, n, t):
    flag, g = case["flag"], case["grids"][t]
    nz, ny, nx = flag.shape
    title = {"X": "X", "Y": "Y", "Z": "Y"}[axis]          # the Z slice's title says Y (postprocess.c:70)
    zone = (ny, nz) if axis == "X" else (nz, nx)           # Y and Z both print nz, nx (postprocess.c:51, 71)
    out = ['TITLE = " %s slice at %04d"\n' % (title, n),
           'Variables = "X", "Y", "Rho", "Ux", "Uy", "Uz", "flag"\n',
           "ZONE I = %4d, J = %4d, F=POINT\n" % zone]
    if axis == "X":
        pts = [(k, j, (k, j, n)) for k in range(nz) for j in range(ny)]
    elif axis == "Y":
        pts = [(k, i, (k, n, i)) for k in range(nz) for i in range(nx)]
    else:
        pts = [(j, i, (n, j, i)) for j in range(ny) for i in range(nx)]
    for a, b, (z, y, x) in pts:
        out.append(ROW % (a, b, g[0, z, y, x], g[1, z, y, x], g[2, z, y, x], g[3, z, y, x], float(flag[z, y, x])))
    return "".join(out)


@pytest.mark.parametrize("axis,n", [("X", 0), ("X", 10), ("Y", 4), ("Y", 8), ("Z", 0), ("Z", 11)])
def test_slices_match_scatter(case, axis, n):
    wd = case["wd"]
    r = run([os.path.join(BIN, "postprocess.x"), axis, str(n), "7", "20"], wd)
    assert r.returncode == 0, r.stderr
    assert r.stdout.splitlines()[0] == "argc = 5"
    assert "Total 2 time seriese data" in r.stdout
    for t in (7, 20):
        got = open(wd / f"Slice_{axis}_{t:07d}.dat").read()
        assert got == expected_slice(case, axis, n, t), (axis, n, t)


def test_full_field_matches_scatter(case):
    wd, flag = case["wd"], case["flag"]
    nz, ny, nx = flag.shape
    r = run([os.path.join(BIN, "postprocess.x"), "A", "20"], wd)
    assert r.returncode == 0, r.stderr
    lines = open(wd / "full_0000020.dat").read().splitlines()
    assert lines[0] == 'TITLE = "SIMPLE DATASET"'
    assert lines[1] == 'VARIABLES = "X" "Y" "Z" "rho" "vx" "vy" "vz" "flag"'   # TECINI112 names
    assert lines[2] == f'ZONE T="Simple Zone", I={nx}, J={ny}, K={nz}, DATAPACKING=BLOCK, SOLUTIONTIME=20'
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    g = case["grids"][20]
    blocks = [x, y, z, g[0], g[1], g[2], g[3], flag]           # X Y Z rho vx vy vz flag, x fastest
    want = ["%.14e" % v for b in blocks for v in np.asarray(b, dtype=float).reshape(-1)]
    assert lines[3:] == want
    # the solid voxels stay 0 and every pore voxel carries its node's macros
    assert np.count_nonzero(g[0]) == len(case["V"])
