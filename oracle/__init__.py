"""CPU oracle for porelb's sparse D3Q19 time step -- TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``oracle/liboracle.so`` (a plain-C restatement of the
reference, see ``plb_oracle.h``) plus helpers that run the reference programs
built from the reference's own sources into ``oracle/_ref``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
this package, and only as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_PREPROCESS = os.path.join(REF_DIR, "ref_preprocess.x")
REF_PORELB = os.path.join(REF_DIR, "ref_porelb.x")
MPIRUN = "/opt/conda/bin/mpirun"

Q = 19
FLUID, SOLID, EXTERN = 0, 1, 2
TAB_BOUNCE, TAB_CONST = -1, -2
BB_REFCOMPAT, BB_HALFWAY = 0, 1

# NEIGHBOR_INFO / NODE_INFO (porelb.h:58-76): 8-byte entries, 152-byte records.
NB_DTYPE = np.dtype([("node_type", "u1"), ("_pad", "u1"), ("node_pid", "<u2"), ("node_id", "<u4")])
NODE_DTYPE = np.dtype([("x", "<u2"), ("y", "<u2"), ("z", "<u2"), ("_pad", "<u2"), ("nb", NB_DTYPE, (18,))])
assert NB_DTYPE.itemsize == 8 and NODE_DTYPE.itemsize == 152

E = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [0, 1, 1], [0, -1, -1], [0, -1, 1], [0, 1, -1], [-1, 0, -1], [1, 0, 1], [-1, 0, 1],
              [1, 0, -1], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 1, 0]], dtype=np.int64)
TP = np.array([1.0 / 3] + [1.0 / 18] * 6 + [1.0 / 36] * 12)
RE = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        p = C.c_void_p
        i64, i32, d = C.c_int64, C.c_int, C.c_double
        L.orc_count_pores.argtypes = [p, i32, i32, i32]
        L.orc_count_pores.restype = i64
        L.orc_build_V.argtypes = [p, i32, i32, i32, i32, p, p]
        L.orc_pull_table.argtypes = [p, i64, p, i32, i32, p]
        L.orc_init_equilibrium.argtypes = [i64, d, p, p]
        L.orc_init_equilibrium.restype = None
        L.orc_push_run.argtypes = [p, i64, p, i32, i32, d, d, i64, i64, p, p, p, p, p]
        L.orc_pull_run.argtypes = [p, i64, d, d, d, i64, i64, p, p, p, p, p]
        L.orc_reduce.argtypes = [p, p, i64, i32, i32, i32, d, d, p]
        L.orc_reduce.restype = None
        L.orc_getK_bcc.argtypes = [d, i32, p]
        L.orc_getK_bcc.restype = None
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def build_V(flag: np.ndarray, num_p: int):
    """Restated build_V (preprocess.c:29-164). flag: uint8 [nz][ny][nx]."""
    flag = np.ascontiguousarray(flag, dtype=np.uint8)
    nz, ny, nx = flag.shape
    n = lib().orc_count_pores(_ptr(flag), nx, ny, nz)
    V = np.zeros(n, dtype=NODE_DTYPE)
    start = np.zeros(num_p, dtype=np.int32)
    rc = lib().orc_build_V(_ptr(flag), nx, ny, nz, num_p, _ptr(V), _ptr(start))
    if rc != 0:
        raise ValueError(f"orc_build_V failed: {rc}")
    return V, start


def pull_table(V, start, mode=BB_REFCOMPAT):
    N = len(V)
    tab = np.zeros((18, N), dtype=np.int32)
    lib().orc_pull_table(_ptr(V), N, _ptr(start), len(start), mode, _ptr(tab))
    return tab


def init_equilibrium(N, rho0=1.0):
    f0 = np.zeros((N, Q))
    f1 = np.zeros((N, Q))
    lib().orc_init_equilibrium(N, rho0, _ptr(f0), _ptr(f1))
    return f0, f1


class Fault(RuntimeError):
    def __init__(self, node, step):
        super().__init__(f"density guard at node {node}, step {step}")
        self.node, self.step = node, step


def push_run(V, start, f0, f1, nsteps, tau, gx, it0=0, mode=BB_REFCOMPAT):
    """Push restatement of porelb.c:426-437; f0/f1 AoS updated in place.
    Returns macros [4][N] of the last step."""
    N = len(V)
    mac = np.zeros((4, N))
    fn, fs = C.c_int64(-1), C.c_int64(-1)
    rc = lib().orc_push_run(_ptr(V), N, _ptr(start), len(start), mode, tau, gx, it0, nsteps,
                            _ptr(f0), _ptr(f1), _ptr(mac), C.byref(fn), C.byref(fs))
    if rc == -6:
        raise Fault(fn.value, fs.value)
    return mac


def pull_run(tab, f0, nsteps, tau, gx, it0=0, rho0=1.0):
    N = f0.shape[0]
    scratch = np.zeros_like(f0)
    mac = np.zeros((4, N))
    fn, fs = C.c_int64(-1), C.c_int64(-1)
    rc = lib().orc_pull_run(_ptr(tab), N, tau, gx, rho0, it0, nsteps, _ptr(f0), _ptr(scratch),
                            _ptr(mac), C.byref(fn), C.byref(fs))
    if rc == -6:
        raise Fault(fn.value, fs.value)
    return mac


def reduce(f0, ux, dims, tau, gx):
    nz, ny, nx = dims
    out = np.zeros(5)
    lib().orc_reduce(_ptr(f0), _ptr(np.ascontiguousarray(ux)), f0.shape[0], nx, ny, nz, tau, gx,
                     _ptr(out))
    return dict(mass=out[0], sum_ux=out[1], mean_ux=out[2], q=out[3], K=out[4])


def getK_bcc(chi, nx):
    out = np.zeros(4)
    lib().orc_getK_bcc(chi, nx, _ptr(out))
    return dict(K1=out[0], K2=out[1], eps=out[2], a=out[3])


# ---------------------------------------------------------------------------
# V.bin / part-file readers (R4 / R11 layouts) for checking outputs.

def read_vbin(path):
    with open(path, "rb") as fh:
        hdr = np.frombuffer(fh.read(20), dtype="<i4")
        num_node, num_p, nx, ny, nz = (int(v) for v in hdr)
        start = np.frombuffer(fh.read(4 * num_p), dtype="<i4").copy()
        V = np.frombuffer(fh.read(152 * num_node), dtype=NODE_DTYPE).copy()
    return dict(num_node=num_node, num_p=num_p, dims=(nz, ny, nx), start=start, V=V)


def semantic_fields(V):
    """The V.bin fields the reference defines: coordinates, neighbour types,
    and pid/id of non-SOLID neighbours (SOLID pid/id and padding are left
    uninitialised by preprocess.c:120-151)."""
    t = V["nb"]["node_type"].astype(np.int64)
    live = t != SOLID
    return (V["x"].astype(np.int64), V["y"].astype(np.int64), V["z"].astype(np.int64), t,
            np.where(live, V["nb"]["node_pid"], 0).astype(np.int64),
            np.where(live, V["nb"]["node_id"], 0).astype(np.int64))


def have_reference_build():
    return os.path.exists(REF_PREPROCESS) and os.path.exists(REF_PORELB)


def run_reference(flag, num_p, nsteps, tau, gx, save_every=-1, print_steps=0, workdir=None,
                  dump=True, timeout=600):
    """Run the reference preprocess + porelb (built from its own sources) on
    `flag` with `num_p` MPI ranks.  Returns V.bin contents, per-rank f0/f1
    dumps concatenated in V order, the final part-file macros and the parsed
    timing line."""
    flag = np.ascontiguousarray(flag, dtype=np.uint8)
    nz, ny, nx = flag.shape
    own = workdir is None
    wd = workdir or tempfile.mkdtemp(prefix="plbref_")
    try:
        flag.tofile(os.path.join(wd, "flag.bin"))
        subprocess.run([REF_PREPROCESS, str(num_p), str(nx), str(ny), str(nz), "flag.bin", "V.bin"],
                       cwd=wd, check=True, capture_output=True, timeout=timeout)
        args = [str(nsteps), repr(float(tau)), repr(float(gx)), str(save_every), str(print_steps)]
        if dump:
            args.append("dump")
        cmd = [MPIRUN, "-np", str(num_p), REF_PORELB] + args
        r = subprocess.run(cmd, cwd=wd, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed rc={r.returncode}: {r.stderr[-2000:]}")
        vb = read_vbin(os.path.join(wd, "V.bin"))
        res = dict(vbin=vb, stdout=r.stdout)
        for line in r.stdout.splitlines():
            if line.startswith("PLB_REF_RESULT"):
                res["timing"] = {k: float(v) for k, v in (t.split("=") for t in line.split()[1:])}
        if dump:
            f0s, f1s = [], []
            for p in range(num_p):
                f0s.append(np.fromfile(os.path.join(wd, f"dump_f0_{p:03d}.bin")).reshape(-1, Q))
                f1s.append(np.fromfile(os.path.join(wd, f"dump_f1_{p:03d}.bin")).reshape(-1, Q))
            res["f0"] = np.concatenate(f0s)
            res["f1"] = np.concatenate(f1s)
        if save_every >= 0:
            parts = []
            for p in range(num_p):
                a = np.fromfile(os.path.join(wd, f"part_{p:03d}_{nsteps:07d}.bin"))
                parts.append(a.reshape(4, -1))
            res["macros"] = np.concatenate(parts, axis=1)
        return res
    finally:
        if own:
            shutil.rmtree(wd, ignore_errors=True)
