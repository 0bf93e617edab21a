"""porelb_amd -- MI355X (gfx950) implementation of porelb's sparse D3Q19 time step.

Python mirror of the C ABI in ``include/plb.h`` (ctypes over the in-tree
``porelb_amd/libplb.so``).  The C command lines in ``porelb_amd/bin`` are
the drop-in for the reference's ``preprocess.x`` / ``porelb.x``; this module
serves tests, the benchmark and scripting.

There is no CPU fallback: if ``libplb.so`` is missing or no GPU is visible,
every compute call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
# Always the in-tree build: the library beside this file.
LIB_PATH = os.path.join(HERE, "libplb.so")
HEADER = os.path.join(REPO, "include", "plb.h")
HEADERS = [HEADER, os.path.join(REPO, "include", "plb_tune.h")]
BIN_DIR = os.path.join(HERE, "bin")

Q = 19
FP64, FP32, FP16 = 0, 1, 2
BB_REFCOMPAT, BB_HALFWAY = 0, 1
LAYOUT_SOA, LAYOUT_AOSOA = 0, 1
ORDER_REFERENCE, ORDER_BRICK, ORDER_PENCIL = 0, 1, 2
INDEX_I32, INDEX_TILE8, INDEX_BITMAP = 0, 1, 2
ROUND_NEAREST, ROUND_STOCHASTIC = 0, 1
STREAM_PULL, STREAM_AA = 0, 1
NOALLOC_MAX_NODES = 1 << 20     # PLB_NOALLOC_MAX_NODES
TYPE_FLUID, TYPE_SOLID, TYPE_EXTERN = 0, 1, 2
ERR_DENSITY = -6

# V.bin record (porelb.h:58-76): 8-byte neighbour entries, 152-byte nodes.
NB_DTYPE = np.dtype([("node_type", "u1"), ("_pad", "u1"), ("node_pid", "<u2"), ("node_id", "<u4")])
NODE_DTYPE = np.dtype([("x", "<u2"), ("y", "<u2"), ("z", "<u2"), ("_pad", "<u2"), ("nb", NB_DTYPE, (18,))])
assert NODE_DTYPE.itemsize == 152


class PlbError(RuntimeError):
    def __init__(self, code, what=""):
        msg = _lib().plb_strerror(code).decode() if _LIB is not None else str(code)
        super().__init__(f"{what}: libplb error {code} ({msg})")
        self.code = code


class Config(C.Structure):
    _fields_ = [("tau", C.c_double), ("gx", C.c_double), ("rho0", C.c_double),
                ("precision", C.c_int), ("bb_mode", C.c_int), ("layout", C.c_int),
                ("ordering", C.c_int), ("brick", C.c_int * 3), ("block_size", C.c_int),
                ("device", C.c_int), ("fp16_scale_log2", C.c_int), ("index_kind", C.c_int),
                ("nodes_per_thread", C.c_int), ("graph_steps", C.c_int), ("waves_per_simd", C.c_int),
                ("fp16_rounding", C.c_int), ("placement_reserve_gib", C.c_int),
                ("stream_mode", C.c_int)]


class Reduction(C.Structure):
    _fields_ = [("mass", C.c_double), ("sum_ux", C.c_double), ("mean_ux", C.c_double),
                ("q", C.c_double), ("K", C.c_double), ("n_nodes", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Fault(C.Structure):
    _fields_ = [("tripped", C.c_int), ("step", C.c_int64), ("node", C.c_int64), ("pid", C.c_int),
                ("local_id", C.c_int64), ("x", C.c_int), ("y", C.c_int), ("z", C.c_int),
                ("rho", C.c_double)]


class StateInfo(C.Structure):
    _fields_ = [("precision", C.c_int), ("elem_bytes", C.c_int), ("fp16_scale_log2", C.c_int),
                ("rho0", C.c_double), ("n_nodes", C.c_int64)]


class Sample(C.Structure):
    _fields_ = [("flag", C.c_void_p), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("tau", C.c_double), ("gx", C.c_double)]


_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C porelb_amd` "
                               "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        p, i32, i64, d = C.c_void_p, C.c_int, C.c_int64, C.c_double
        sig = {
            "plb_config_default": ([p], None),
            "plb_create": ([p, p, p, i64, p, i32, i32, i32, i32], i32),
            "plb_create_from_flag": ([p, p, p, i32, i32, i32, i32], i32),
            "plb_destroy": ([p], None),
            "plb_init_equilibrium": ([p], i32),
            "plb_step": ([p, i64, i64], i32),
            "plb_step_async": ([p, i64, i64], i32),
            "plb_sync": ([p], i32),
            "plb_step_timed": ([p, i64, i64, p], i32),
            "plb_read_macros": ([p, p, p, p, p], i32),
            "plb_read_populations": ([p, p], i32),
            "plb_write_populations": ([p, p], i32),
            "plb_read_table": ([p, p], i32),
            "plb_read_macros_noalloc": ([p, i64, i64, p, p, p, p], i32),
            "plb_get_state_info": ([p, p], i32),
            "plb_read_state": ([p, p], i32),
            "plb_write_state": ([p, p], i32),
            "plb_num_nodes": ([p], i64),
            "plb_reduce": ([p, p], i32),
            "plb_last_fault": ([p, p], i32),
            "plb_bytes_per_update": ([p], d),
            "plb_strerror": ([i32], C.c_char_p),
            "plb_host_alloc": ([C.c_size_t], C.c_void_p),
            "plb_host_free": ([C.c_void_p], None),
            "plb_preprocess": ([p, i32, i32, i32, i32, i32, p, p, p], i32),
            "plb_batch_create": ([p, p, p, i32], i32),
            "plb_batch_run": ([p, i64, i64, d, p, p, p], i32),
            "plb_batch_run_persistent": ([p, i64, i64, d, p, p, p], i32),
            "plb_batch_reduce": ([p, p], i32),
            "plb_run_to_convergence": ([p, i64, i64, i64, d, p, p], i32),
            "plb_batch_read_macros": ([p, i32, p, p, p, p], i32),
            "plb_batch_num_nodes": ([p, i32], i64),
            "plb_batch_last_fault": ([p, p, p], i32),
            "plb_batch_destroy": ([p], None),
            "plb_mask_noise": ([p, i32, i32, i32, d, d, C.c_uint64], i32),
            "plb_mask_spheres": ([p, i32, i32, i32, d, d, d, C.c_uint64], i32),
            "plb_mask_bcc": ([p, i32, d], i32),
            "plb_copy_bandwidth": ([i32, i64, i32, p], i32),
            "plb_autotune": ([p, p, i32, i32, i32, i32, i32, C.c_char_p, p, p, i32, p], i32),
            "plb_batch_autotune": ([p, p, i32, i32, i32, p, p, i32, p], i32),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def lib():
    return _lib()


def header_symbols(paths=None):
    """Names of the functions include/plb.h and include/plb_tune.h declare."""
    names = set()
    for path in paths or HEADERS:
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(plb_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc, what):
    if rc != 0:
        raise PlbError(rc, what)


def default_config(**kw) -> Config:
    cfg = Config()
    _lib().plb_config_default(C.byref(cfg))
    for k, v in kw.items():
        if k == "brick":
            cfg.brick[0], cfg.brick[1], cfg.brick[2] = v
        else:
            setattr(cfg, k, v)
    return cfg


# ------------------------------------------------------------------ masks
def mask_noise(dims, sigma, porosity, seed):
    nz, ny, nx = dims
    f = np.zeros(dims, dtype=np.uint8)
    _check(_lib().plb_mask_noise(_ptr(f), nx, ny, nz, sigma, porosity, seed), "plb_mask_noise")
    return f


def mask_spheres(dims, rmin, rmax, porosity, seed):
    nz, ny, nx = dims
    f = np.zeros(dims, dtype=np.uint8)
    _check(_lib().plb_mask_spheres(_ptr(f), nx, ny, nz, rmin, rmax, porosity, seed), "plb_mask_spheres")
    return f


def mask_bcc(n, chi):
    f = np.zeros((n, n, n), dtype=np.uint8)
    _check(_lib().plb_mask_bcc(_ptr(f), n, chi), "plb_mask_bcc")
    return f


# ------------------------------------------------------------------ preprocess
def preprocess(flag, num_p, device=0):
    """Device build_V (preprocess.c:29-164): returns (V records, V_sub_start)."""
    flag = np.ascontiguousarray(flag, dtype=np.uint8)
    nz, ny, nx = flag.shape
    n = C.c_int64(0)
    start = np.zeros(num_p, dtype=np.int32)
    _check(_lib().plb_preprocess(_ptr(flag), nx, ny, nz, num_p, device, None, None, C.byref(n)),
           "plb_preprocess(count)")
    V = np.zeros(n.value, dtype=NODE_DTYPE)
    _check(_lib().plb_preprocess(_ptr(flag), nx, ny, nz, num_p, device, _ptr(V), _ptr(start),
                                 C.byref(n)), "plb_preprocess")
    return V, start


def host_array(shape, dtype=np.float64):
    """A numpy array in page-locked host memory (plb_host_alloc), freed when
    the array is garbage-collected; read-backs into it run at the PCIe rate."""
    import weakref

    nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
    p = _lib().plb_host_alloc(nbytes)
    if not p:
        raise MemoryError(f"plb_host_alloc({nbytes}) failed")
    raw = (C.c_char * max(nbytes, 1)).from_address(p)
    weakref.finalize(raw, _lib().plb_host_free, p)
    return np.frombuffer(raw, dtype=dtype, count=int(np.prod(shape))).reshape(shape)


# ------------------------------------------------------------------ lattice
class Lattice:
    """One pore-node lattice resident on the GPU (a plb_ctx)."""

    def __init__(self, handle, dims, cfg):
        self._h = handle
        self.dims = dims
        self.cfg = cfg
        self.N = int(_lib().plb_num_nodes(self._h))

    @classmethod
    def from_V(cls, V, start, dims, cfg=None, **kw):
        cfg = cfg or default_config(**kw)
        nz, ny, nx = dims
        V = np.ascontiguousarray(V, dtype=NODE_DTYPE)
        start = np.ascontiguousarray(start, dtype=np.int32)
        h = C.c_void_p()
        _check(_lib().plb_create(C.byref(h), C.byref(cfg), _ptr(V), len(V), _ptr(start), len(start),
                                 nx, ny, nz), "plb_create")
        return cls(h, dims, cfg)

    @classmethod
    def from_flag(cls, flag, num_p=1, cfg=None, **kw):
        cfg = cfg or default_config(**kw)
        flag = np.ascontiguousarray(flag, dtype=np.uint8)
        nz, ny, nx = flag.shape
        h = C.c_void_p()
        _check(_lib().plb_create_from_flag(C.byref(h), C.byref(cfg), _ptr(flag), nx, ny, nz, num_p),
               "plb_create_from_flag")
        return cls(h, flag.shape, cfg)

    def close(self):
        if self._h:
            _lib().plb_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_equilibrium(self):
        _check(_lib().plb_init_equilibrium(self._h), "plb_init_equilibrium")

    def step(self, it0, nsteps):
        """Returns 0 or raises; a density-guard trip raises PlbError(code=-6)."""
        _check(_lib().plb_step(self._h, it0, nsteps), "plb_step")

    def step_rc(self, it0, nsteps):
        return _lib().plb_step(self._h, it0, nsteps)

    def step_timed(self, it0, nsteps):
        ms = C.c_double(0.0)
        _check(_lib().plb_step_timed(self._h, it0, nsteps, C.byref(ms)), "plb_step_timed")
        return ms.value

    def read_macros(self, out=None):
        """(4, N) rho, ux, uy, uz in V order; `out` may be a host_array."""
        if out is None:
            out = np.zeros((4, self.N))
        assert out.shape == (4, self.N) and out.dtype == np.float64 and out.flags.c_contiguous
        _check(_lib().plb_read_macros(self._h, _ptr(out[0]), _ptr(out[1]), _ptr(out[2]), _ptr(out[3])),
               "plb_read_macros")
        return out

    def read_populations(self):
        f = np.zeros((self.N, Q))
        _check(_lib().plb_read_populations(self._h, _ptr(f)), "plb_read_populations")
        return f

    def write_populations(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == (self.N, Q)
        _check(_lib().plb_write_populations(self._h, _ptr(f)), "plb_write_populations")

    STATE_DTYPE = {FP64: np.float64, FP32: np.float32, FP16: np.uint16}

    def state_info(self):
        si = StateInfo()
        _check(_lib().plb_get_state_info(self._h, C.byref(si)), "plb_get_state_info")
        return {k: getattr(si, k) for k, _ in si._fields_}

    def read_state(self):
        """The stored population elements (N, 19) in V order: float64
        populations, float32 deviations, or the uint16 bits of FP16 scaled
        deviations (plb_read_state)."""
        info = self.state_info()
        s = np.zeros((self.N, Q), dtype=self.STATE_DTYPE[info["precision"]])
        _check(_lib().plb_read_state(self._h, _ptr(s)), "plb_read_state")
        return s

    def write_state(self, s):
        info = self.state_info()
        s = np.ascontiguousarray(s, dtype=self.STATE_DTYPE[info["precision"]])
        assert s.shape == (self.N, Q)
        _check(_lib().plb_write_state(self._h, _ptr(s)), "plb_write_state")

    def read_table(self):
        t = np.zeros((18, self.N), dtype=np.int32)
        _check(_lib().plb_read_table(self._h, _ptr(t)), "plb_read_table")
        return t

    def reduce(self):
        r = Reduction()
        _check(_lib().plb_reduce(self._h, C.byref(r)), "plb_reduce")
        return r.as_dict()

    def run_to_convergence(self, it0, max_steps, check_every, tol):
        """plb_run_to_convergence: (steps taken, reductions of the last step)."""
        n = C.c_int64(0)
        r = Reduction()
        _check(_lib().plb_run_to_convergence(self._h, it0, max_steps, check_every, tol, C.byref(n),
                                             C.byref(r)), "plb_run_to_convergence")
        return n.value, r.as_dict()

    def last_fault(self):
        f = Fault()
        _check(_lib().plb_last_fault(self._h, C.byref(f)), "plb_last_fault")
        return {k: getattr(f, k) for k, _ in f._fields_}

    def bytes_per_update(self):
        return float(_lib().plb_bytes_per_update(self._h))


# ------------------------------------------------------------------ tuning
class TuneResult(C.Structure):
    _fields_ = [("cfg", Config), ("mlups", C.c_double), ("ms", C.c_double)]


def copy_bandwidth(nbytes=8 << 30, iters=5, device=0):
    """Practical HBM ceiling: GB/s of a streaming copy of nbytes."""
    g = C.c_double(0.0)
    _check(_lib().plb_copy_bandwidth(device, int(nbytes), iters, C.byref(g)), "plb_copy_bandwidth")
    return g.value


def autotune(flag, num_p=1, steps=20, cache_path=None, cfg=None, **kw):
    """Fastest (index encoding, layout, node order, block size, nodes per
    thread, waves per SIMD) for this mask, by plb_autotune's two-stage search;
    returns (Config, results)."""
    cfg = cfg or default_config(**kw)
    flag = np.ascontiguousarray(flag, dtype=np.uint8)
    nz, ny, nx = flag.shape
    best = Config()
    res = (TuneResult * 64)()
    n = C.c_int(0)
    _check(_lib().plb_autotune(C.byref(cfg), _ptr(flag), nx, ny, nz, num_p, steps,
                               cache_path.encode() if cache_path else None, C.byref(best), res, 64,
                               C.byref(n)), "plb_autotune")
    return best, _tune_rows(res[:n.value])


def _tune_rows(res):
    out = []
    for r in res:
        c = r.cfg
        out.append(dict(index_kind=c.index_kind, layout=c.layout, ordering=c.ordering, brick=tuple(c.brick),
                        block=c.block_size, npt=c.nodes_per_thread, waves=c.waves_per_simd,
                        mlups=r.mlups, ms=r.ms))
    return out


def _sample_array(flags, taus, gxs):
    flags = [np.ascontiguousarray(f, dtype=np.uint8) for f in flags]
    arr = (Sample * len(flags))()
    for s, (f, tau, gx) in enumerate(zip(flags, taus, gxs)):
        nz, ny, nx = f.shape
        arr[s] = Sample(f.ctypes.data, nx, ny, nz, tau, gx)
    return flags, arr


def batch_autotune(flags, taus, gxs, steps=20, cfg=None, persistent=False, **kw):
    """Fastest (index encoding, layout, node order, block size, waves per
    SIMD) for a batch of these samples, by plb_batch_autotune's three-stage
    search, timed with the per-step driver or (persistent=True) the
    persistent launch; returns (Config, results)."""
    cfg = cfg or default_config(**kw)
    keep, arr = _sample_array(flags, taus, gxs)
    best = Config()
    res = (TuneResult * 64)()
    n = C.c_int(0)
    _check(_lib().plb_batch_autotune(C.byref(cfg), arr, len(keep), steps, int(bool(persistent)), C.byref(best),
                                     res, 64, C.byref(n)), "plb_batch_autotune")
    return best, _tune_rows(res[:n.value])


# ------------------------------------------------------------------ batch
class Batch:
    """Many independent samples advanced together (plb_batch_*)."""

    def __init__(self, flags, taus, gxs, cfg=None, **kw):
        cfg = cfg or default_config(**kw)
        self._flags, arr = _sample_array(flags, taus, gxs)
        h = C.c_void_p()
        _check(_lib().plb_batch_create(C.byref(h), C.byref(cfg), arr, len(self._flags)),
               "plb_batch_create")
        self._h = h
        self.n = len(self._flags)
        self.nodes = [int(_lib().plb_batch_num_nodes(h, s)) for s in range(self.n)]

    def run(self, max_steps, check_every, tol, persistent=False):
        """(steps per sample, node updates, device ms); persistent=True runs
        plb_batch_run_persistent (one launch) instead of one launch per step."""
        steps = np.zeros(self.n, dtype=np.int64)
        upd = C.c_int64(0)
        ms = C.c_double(0.0)
        fn = _lib().plb_batch_run_persistent if persistent else _lib().plb_batch_run
        _check(fn(self._h, max_steps, check_every, tol, _ptr(steps), C.byref(upd), C.byref(ms)),
               "plb_batch_run_persistent" if persistent else "plb_batch_run")
        return steps, upd.value, ms.value

    def reduce(self):
        arr = (Reduction * self.n)()
        _check(_lib().plb_batch_reduce(self._h, arr), "plb_batch_reduce")
        return [arr[s].as_dict() for s in range(self.n)]

    def last_fault(self):
        """(sample, fault dict) of the first density-guard trip (sample -1: none)."""
        s = C.c_int(-1)
        f = Fault()
        _check(_lib().plb_batch_last_fault(self._h, C.byref(s), C.byref(f)), "plb_batch_last_fault")
        return s.value, {k: getattr(f, k) for k, _ in f._fields_}

    def read_macros(self, sample):
        """(4, n) rho, ux, uy, uz of one sample in its V order."""
        out = np.zeros((4, self.nodes[sample]))
        _check(_lib().plb_batch_read_macros(self._h, sample, _ptr(out[0]), _ptr(out[1]), _ptr(out[2]),
                                            _ptr(out[3])), "plb_batch_read_macros")
        return out

    def close(self):
        if self._h:
            _lib().plb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ files
def read_vbin(path):
    """V.bin (preprocess.c:166-193 layout)."""
    with open(path, "rb") as fh:
        num_node, num_p, nx, ny, nz = (int(v) for v in np.frombuffer(fh.read(20), dtype="<i4"))
        start = np.frombuffer(fh.read(4 * num_p), dtype="<i4").copy()
        V = np.frombuffer(fh.read(152 * num_node), dtype=NODE_DTYPE).copy()
    return V, start, (nz, ny, nx)


def write_vbin(path, V, start, dims):
    nz, ny, nx = dims
    with open(path, "wb") as fh:
        fh.write(np.array([len(V), len(start), nx, ny, nz], dtype="<i4").tobytes())
        fh.write(np.asarray(start, dtype="<i4").tobytes())
        fh.write(np.asarray(V, dtype=NODE_DTYPE).tobytes())
