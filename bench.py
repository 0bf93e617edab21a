#!/usr/bin/env python3
"""Benchmark of the sparse D3Q19 time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one D3Q19 time step (pull-stream + bounce-back + BGK with body
force) over every pore node of the domain.  Default workload: a 1024^3
periodic sphere pack (radius U[6,12], porosity 0.20, ~215 M pore nodes),
tau 0.6, g_x 5e-7, populations stored as FP32 deviations with FP32 collide
and half-way bounce-back.  The plain FP64 path (reference bounce order) is
timed on the same domain for the speed ratio.

Prints one JSON line with pore-node MLUPS (whole job), the roofline of the
step kernel (algorithmic bytes per update x nodes / average kernel time from
HIP events on the kernel's stream, against 8 TB/s) and a CPU baseline: the
reference solver, built from its own sources, timed with its own MPI_Wtime
method under mpirun on a bounded sample of the same kind of domain.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
STORAGE_BYTES = {0: 8, 1: 4, 2: 2}   # population storage bytes per precision

# Launch shapes measured on MI355X (plb_autotune candidates and a finer
# sweep): layout 1 = AoSoA tiles of 64 nodes; index 1 = TILE8 pull table;
# ordering 2 = pencils of `brick` = (nx, Py, 1) rows visited plane by plane.
# reserve_gib: device memory the timed lattice allocates first and holds (plb_config.placement_reserve_gib);
# on a freshly booted MI355X the first ~40 GiB the first process allocates stream about 6 % slower.
SPHERES1024 = dict(kind="spheres", dims=(1024, 1024, 1024), rmin=6.0, rmax=12.0, porosity=0.20, seed=4,
                   tau=0.6, gx=5e-7, reserve_gib=40, label="1024^3 sphere pack, porosity 0.20, tau 0.6, g_x 5e-7")
WORKLOADS = {
    "spheres1024-fp32": dict(SPHERES1024, precision=1, bb=1, index=1, layout=1, ordering=2,
                             brick=(1024, 128, 1), block=256, waves=5),
    "spheres1024-fp16": dict(SPHERES1024, precision=2, bb=1, index=1, layout=1, ordering=2,
                             brick=(1024, 64, 1), block=128, fp16_rounding=1),
    "spheres1024-fp64": dict(SPHERES1024, precision=0, bb=0, index=0, layout=1, ordering=2,
                             brick=(1024, 64, 1), block=256),
    "spheres256-fp64": dict(kind="spheres", dims=(256, 256, 256), rmin=6.0, rmax=12.0, porosity=0.30,
                            seed=2, tau=0.8, gx=1e-6, precision=0, bb=0, index=0, layout=1, ordering=2,
                            brick=(256, 64, 1), block=256,
                            label="256^3 sphere pack, porosity 0.30, tau 0.8, g_x 1e-6"),
    "noise512-fp32": dict(kind="noise", dims=(512, 512, 512), sigma=3.0, porosity=0.25, seed=3,
                          tau=1.0, gx=1e-6, precision=1, bb=1, index=1, layout=1, ordering=2,
                          brick=(512, 64, 1), block=256,
                          label="512^3 smoothed noise sigma 3, porosity 0.25, tau 1.0, g_x 1e-6"),
}
# The plain FP64 path the reduced-precision paths are compared with: the
# reference bounce order and the int32 pull table, in its fastest launch shape.
FP64_PLAIN = dict(precision=0, bb=0, index=0, layout=1, ordering=2, brick=(1024, 64, 1), block=256, waves=5)
PREC_NAME = {0: "f64", 1: "f32-store/f32-collide", 2: "f16-store/f32-collide"}
# The paths timed beside the line (plain FP64, FP16 storage) take a fixed
# window whatever --steps and --warmup are (about 1.4 s and 0.4 s of GPU time).
SECONDARY_STEPS, SECONDARY_WARMUP = 100, 10


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_mask(P, w):
    """Seeded mask; cached per machine under $TMPDIR (a 1024^3 sphere pack
    takes ~15 s to generate on one host core)."""
    key = "_".join(str(w[k]) for k in ("kind", "dims", "porosity", "seed") if k in w)
    key += "_" + "_".join(str(w.get(k)) for k in ("rmin", "rmax", "sigma"))
    path = os.path.join(tempfile.gettempdir(), "plb_mask_" + "".join(c if c.isalnum() else "_" for c in key) + ".u8")
    nvox = int(np.prod(w["dims"]))
    if os.path.exists(path) and os.path.getsize(path) == nvox:
        return np.fromfile(path, dtype=np.uint8).reshape(w["dims"])
    if w["kind"] == "spheres":
        flag = P.mask_spheres(w["dims"], w["rmin"], w["rmax"], w["porosity"], w["seed"])
    else:
        flag = P.mask_noise(w["dims"], w["sigma"], w["porosity"], w["seed"])
    try:
        flag.tofile(path + ".tmp")
        os.replace(path + ".tmp", path)
    except OSError:
        pass
    return flag


def time_path(P, flag, w, precision, bb, steps, warmup, layout=0, ordering=0, block=256, index=0,
              brick=(16, 4, 4), waves=0, fp16_rounding=1, keep=False, reserve_gib=0):
    """Times `steps` steps after `warmup` on a new lattice.  keep=True returns
    the lattice open (result["lattice"]) for the precision gate, so the gate's
    steps run on the allocation that was timed."""
    cfg = P.default_config(tau=w["tau"], gx=w["gx"], precision=precision, bb_mode=bb, layout=layout,
                           ordering=ordering, block_size=block, index_kind=index, brick=brick,
                           waves_per_simd=waves, fp16_rounding=fp16_rounding, placement_reserve_gib=reserve_gib)
    t0 = time.perf_counter()
    lat = P.Lattice.from_flag(flag, 1, cfg=cfg)
    log(f"lattice built: {lat.N} pore nodes, precision {precision}, {time.perf_counter() - t0:.1f} s")
    try:
        lat.step(0, warmup)
        t0 = time.perf_counter()
        ev_ms = lat.step_timed(warmup, steps)      # HIP events on the context stream; syncs
        wall = time.perf_counter() - t0
        red = lat.reduce()
        bpu = lat.bytes_per_update()
        n = lat.N
    except BaseException:
        lat.close()
        raise
    out = dict(nodes=n, wall_s=wall, event_ms=ev_ms, bytes_per_update=bpu, reduction=red)
    if keep:
        out["lattice"] = lat
    else:
        lat.close()
    return out


def cpu_baseline(P, w, max_seconds=20.0):
    """The reference porelb.x (oracle/_ref, built from the reference's sources)
    under mpirun on all usable host cores, timed by its own MPI_Wtime loop
    timer, on a 256^3 domain of the same kind and porosity."""
    import oracle as O

    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    host = {"os_cpu_count": os.cpu_count(), "affinity": affinity}
    cores = min(16, affinity)
    ranks = 1
    while ranks * 2 <= cores:
        ranks *= 2
    dims = (256, 256, 256)
    sample = dict(w, dims=dims)
    flag = make_mask(P, sample)
    npore = int(flag.sum())
    if O.have_reference_build() and os.path.exists(O.MPIRUN):
        per_rank_mlups = 8.0
        steps = int(max(10, min(1000, max_seconds * ranks * per_rank_mlups * 1e6 / npore)))
        wd = tempfile.mkdtemp(prefix="plbcpu_")
        log(f"cpu baseline: reference porelb.x, mpirun -np {ranks}, {npore} nodes, {steps} steps")
        r = O.run_reference(flag, ranks, steps, w["tau"], w["gx"], save_every=1000, print_steps=1,
                            workdir=wd, dump=False, timeout=900)
        subprocess.run(["rm", "-rf", wd])
        t = r["timing"]
        return {"value": round(t["mlups"], 3), "unit": "pore-node MLUPS", "cores": ranks, "ranks": ranks,
                "host_cores": host, "kind": "reference",
                "sample": f"{dims[0]}^3 {w['kind']} domain, porosity {w['porosity']}, {npore} pore nodes, "
                          f"{steps} FP64 steps, mpirun -np {ranks} (reference porelb.c, MPI_Wtime loop "
                          f"timer incl. per-step print and it%1000 save)"}
    # Fallback: the C restatement in oracle/, one thread.
    V, start = O.build_V(flag, 1)
    tab = O.pull_table(V, start, O.BB_REFCOMPAT)
    f0, _ = O.init_equilibrium(len(V))
    steps = 5
    t0 = time.perf_counter()
    O.pull_run(tab, f0, steps, w["tau"], w["gx"])
    dt = time.perf_counter() - t0
    return {"value": round(len(V) * steps / dt / 1e6, 3), "unit": "pore-node MLUPS", "cores": 1, "ranks": 1,
            "host_cores": host, "kind": "port", "sample": f"{dims[0]}^3 domain, {len(V)} pore nodes, {steps} FP64 steps, 1 thread"}


PRECISION_VARIANTS = (("fp64", 0, 0), ("fp32_store", 1, 0), ("fp16_store_sr", 2, 1), ("fp16_store_rn", 2, 0))


def study_precision(P, max_steps, check, tol, variants=PRECISION_VARIANTS):
    """Config 3: 512^3 smoothed noise (sigma 3, porosity 0.25), tau 1.0,
    g_x 1e-6, half-way bounce-back; FP64 vs FP32-storage vs FP16-storage
    (stochastic and round-to-nearest stores) run to steady state
    (|d<u_x>| <= tol |<u_x>| between checks, or max_steps).  Reports the
    relative L2 error of the velocity field and of K against FP64."""
    w = WORKLOADS["noise512-fp32"]
    flag = make_mask(P, w)
    log(f"mask {w['dims']} porosity {flag.mean():.4f}")
    runs = {}
    for name, prec, rnd in variants:
        cfg = P.default_config(tau=w["tau"], gx=w["gx"], precision=prec, bb_mode=1, index_kind=1, layout=1,
                               ordering=2, brick=w["brick"], block_size=256, fp16_rounding=rnd,
                               waves_per_simd=0 if prec == 2 else 5)  # the cap slows FP16 only
        with P.Lattice.from_flag(flag, 1, cfg=cfg) as lat:
            m0 = lat.reduce()["mass"]
            t0 = time.perf_counter()
            prev, it = None, 0
            while it < max_steps:
                lat.step(it, check)
                it += check
                r = lat.reduce()
                if (it // check) % 20 == 0:
                    log(f"{name}: step {it}, <u_x> {r['mean_ux']:.12e}")
                if prev is not None and abs(r["mean_ux"] - prev) <= tol * abs(r["mean_ux"]):
                    break
                prev = r["mean_ux"]
            wall = time.perf_counter() - t0
            mac = lat.read_macros()
            runs[name] = dict(steps=it, K=r["K"], mean_ux=r["mean_ux"], mass_drift=(r["mass"] - m0) / m0,
                              wall_s=wall, mlups=lat.N * it / wall / 1e6, u=mac[1:].copy(), nodes=lat.N,
                              converged=it < max_steps)
        log(f"{name}: {runs[name]['steps']} steps, K {runs[name]['K']:.9e}")
    ref = runs["fp64"]
    out = {"study": "precision", "workload": w["label"], "pore_nodes": ref["nodes"], "tol": tol,
           "check_every": check, "max_steps": max_steps, "results": {}}
    for name, _, _ in variants:
        r = runs[name]
        out["results"][name] = {
            "steps": r["steps"], "converged": r["converged"], "K": r["K"], "mean_ux": r["mean_ux"],
            "mass_rel_drift": r["mass_drift"], "mlups_incl_checks": round(r["mlups"], 1),
            "velocity_rel_l2_err": float(np.linalg.norm(r["u"] - ref["u"]) / np.linalg.norm(ref["u"])),
            "K_rel_err": abs(r["K"] / ref["K"] - 1.0)}
    print(json.dumps(out), flush=True)


def precision_runs(P, flag, w, precs, steps, chunk=500, lattices=None):
    """The same number of steps from rest in each storage precision of
    `precs` (the first one is the FP64 reference), all with half-way
    bounce-back.  A precision found in `lattices` (an open half-way lattice of
    that precision) is restarted from rest on it; the others run on a new
    lattice in the workload's TILE8 / pencil shape.  Launch shapes do not
    change results.  Returns per precision the FP64 reductions after the last
    step, the mass drift, the velocity field (V order) and the MLUPS of the
    timed steps."""
    import contextlib

    runs = {}
    for prec in precs:
        cfg = P.default_config(tau=w["tau"], gx=w["gx"], precision=prec, bb_mode=1, index_kind=1, layout=1,
                               ordering=2, brick=w["brick"], block_size=256,
                               waves_per_simd=0 if prec == 2 else w.get("waves", 5),
                               fp16_rounding=w.get("fp16_rounding", 1))
        given = (lattices or {}).get(prec)
        if given is not None:
            given.init_equilibrium()
        with (contextlib.nullcontext(given) if given is not None else P.Lattice.from_flag(flag, 1, cfg=cfg)) as lat:
            m0 = lat.reduce()["mass"]
            ms, it = 0.0, 0
            while it < steps:
                n = min(chunk, steps - it)
                ms += lat.step_timed(it, n)
                it += n
                log(f"precision {prec}: step {it}")
            r = lat.reduce()
            mlups = lat.N * steps / (ms * 1e-3) / 1e6
            runs[prec] = dict(K=r["K"], mean_ux=r["mean_ux"], mass_drift=(r["mass"] - m0) / m0, mlups=mlups,
                              roofline_frac=mlups * 1e6 * (2 * 19 * STORAGE_BYTES[prec] + 19 * 4) / (HBM_PEAK_GBS * 1e9),
                              u=lat.read_macros()[1:].copy(), nodes=lat.N)
        log(f"precision {prec}: {mlups:.0f} MLUPS, K {r['K']:.9e}")
    return runs


def errors_vs_fp64(runs, prec):
    u64 = runs[0]["u"]
    return {"K_rel_err": abs(runs[prec]["K"] / runs[0]["K"] - 1.0),
            "velocity_rel_l2_err": float(np.linalg.norm(runs[prec]["u"] - u64) / np.linalg.norm(u64))}


def study_config4(P, steps):
    """Config 4: the 1024^3 sphere pack (porosity 0.20, tau 0.6, g_x 5e-7),
    the same number of steps in FP64 and in FP32 storage (both half-way
    bounce-back, TILE8, pencils, at most 5 waves per SIMD).  Reports MLUPS and
    the roofline fraction of each run, and the FP32 errors in K and in the
    velocity field against FP64."""
    w = WORKLOADS["spheres1024-fp32"]
    runs = precision_runs(P, make_mask(P, w), w, (0, 1), steps)
    out = {"study": "config4", "workload": w["label"], "pore_nodes": runs[0]["nodes"], "steps": steps,
           "bounce_back": "halfway", "results": {}}
    for prec, name in ((0, "fp64"), (1, "fp32_store")):
        r = runs[prec]
        out["results"][name] = dict({
            "mlups": round(r["mlups"], 1), "roofline_frac": round(r["roofline_frac"], 4), "K": r["K"],
            "mean_ux": r["mean_ux"], "mass_rel_drift": r["mass_drift"]}, **errors_vs_fp64(runs, prec))
    print(json.dumps(out), flush=True)


def study_preprocess(P):
    """The index build (preprocess.c:29-193 and the pull table): the
    reference's own build_V and save_V (one host core; 256^3 only, larger
    masks take many minutes) against the GPU build -- plb_preprocess with the
    V records copied to host memory, preprocess.x end to end including the
    V.bin write, and plb_create_from_flag for the benchmark's step shape."""
    import oracle as O

    cli = os.path.join(REPO, "porelb_amd", "bin", "preprocess.x")
    P.preprocess(np.ones((4, 4, 4), dtype=np.uint8), 1)  # runtime start-up outside the timings
    out = {"study": "preprocess", "results": []}
    for key, num_p, ref in (("spheres256-fp64", 16, True), ("noise512-fp32", 1, False),
                            ("spheres1024-fp32", 1, False)):
        w = WORKLOADS[key]
        flag = make_mask(P, w)
        nz, ny, nx = flag.shape
        res = {"mask": w["label"].split(",")[0] + f", porosity {w['porosity']}", "num_p": num_p}
        t0 = time.perf_counter()
        V, start = P.preprocess(flag, num_p)
        res["gpu_build_V_to_host_s"] = round(time.perf_counter() - t0, 3)
        res["pore_nodes"] = int(len(V))
        del V
        wd = tempfile.mkdtemp(prefix="plbpre_")
        try:
            flag.tofile(os.path.join(wd, "flag.bin"))
            vfile = "V.bin" if nx * ny * nz <= 512 ** 3 else "/dev/null"
            t0 = time.perf_counter()
            subprocess.run([cli, str(num_p), str(nx), str(ny), str(nz), "flag.bin", vfile], cwd=wd, check=True,
                           capture_output=True, timeout=900)
            res["gpu_preprocess_x_s"] = round(time.perf_counter() - t0, 3)
            res["gpu_preprocess_x_writes"] = vfile
            if ref and os.path.exists(O.REF_PREPROCESS):
                log(f"reference preprocess on {key}")
                r = subprocess.run([O.REF_PREPROCESS, str(num_p), str(nx), str(ny), str(nz), "flag.bin", "Vref.bin"],
                                   cwd=wd, check=True, capture_output=True, text=True, timeout=1800)
                line = [ln for ln in r.stdout.splitlines() if ln.startswith("PLB_REF_PREPROCESS")][0]
                t = {k: float(v) for k, v in (tok.split("=") for tok in line.split()[1:])}
                res["ref_build_V_s"], res["ref_save_V_s"] = t["build_s"], t["save_s"]
                # SOLID entries and padding are uninitialised in the reference's
                # records, so the files are compared field by field.
                a, b = O.read_vbin(os.path.join(wd, "V.bin")), O.read_vbin(os.path.join(wd, "Vref.bin"))
                res["V_semantic_fields_equal"] = bool(
                    np.array_equal(a["start"], b["start"]) and
                    all(np.array_equal(x, y) for x, y in zip(O.semantic_fields(a["V"]), O.semantic_fields(b["V"]))))
        finally:
            subprocess.run(["rm", "-rf", wd])
        sw = WORKLOADS["spheres1024-fp32"]
        t0 = time.perf_counter()
        lat = P.Lattice.from_flag(flag, 1, tau=w["tau"], gx=w["gx"], precision=1, bb_mode=1, index_kind=1, layout=1,
                                  ordering=2, brick=(nx,) + tuple(sw["brick"][1:]), block_size=256)
        res["gpu_lattice_create_s"] = round(time.perf_counter() - t0, 3)
        lat.close()
        log(json.dumps(res))
        out["results"].append(res)
    print(json.dumps(out), flush=True)


def study_outputs(P, reps=5):
    """The output side of a step at 1024^3 (FP32 bench shape): the FP64
    reduction (plb_reduce: populations and u_x read once, 19*4 + 8 B per node)
    and the macro read-back that save_data needs (plb_read_macros: 4 x 8 B per
    node gathered to V order and copied to pageable, then to page-locked host
    memory)."""
    w = WORKLOADS["spheres1024-fp32"]
    flag = make_mask(P, w)
    with P.Lattice.from_flag(flag, 1, tau=w["tau"], gx=w["gx"], precision=1, bb_mode=1, index_kind=1, layout=1,
                             ordering=2, brick=w["brick"], block_size=256) as lat:
        lat.step(0, 10)
        n = lat.N
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            lat.reduce()
            t.append(time.perf_counter() - t0)
        red = min(t)
        t = []
        for _ in range(2):
            t0 = time.perf_counter()
            lat.read_macros()
            t.append(time.perf_counter() - t0)
        mac = min(t)
        pinned = P.host_array((4, n))
        t = []
        for _ in range(2):
            t0 = time.perf_counter()
            lat.read_macros(out=pinned)
            t.append(time.perf_counter() - t0)
        mac_pinned = min(t)
        del pinned
    out = {"study": "outputs", "workload": w["label"], "pore_nodes": n,
           "reduce_s": round(red, 5), "reduce_bytes": n * (19 * 4 + 8),
           "reduce_gbs": round(n * (19 * 4 + 8) / red / 1e9, 1),
           "reduce_roofline_frac": round(n * (19 * 4 + 8) / red / (HBM_PEAK_GBS * 1e9), 3),
           "read_macros_s": round(mac, 4), "read_macros_host_gbs": round(n * 32 / mac / 1e9, 2),
           "read_macros_pinned_s": round(mac_pinned, 4),
           "read_macros_pinned_gbs": round(n * 32 / mac_pinned / 1e9, 2)}
    print(json.dumps(out), flush=True)


def batch_samples(P, n_samples=64, seed=100):
    """The batching workload: n samples with edge L ~ U{64..192} and
    porosity ~ U[0.15, 0.45], alternately sphere packs (r U[6,12]) and
    smoothed noise (sigma 2), seeded; every sample runs at tau 1.0, g_x 1e-6."""
    rng = np.random.default_rng(seed)
    flags = []
    for k in range(n_samples):
        L = int(rng.integers(64, 193))
        phi = float(rng.uniform(0.15, 0.45))
        if k % 2 == 0:
            flags.append(P.mask_spheres((L, L, L), 6.0, 12.0, phi, seed + k))
        else:
            flags.append(P.mask_noise((L, L, L), 2.0, phi, seed + k))
    return flags


def study_batch(P, n_samples, max_steps, check, tol, seed=100, waves=0, tune_steps=20):
    """Config 5: n samples, edge L ~ U{64..192}, porosity ~ U[0.15, 0.45], half
    sphere packs (r U[6,12]) and half smoothed noise (sigma 2), FP32 storage,
    half-way bounce-back, tau 1.0, g_x 1e-6.  One batched launch per step over
    all active samples (per-sample convergence stop) vs the same samples run
    one after another (one sample per launch, same kernels and convergence
    rule); at most `waves` step-kernel waves per SIMD in both (TILE8, pencils
    of 32 rows, 256 threads).  Then the same per-step batch in the shape
    plb_batch_autotune picks for the whole batch, the persistent launch in the
    shape tuned for it, and the per-step batch as one batch per size class
    (L <= 128, L > 128), each in the shape tuned on that class."""
    flags = batch_samples(P, n_samples, seed)
    taus, gxs = [1.0] * n_samples, [1e-6] * n_samples
    kw = dict(precision=1, bb_mode=1, index_kind=1, layout=1, ordering=2, brick=(192, 32, 1), block_size=256,
              waves_per_simd=waves)
    modes = {}

    def run_batch(idx, cfg, persistent=False):
        b = P.Batch([flags[k] for k in idx], [taus[k] for k in idx], [gxs[k] for k in idx], cfg=cfg)
        try:
            steps, upd, ms = b.run(max_steps, check, tol, persistent=persistent)
            red = b.reduce()
            nodes = list(b.nodes)
        finally:
            b.close()
        return dict(steps=[int(x) for x in steps], upd=int(upd), ms=ms, K=[r["K"] for r in red], nodes=nodes)

    everyone = list(range(n_samples))
    for persistent in (False, True):
        m = run_batch(everyone, P.default_config(**kw), persistent)
        nodes = m.pop("nodes")
        modes["persistent" if persistent else "per_step"] = m
        log(f"batch {'persistent' if persistent else 'per-step'}: {m['ms']:.1f} ms, "
            f"{m['upd'] / (m['ms'] * 1e-3) / 1e6:.0f} MLUPS")

    def shape(cfg):
        return {"index_kind": cfg.index_kind, "layout": cfg.layout, "ordering": cfg.ordering,
                "pencil_rows": cfg.brick[1] if cfg.ordering == 2 else None, "block": cfg.block_size,
                "waves_per_simd": cfg.waves_per_simd}

    base = P.default_config(**kw)
    t0 = time.perf_counter()
    tuned_cfg, _ = P.batch_autotune(flags, taus, gxs, tune_steps, cfg=base)
    tune_s = time.perf_counter() - t0
    tuned = run_batch(everyone, tuned_cfg)
    tuned.pop("nodes")
    log(f"batch tuned {shape(tuned_cfg)}: {tuned['ms']:.1f} ms ({tune_s:.1f} s of tuning)")
    t0 = time.perf_counter()
    ptuned_cfg, _ = P.batch_autotune(flags, taus, gxs, tune_steps, cfg=base, persistent=True)
    ptune_s = time.perf_counter() - t0
    ptuned = run_batch(everyone, ptuned_cfg, persistent=True)
    ptuned.pop("nodes")
    log(f"batch persistent tuned {shape(ptuned_cfg)}: {ptuned['ms']:.1f} ms ({ptune_s:.1f} s of tuning)")
    classes = {"L<=128": [k for k in everyone if max(flags[k].shape) <= 128],
               "L>128": [k for k in everyone if max(flags[k].shape) > 128]}
    per_class = {"updates": 0, "ms": 0.0, "classes": {}}
    class_steps, class_K = [0] * n_samples, [0.0] * n_samples
    for name, idx in classes.items():
        if not idx:
            continue
        ccfg, _ = P.batch_autotune([flags[k] for k in idx], [taus[k] for k in idx], [gxs[k] for k in idx],
                                   tune_steps, cfg=base)
        m = run_batch(idx, ccfg)
        for j, k in enumerate(idx):
            class_steps[k], class_K[k] = m["steps"][j], m["K"][j]
        per_class["updates"] += m["upd"]
        per_class["ms"] += m["ms"]
        per_class["classes"][name] = dict(samples=len(idx), shape=shape(ccfg),
                                          mlups=round(m["upd"] / (m["ms"] * 1e-3) / 1e6, 1))
        log(f"batch class {name} ({len(idx)} samples) {shape(ccfg)}: {m['ms']:.1f} ms")
    seq_upd, seq_ms, seq_steps = 0, 0.0, []
    for k, f in enumerate(flags):
        one = P.Batch([f], [taus[k]], [gxs[k]], **kw)
        st, u, m = one.run(max_steps, check, tol)
        seq_upd += u
        seq_ms += m
        seq_steps.append(int(st[0]))
        one.close()
    log(f"batch sequential: {seq_ms:.1f} ms")

    def rate(u, m):
        return {"updates": int(u), "ms": round(m, 1), "mlups": round(u / (m * 1e-3) / 1e6, 1)}

    pst, per = modes["persistent"], modes["per_step"]
    out = {"study": "batch", "samples": n_samples, "pore_nodes_total": int(sum(nodes)), "waves_per_simd": waves,
           "tol": tol, "check_every": check, "max_steps": max_steps,
           "per_step": rate(per["upd"], per["ms"]), "persistent": rate(pst["upd"], pst["ms"]),
           "sequential": rate(seq_upd, seq_ms),
           "steps_min": int(min(per["steps"])), "steps_max": int(max(per["steps"])),
           "same_steps_all_three": per["steps"] == pst["steps"] == seq_steps,
           "same_K_persistent_per_step": per["K"] == pst["K"],
           "K_range": [float(min(per["K"])), float(max(per["K"]))]}
    out["speedup_per_step_vs_sequential"] = round(out["per_step"]["mlups"] / out["sequential"]["mlups"], 3)
    out["speedup_persistent_vs_sequential"] = round(out["persistent"]["mlups"] / out["sequential"]["mlups"], 3)
    # Node values are shape-independent; the FP64 sums behind K and the stop
    # test follow the node order, so K may differ in its last bits.
    def kdiff(K):
        return float(max(abs(a - b) / abs(b) for a, b in zip(K, per["K"]) if b != 0.0))

    out["tuned"] = dict(rate(tuned["upd"], tuned["ms"]), shape=shape(tuned_cfg), tuning_s=round(tune_s, 1),
                        same_steps=tuned["steps"] == per["steps"], K_max_rel_diff=kdiff(tuned["K"]))
    out["persistent_tuned"] = dict(rate(ptuned["upd"], ptuned["ms"]), shape=shape(ptuned_cfg),
                                   tuning_s=round(ptune_s, 1), same_steps=ptuned["steps"] == per["steps"],
                                   K_max_rel_diff=kdiff(ptuned["K"]))
    out["tuned_per_class"] = dict(rate(per_class["updates"], per_class["ms"]), classes=per_class["classes"],
                                  same_steps=class_steps == per["steps"], K_max_rel_diff=kdiff(class_K))
    print(json.dumps(out), flush=True)


def study_batch_steps(P, steps, waves=4, seed=100):
    """The 64-sample batch for a fixed number of steps (tol 0: no sample
    stops), once per step and as one persistent launch, in one shape (FP32,
    TILE8, AoSoA, pencils of 32 rows, 256 threads, at most `waves` waves per
    SIMD): a short run for counter passes that compare the two drivers."""
    flags = batch_samples(P, 64, seed)
    cfg = P.default_config(precision=1, bb_mode=1, index_kind=1, layout=1, ordering=2, brick=(192, 32, 1),
                           block_size=256, waves_per_simd=waves)
    b = P.Batch(flags, [1.0] * 64, [1e-6] * 64, cfg=cfg)
    try:
        _, u1, m1 = b.run(steps, steps, 0.0)
        _, u2, m2 = b.run(steps, steps, 0.0, persistent=True)
    finally:
        b.close()
    print(json.dumps({"study": "batch-steps", "pore_nodes_total": int(sum(b.nodes)), "steps": steps,
                      "waves_per_simd": waves,
                      "per_step": {"ms": round(m1, 1), "mlups": round(u1 / (m1 * 1e-3) / 1e6, 1)},
                      "persistent": {"ms": round(m2, 1), "mlups": round(u2 / (m2 * 1e-3) / 1e6, 1)}}), flush=True)


def study_aa(P, steps=100, warmup=10):
    """The 1024^3 sphere pack in both stream modes: two buffers with pull
    steps, and one buffer with the AA pattern (PLB_STREAM_AA: local and
    neighbour steps alternating), FP32 and FP16 storage in their benchmark
    shapes, `steps` timed steps each after `warmup` (an even count, so both
    AA kernels run equally often), two rounds in alternation."""
    w = SPHERES1024
    flag = make_mask(P, w)
    out = {"study": "aa", "workload": w["label"], "steps": steps, "runs": []}
    for name in ("spheres1024-fp32", "spheres1024-fp16"):
        wk = WORKLOADS[name]
        for mode in (0, 1, 0, 1):
            cfg = P.default_config(tau=w["tau"], gx=w["gx"], precision=wk["precision"], bb_mode=1,
                                   index_kind=wk["index"], layout=wk["layout"], ordering=wk["ordering"],
                                   brick=wk["brick"], block_size=wk["block"], waves_per_simd=wk.get("waves", 0),
                                   fp16_rounding=wk.get("fp16_rounding", 1), placement_reserve_gib=w["reserve_gib"],
                                   stream_mode=mode)
            lat = P.Lattice.from_flag(flag, 1, cfg=cfg)
            try:
                lat.step(0, warmup)
                ms = lat.step_timed(warmup, steps)
                run = {"workload": name, "stream": "aa" if mode else "pull", "mlups": round(lat.N * steps / ms / 1e3, 1),
                       "ms_per_step": round(ms / steps, 4), "bytes_per_update": round(lat.bytes_per_update(), 2)}
            finally:
                lat.close()
            log(json.dumps(run))
            out["runs"].append(run)
    print(json.dumps(out), flush=True)


def study_placement(P, reserve_gib, steps=200, warmup=20):
    """Which allocations of the FP32 1024^3 lattice stream slower: after the
    benchmark's mask and copy-bandwidth run, hold `reserve_gib` GiB, then time
    lattice A, lattice B while A is alive, and lattice C after A is freed.
    Run it as the first and later processes on a freshly booted box."""
    import ctypes

    w = WORKLOADS["spheres1024-fp32"]
    flag = make_mask(P, w)
    P.copy_bandwidth(4 << 30, 5)
    hip = ctypes.CDLL("libamdhip64.so")
    held = ctypes.c_void_p()
    nbytes = ctypes.c_size_t(int(reserve_gib) << 30)
    if reserve_gib > 0:
        if hip.hipMalloc(ctypes.byref(held), nbytes) != 0 or hip.hipMemset(held, 0, nbytes) != 0:
            raise RuntimeError("hipMalloc of the held block failed")
        hip.hipDeviceSynchronize()

    def timed():
        cfg = P.default_config(tau=w["tau"], gx=w["gx"], precision=1, bb_mode=1, layout=1, ordering=2,
                               block_size=256, index_kind=1, brick=w["brick"], waves_per_simd=5)
        lat = P.Lattice.from_flag(flag, 1, cfg=cfg)
        lat.step(0, warmup)
        ms = lat.step_timed(warmup, steps)
        return lat, round(lat.N * steps / (ms * 1e-3) / 1e6, 1)

    out = {"study": "placement", "held_gib": reserve_gib, "workload": w["label"], "steps": steps}
    a, out["A_mlups"] = timed()
    b, out["B_A_alive_mlups"] = timed()
    a.close()
    c, out["C_A_freed_mlups"] = timed()
    b.close()
    c.close()
    if reserve_gib > 0:
        hip.hipFree(held)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="spheres1024-fp32", choices=sorted(WORKLOADS))
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-fp16", action="store_true")
    ap.add_argument("--layout", type=int, default=None)
    ap.add_argument("--ordering", type=int, default=None)
    ap.add_argument("--block", type=int, default=None)
    ap.add_argument("--index", type=int, default=None)
    ap.add_argument("--waves", type=int, default=None, help="step-kernel waves per SIMD (0 = no cap)")
    ap.add_argument("--study", choices=["precision", "batch", "batch-steps", "config4", "preprocess", "outputs",
                                        "placement", "aa"],
                    default=None,
                    help="run a study instead of the benchmark line")
    ap.add_argument("--max-steps", type=int, default=100000)
    ap.add_argument("--check", type=int, default=200)
    ap.add_argument("--tol", type=float, default=None)
    ap.add_argument("--samples", type=int, default=64)
    ap.add_argument("--reserve-gib", type=int, default=0, help="placement study: GiB held before the lattices")
    ap.add_argument("--gate-steps", type=int, default=5000,
                    help="steps of the FP64 vs reduced-precision comparison in the line (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the step runs on one GPU; --gpus {args.gpus} is not supported "
              "(the domain of every configuration fits one MI355X)", file=sys.stderr)
        sys.exit(2)

    import porelb_amd as P

    if args.study == "precision":
        study_precision(P, args.max_steps, args.check, args.tol if args.tol is not None else 1e-8)
        return
    if args.study == "batch":
        study_batch(P, args.samples, args.max_steps, args.check, args.tol if args.tol is not None else 1e-6,
                    waves=args.waves if args.waves is not None else 5)
        return
    if args.study == "aa":
        study_aa(P)
        return
    if args.study == "batch-steps":
        study_batch_steps(P, args.max_steps if args.max_steps != 100000 else 200)
        return
    if args.study == "placement":
        study_placement(P, args.reserve_gib)
        return
    if args.study == "outputs":
        study_outputs(P)
        return
    if args.study == "preprocess":
        study_preprocess(P)
        return
    if args.study == "config4":
        study_config4(P, args.max_steps if args.max_steps != 100000 else 5000)
        return
    w = WORKLOADS[args.workload]
    cpu = None
    if not args.skip_cpu:
        try:
            cpu = cpu_baseline(P, w)
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline failed:", repr(e))
    t0 = time.perf_counter()
    flag = make_mask(P, w)
    # the mask's flag.bin bytes (u8 [nz][ny][nx], x fastest) identify the input
    mask_sha256 = hashlib.sha256(memoryview(flag).cast("B")).hexdigest()
    log(f"mask {w['dims']} porosity {flag.mean():.4f} in {time.perf_counter() - t0:.1f} s")
    # Practical ceiling: a streaming copy of 4 GiB (its rocprofv3 FETCH_SIZE /
    # WRITE_SIZE also calibrate the counter scale for the step kernel).
    copy_gbs = P.copy_bandwidth(4 << 30, 5)
    log(f"copy ceiling {copy_gbs:.0f} GB/s")

    shape = {k: (getattr(args, k) if getattr(args, k) is not None else w[k])
             for k in ("layout", "ordering", "block", "index")}
    shape["waves"] = args.waves if args.waves is not None else w.get("waves", 0)
    # The timed lattice stays open: the precision gate below restarts it from
    # rest, so every launch of this kernel in the run is on one allocation.
    gate = args.gate_steps > 0 and w["precision"] != 0 and w["bb"] == 1
    main_run = time_path(P, flag, w, w["precision"], w["bb"], args.steps, args.warmup, shape["layout"],
                         shape["ordering"], shape["block"], shape["index"], w.get("brick", (16, 4, 4)),
                         shape["waves"], w.get("fp16_rounding", 1), keep=gate,
                         reserve_gib=w.get("reserve_gib", 0))
    kept = {w["precision"]: main_run.pop("lattice")} if gate else {}
    n, steps = main_run["nodes"], args.steps
    mlups = n * steps / main_run["wall_s"] / 1e6
    avg_ms = main_run["event_ms"] / steps
    # Algorithmic bytes per pore-node update: both population copies plus a
    # 19-entry int32 pull table, B(s) = 2*19*s + 19*4 (228 B for FP32 storage).
    # The layout's own bytes (TILE8 index, 18 entries) are reported beside it.
    algo_bpu = 2 * 19 * STORAGE_BYTES[w["precision"]] + 19 * 4
    bpu = main_run["bytes_per_update"]
    achieved = algo_bpu * n / (avg_ms * 1e-3) / 1e9
    layout_gbs = bpu * n / (avg_ms * 1e-3) / 1e9
    # HBM bytes per launch from the committed rocprofv3 --pmc summary of this
    # workload (FETCH_SIZE and WRITE_SIZE passes, scaled by a copy kernel of
    # known size); counters cannot be read inside this process, so the
    # figure and the commit it was measured on are named in the line.
    traffic, traffic_commit = None, None
    tname = f"profiles/step_{args.workload}_pmc.md"
    tfile = os.path.join(REPO, tname)
    if os.path.exists(tfile):
        for line in open(tfile):
            if line.startswith("hbm_bytes_per_launch"):
                traffic = float(line.split("=")[1])
            if line.startswith("measured_at_commit"):
                traffic_commit = line.split("=")[1].strip()

    secondary = {}
    if not args.skip_fp64 and w["precision"] != 0:
        fp = FP64_PLAIN
        brick = (w["dims"][2],) + fp["brick"][1:]
        fp64 = time_path(P, flag, w, 0, 0, SECONDARY_STEPS, SECONDARY_WARMUP, fp["layout"],
                         fp["ordering"], fp["block"], fp["index"], brick, fp["waves"])
        fp64_mlups = fp64["nodes"] * SECONDARY_STEPS / fp64["wall_s"] / 1e6
        secondary = {"fp64_plain_mlups": round(fp64_mlups, 1),
                     "fp64_plain_roofline_frac": round((2 * 19 * 8 + 19 * 4) * fp64["nodes"] /
                                                       (fp64["event_ms"] / SECONDARY_STEPS * 1e-3) /
                                                       1e9 / HBM_PEAK_GBS, 4),
                     "speedup_vs_fp64_plain": round(mlups / fp64_mlups, 3),
                     "fp64_plain_shape": "int32 table, AoSoA, pencils of 64 rows, 256 threads, "
                                         "at most 5 waves per SIMD"}

    # FP16 storage (stochastic rounding) in its own fastest shape, beside the line.
    if not args.skip_fp16 and w["precision"] == 1:
        w16 = WORKLOADS["spheres1024-fp16"]
        st16 = SECONDARY_STEPS
        r16 = time_path(P, flag, w, 2, 1, st16, SECONDARY_WARMUP, w16["layout"], w16["ordering"],
                        w16["block"], w16["index"], w16["brick"], 0, w16["fp16_rounding"], keep=gate)
        if gate:
            kept[2] = r16.pop("lattice")
        m16 = r16["nodes"] * st16 / r16["wall_s"] / 1e6
        secondary.update({"fp16_store_mlups": round(m16, 1),
                          "fp16_store_roofline_frac": round((2 * 19 * 2 + 19 * 4) * r16["nodes"] /
                                                            (r16["event_ms"] / st16 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "fp16_store_shape": "stochastic rounding, TILE8, AoSoA, pencils of 64 rows, 128 threads"})

    # The metric's precision gate: the same steps from rest in FP64 and in the
    # reduced storage forms (both half-way bounce-back), K = nu q / g_x
    # (nu = (tau - 1/2)/3, porelb.c:89) and the velocity field compared.
    precision = None
    if gate:
        t0 = time.perf_counter()
        precs = (0, 1, 2) if (w["precision"] == 1 and not args.skip_fp16) else (0, w["precision"])
        runs = precision_runs(P, flag, w, precs, args.gate_steps, lattices=kept)
        precision = {"compare": f"stored precision vs f64, both half-way bounce-back, {args.gate_steps} steps "
                                "from rest on this workload; FP16 with stochastic rounding",
                     "K_f64": runs[0]["K"], "mass_rel_drift_f64": runs[0]["mass_drift"], "gate_K_rel_err": 1e-4}
        for prec in precs[1:]:
            err = errors_vs_fp64(runs, prec)
            precision[{1: "f32", 2: "f16"}[prec]] = {
                "K": runs[prec]["K"], "K_rel_err": err["K_rel_err"], "pass": err["K_rel_err"] <= 1e-4,
                "velocity_rel_l2_err": err["velocity_rel_l2_err"], "mass_rel_drift": runs[prec]["mass_drift"]}
        precision["wall_s"] = round(time.perf_counter() - t0, 1)
        # The gate's FP64 run is the fastest FP64 form (TILE8 index, half-way
        # bounce-back), timed by device events over its steps: the line also
        # states its speedup over that one, beside the plain FP64 path's.
        secondary["fp64_tile8_halfway_mlups"] = round(runs[0]["mlups"], 1)
        secondary["speedup_vs_fp64_tile8_halfway"] = round(mlups / runs[0]["mlups"], 3)
        del runs
    for lat in kept.values():
        lat.close()

    out = {
        "metric": "pore-node MLUPS",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_run["wall_s"] / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {0: "f64", 1: "f32", 2: "f16-store/f32"}[w["precision"]],
        "data": "synthetic (seeded periodic mask)",
        "config": {"workload": w["label"], "pore_nodes": n, "storage": PREC_NAME[w["precision"]],
                   "bounce_back": "halfway" if w["bb"] else "reference order",
                   "pull_table": "TILE8" if shape["index"] else "int32",
                   "layout": "AoSoA64" if shape["layout"] else "SoA",
                   "node_order": ["reference", "bricks", "pencils"][shape["ordering"]],
                   "brick": list(w.get("brick", (16, 4, 4))), "block": shape["block"],
                   "waves_per_simd": shape["waves"], "placement_reserve_gib": w.get("reserve_gib", 0),
                   "mask_seed": w["seed"], "flag_bin_sha256": mask_sha256},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": (f"{tname}, rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, measured at commit "
                                        f"{traffic_commit}" if traffic else None),
                     "algorithmic_bytes_per_update": algo_bpu, "kernel_ms": round(avg_ms, 4),
                     "layout_bytes_per_update": round(bpu, 3), "layout_gbs": round(layout_gbs, 1),
                     "traffic_gbs": round(traffic / (avg_ms * 1e-3) / 1e9, 1) if traffic else None,
                     "traffic_frac_of_peak": (round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                              if traffic else None),
                     "copy_ceiling_gbs": round(copy_gbs, 1),
                     "traffic_frac_of_copy": (round(traffic / (avg_ms * 1e-3) / 1e9 / copy_gbs, 4)
                                              if traffic else None)},
        "cpu_baseline": cpu,
        "secondary": secondary,
        "precision": precision,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
