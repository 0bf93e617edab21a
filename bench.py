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

WORKLOADS = {
    # name: mask kind, dims (nz,ny,nx), mask params, tau, gx, precision, bb
    "spheres1024-fp32": dict(kind="spheres", dims=(1024, 1024, 1024), rmin=6.0, rmax=12.0, porosity=0.20,
                             seed=4, tau=0.6, gx=5e-7, precision=1, bb=1,
                             label="1024^3 sphere pack, porosity 0.20, tau 0.6, g_x 5e-7"),
    "spheres1024-fp16": dict(kind="spheres", dims=(1024, 1024, 1024), rmin=6.0, rmax=12.0, porosity=0.20,
                             seed=4, tau=0.6, gx=5e-7, precision=2, bb=1,
                             label="1024^3 sphere pack, porosity 0.20, tau 0.6, g_x 5e-7"),
    "spheres1024-fp64": dict(kind="spheres", dims=(1024, 1024, 1024), rmin=6.0, rmax=12.0, porosity=0.20,
                             seed=4, tau=0.6, gx=5e-7, precision=0, bb=0,
                             label="1024^3 sphere pack, porosity 0.20, tau 0.6, g_x 5e-7"),
    "spheres256-fp64": dict(kind="spheres", dims=(256, 256, 256), rmin=6.0, rmax=12.0, porosity=0.30,
                            seed=2, tau=0.8, gx=1e-6, precision=0, bb=0,
                            label="256^3 sphere pack, porosity 0.30, tau 0.8, g_x 1e-6"),
    "noise512-fp32": dict(kind="noise", dims=(512, 512, 512), sigma=3.0, porosity=0.25, seed=3,
                          tau=1.0, gx=1e-6, precision=1, bb=1,
                          label="512^3 smoothed noise sigma 3, porosity 0.25, tau 1.0, g_x 1e-6"),
}
PREC_NAME = {0: "f64", 1: "f32-store/f32-collide", 2: "f16-store/f32-collide"}


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
              brick=(16, 4, 4)):
    cfg = P.default_config(tau=w["tau"], gx=w["gx"], precision=precision, bb_mode=bb, layout=layout,
                           ordering=ordering, block_size=block, index_kind=index, brick=brick)
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
    finally:
        lat.close()
    return dict(nodes=n, wall_s=wall, event_ms=ev_ms, bytes_per_update=bpu, reduction=red)


def cpu_baseline(P, w, max_seconds=20.0):
    """The reference porelb.x (oracle/_ref, built from the reference's sources)
    under mpirun on all usable host cores, timed by its own MPI_Wtime loop
    timer, on a 256^3 domain of the same kind and porosity."""
    import oracle as O

    cores = min(16, os.cpu_count() or 1)
    ranks = 1
    while ranks * 2 <= cores:
        ranks *= 2
    dims = (256, 256, 256)
    sample = dict(w, dims=dims)
    flag = make_mask(P, sample)
    npore = int(flag.sum())
    if O.have_reference_build() and os.path.exists(O.MPIRUN):
        per_rank_mlups = 8.0
        steps = int(max(10, min(400, max_seconds * ranks * per_rank_mlups * 1e6 / npore)))
        wd = tempfile.mkdtemp(prefix="plbcpu_")
        log(f"cpu baseline: reference porelb.x, mpirun -np {ranks}, {npore} nodes, {steps} steps")
        r = O.run_reference(flag, ranks, steps, w["tau"], w["gx"], save_every=1000, print_steps=1,
                            workdir=wd, dump=False, timeout=900)
        subprocess.run(["rm", "-rf", wd])
        t = r["timing"]
        return {"value": round(t["mlups"], 3), "unit": "pore-node MLUPS", "cores": ranks, "kind": "reference",
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
    return {"value": round(len(V) * steps / dt / 1e6, 3), "unit": "pore-node MLUPS", "cores": 1,
            "kind": "port", "sample": f"{dims[0]}^3 domain, {len(V)} pore nodes, {steps} FP64 steps, 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="spheres1024-fp32", choices=sorted(WORKLOADS))
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--layout", type=int, default=0)
    ap.add_argument("--ordering", type=int, default=0)
    ap.add_argument("--block", type=int, default=256)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the domain of every configuration fits one MI355X)", file=sys.stderr)
        sys.exit(2)

    import porelb_amd as P

    w = WORKLOADS[args.workload]
    cpu = None
    if not args.skip_cpu:
        try:
            cpu = cpu_baseline(P, w)
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline failed:", repr(e))
    t0 = time.perf_counter()
    flag = make_mask(P, w)
    log(f"mask {w['dims']} porosity {flag.mean():.4f} in {time.perf_counter() - t0:.1f} s")
    # Practical ceiling: a streaming copy of 4 GiB (its rocprofv3 FETCH_SIZE /
    # WRITE_SIZE also calibrate the counter scale for the step kernel).
    copy_gbs = P.copy_bandwidth(4 << 30, 5)
    log(f"copy ceiling {copy_gbs:.0f} GB/s")

    main_run = time_path(P, flag, w, w["precision"], w["bb"], args.steps, args.warmup, args.layout,
                         args.ordering, args.block)
    n, steps = main_run["nodes"], args.steps
    mlups = n * steps / main_run["wall_s"] / 1e6
    avg_ms = main_run["event_ms"] / steps
    bpu = main_run["bytes_per_update"]
    achieved = bpu * n / (avg_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(REPO, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")

    secondary = {}
    if not args.skip_fp64 and w["precision"] != 0:
        fp64 = time_path(P, flag, w, 0, 0, max(20, steps // 2), max(5, args.warmup // 2))
        fp64_mlups = fp64["nodes"] * max(20, steps // 2) / fp64["wall_s"] / 1e6
        secondary = {"fp64_plain_mlups": round(fp64_mlups, 1),
                     "fp64_plain_roofline_frac": round(fp64["bytes_per_update"] * fp64["nodes"] /
                                                       (fp64["event_ms"] / max(20, steps // 2) * 1e-3) /
                                                       1e9 / HBM_PEAK_GBS, 4),
                     "speedup_vs_fp64_plain": round(mlups / fp64_mlups, 3)}

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
        "dtype": "f64" if w["precision"] == 0 else "f32",
        "data": "synthetic (seeded periodic mask)",
        "config": {"workload": w["label"], "pore_nodes": n, "storage": PREC_NAME[w["precision"]],
                   "bounce_back": "halfway" if w["bb"] else "reference order", "layout": args.layout,
                   "ordering": args.ordering, "block": args.block},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_update": bpu, "kernel_ms": round(avg_ms, 4),
                     "copy_ceiling_gbs": round(copy_gbs, 1), "frac_of_copy": round(achieved / copy_gbs, 4)},
        "cpu_baseline": cpu,
        "secondary": secondary,
        "K": main_run["reduction"]["K"],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
