"""Writes tests/golden/config2_20000.json: config 2 as specified (256^3 sphere
pack, radii U[6,12], porosity 0.30, seed 2; tau 0.8, g_x 1e-6; 20 000 steps)
run by the reference porelb.c built from its own sources (oracle/_ref) under
`mpirun -np 16`.  Stores the SHA-256 of the final populations (V order,
[N][19] float64) and of the final part-file macros ([4][N] float64), plus a
few sums, so that a GPU run can be checked without re-running the reference.

Run from the repository root on a machine with the reference build:
    python tests/golden/make_config2_golden.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402
import porelb_amd as P  # noqa: E402

DIMS, RMIN, RMAX, PHI, SEED = (256, 256, 256), 6.0, 12.0, 0.30, 2
TAU, GX, STEPS, RANKS = 0.8, 1e-6, 20000, 16


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    flag = P.mask_spheres(DIMS, RMIN, RMAX, PHI, SEED)
    print("mask", sha(flag), int(flag.sum()), flush=True)
    t0 = time.time()
    ref = O.run_reference(flag, RANKS, STEPS, TAU, GX, save_every=1000, dump=True, timeout=6 * 3600)
    out = {
        "mask": {"kind": "spheres", "dims": list(DIMS), "rmin": RMIN, "rmax": RMAX, "porosity": PHI,
                 "seed": SEED, "sha256": sha(flag), "pores": int(flag.sum())},
        "tau": TAU, "gx": GX, "steps": STEPS, "ranks": RANKS,
        "f0_sha256": sha(ref["f0"]), "macros_sha256": sha(ref["macros"]),
        "mass": float(ref["f0"].sum()), "sum_ux": float(ref["macros"][1].sum()),
        "reference_timing": ref.get("timing"), "host_cpus": os.cpu_count(),
        "wall_s": round(time.time() - t0, 1),
    }
    path = os.path.join(ROOT, "tests", "golden", "config2_20000.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
