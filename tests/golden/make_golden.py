"""Writes the golden fixtures in tests/golden from the reference programs.

The reference's preprocess.c and porelb.c are compiled from their own
sources (``make -C oracle ref`` -> oracle/_ref) and run here, in the build
container, through the harnesses in oracle/ (same call sequence as the
reference main, tau / body force / save interval from the command line).
Their outputs are stored as data: V.bin semantic fields, V_sub_start, final
populations f0 (AoS, V order), final macros, and for the 64^3 case the
SHA-256 of those arrays.  Run:

    make -C oracle all ref && make -C porelb_amd libplb.so
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

import oracle as O  # noqa: E402
import porelb_amd as P  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def tiny_mask(dims, p, seed):
    return (np.random.default_rng(seed).random(dims) < p).astype(np.uint8)


def vfields(V):
    x, y, z, t, pid, nid = O.semantic_fields(V)
    return dict(x=x.astype(np.uint16), y=y.astype(np.uint16), z=z.astype(np.uint16),
                type=t.astype(np.uint8), pid=pid.astype(np.uint16), id=nid.astype(np.uint32))


def tiny_cases():
    return [
        # name, dims (nz,ny,nx), pore prob, seed, ranks, tau, gx, steps
        ("tiny_a", (8, 7, 6), 0.6, 1, [1, 2, 4], 0.8, 1e-3, 30),
        ("tiny_b", (9, 5, 6), 0.7, 2, [1, 3], 1.0, 1e-5, 50),
        ("tiny_c", (4, 4, 4), 1.01, 3, [1, 2], 0.6, 1e-4, 20),   # all-fluid box
        ("tiny_d", (6, 6, 12), 0.45, 4, [1, 2, 3], 0.9, 5e-4, 40),
    ]


def make_tiny():
    for name, dims, p, seed, ranks, tau, gx, steps in tiny_cases():
        flag = tiny_mask(dims, p, seed)
        out = dict(flag=flag, tau=tau, gx=gx, steps=steps, ranks=np.array(ranks))
        for np_ in ranks:
            r = O.run_reference(flag, np_, steps, tau, gx, save_every=1000)
            vb = r["vbin"]
            for k, v in vfields(vb["V"]).items():
                out[f"np{np_}_V_{k}"] = v
            out[f"np{np_}_start"] = vb["start"]
            out[f"np{np_}_f0"] = r["f0"]
            out[f"np{np_}_macros"] = r["macros"]
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(name, dims, "nodes", len(vb["V"]))


def make_fault():
    """A body force large enough to trip the density guard (porelb.c:338-342):
    the reference's exit status and message are the fixture."""
    flag = tiny_mask((8, 8, 8), 0.55, 7)
    wd = tempfile.mkdtemp(prefix="plbfault_")
    flag.tofile(os.path.join(wd, "flag.bin"))
    subprocess.run([O.REF_PREPROCESS, "1", "8", "8", "8", "flag.bin", "V.bin"], cwd=wd, check=True,
                   capture_output=True)
    r = subprocess.run([O.MPIRUN, "-np", "1", O.REF_PORELB, "200", "0.8", "0.5", "-1", "0"], cwd=wd,
                       capture_output=True, text=True)
    msg = [l for l in r.stderr.splitlines() if "Macro denstity" in l]
    info = dict(returncode=r.returncode, message=msg[0] if msg else "")
    np.savez_compressed(os.path.join(HERE, "fault.npz"), flag=flag, tau=0.8, gx=0.5,
                        returncode=r.returncode, message=np.array(info["message"]))
    print("fault", info)


def make_config1():
    """64^3 smoothed noise (sigma 2, porosity 0.35, seed 1); tau 1.0, g_x 1e-6;
    1000 steps on 1 rank, and 200 steps on 4 ranks."""
    flag = P.mask_noise((64, 64, 64), 2.0, 0.35, 1)
    meta = dict(mask=dict(kind="noise", dims=[64, 64, 64], sigma=2.0, porosity=0.35, seed=1,
                          sha256=sha(flag), pores=int(flag.sum())))
    for np_, steps in [(1, 1000), (4, 200)]:
        r = O.run_reference(flag, np_, steps, 1.0, 1e-6, save_every=1000)
        f0, mac = r["f0"], r["macros"]
        idx = np.random.default_rng(0).choice(f0.shape[0], 64, replace=False)
        meta[f"np{np_}"] = dict(steps=steps, f0_sha256=sha(f0), macros_sha256=sha(mac),
                                mass=float(f0.sum()), sum_ux=float(mac[1].sum()),
                                sample_idx=idx.tolist(), sample_f0=f0[idx].tolist(),
                                sample_ux=mac[1][idx].tolist(),
                                seconds=r.get("timing", {}).get("seconds"))
        print("config1 np", np_, "nodes", f0.shape[0], meta[f"np{np_}"]["seconds"])
    with open(os.path.join(HERE, "config1.json"), "w") as fh:
        json.dump(meta, fh, indent=1)


if __name__ == "__main__":
    if not O.have_reference_build():
        sys.exit("oracle/_ref missing: run `make -C oracle ref` (needs the reference sources)")
    make_tiny()
    make_fault()
    make_config1()
