"""The C ABI library loads and exports every symbol include/plb.h declares.
No compute call is made here without a GPU."""
import ctypes as C
import os
import subprocess

import pytest

from conftest import REPO, have_gpu_device


def test_library_exports_every_declared_symbol(plb):
    declared = plb.header_symbols()
    assert len(declared) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", plb.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing


def test_config_default_is_reference_parameters(plb):
    cfg = plb.default_config()
    assert cfg.tau == 0.8 and cfg.gx == 1e-3 and cfg.rho0 == 1.0  # porelb.c:87-91
    assert cfg.precision == plb.FP64 and cfg.bb_mode == plb.BB_REFCOMPAT
    assert cfg.block_size == 256


def test_strerror_codes(plb):
    L = plb.lib()
    assert b"density" in L.plb_strerror(-6)
    assert L.plb_strerror(0) == b"ok"
    assert L.plb_strerror(12345) == b"unknown error"


def test_invalid_config_rejected_before_device(plb):
    L = plb.lib()
    h = C.c_void_p()
    flag = (C.c_uint8 * 8)(*([1] * 8))
    for field, value in [("tau", 0.0), ("precision", 7), ("block_size", 100), ("layout", 3),
                         ("waves_per_simd", 9), ("waves_per_simd", -1), ("graph_steps", 3)]:
        cfg = plb.default_config(**{field: value})
        rc = L.plb_create_from_flag(C.byref(h), C.byref(cfg), flag, 2, 2, 2, 1)
        assert rc == -20, field
    cfg = plb.default_config()
    assert L.plb_create_from_flag(C.byref(h), C.byref(cfg), flag, 2, 2, 2, 3) == -2  # z-slab rule


@pytest.mark.skipif(have_gpu_device(), reason="checks the no-GPU error path")
def test_no_gpu_fails_loudly(plb):
    """There is no CPU fallback: without a GPU creation reports an error."""
    import numpy as np

    flag = np.ones((4, 4, 4), dtype=np.uint8)
    with pytest.raises(plb.PlbError) as ei:
        plb.Lattice.from_flag(flag)
    assert ei.value.code in (-24, -21)


def test_ctypes_structs_match_the_header(plb, tmp_path):
    """The Python mirrors of the C structs have the header's size and field
    offsets (compiled and printed by gcc from include/plb.h and include/plb_tune.h)."""
    structs = {"plb_config": plb.Config, "plb_reduction": plb.Reduction, "plb_fault": plb.Fault,
               "plb_sample": plb.Sample, "plb_tune_result": plb.TuneResult,
               "plb_state_info": plb.StateInfo}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "plb.h"', '#include "plb_tune.h"',
             "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        name, field, val = line.split()
        got[(name, field)] = int(val)
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
