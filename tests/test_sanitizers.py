"""The host C of the command lines (mask generators, file I/O) under
AddressSanitizer, LeakSanitizer and UndefinedBehaviorSanitizer, built with
gcc; no GPU code is involved."""
import os
import shutil
import subprocess

import pytest

from conftest import REPO


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc absent")
def test_host_code_clean_under_sanitizers(tmp_path):
    exe = str(tmp_path / "host_sanitize")
    host = os.path.join(REPO, "porelb_amd", "host")
    cmd = ["gcc", "-std=c11", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I" + os.path.join(REPO, "include"), "-I" + host,
           os.path.join(REPO, "tests", "host_sanitize.c"), os.path.join(host, "plb_masks.c"),
           os.path.join(host, "plb_files.c"), "-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
