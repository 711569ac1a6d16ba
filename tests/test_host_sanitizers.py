"""The host structure builders (threaded colouring and symmetry check, level
sets, BEllPack, RB-CSR, RB-DIA) built with AddressSanitizer + UndefinedBehavior
Sanitizer and with ThreadSanitizer from their sources by g++, and run on a 64^3
7-point grid: no sanitizer report, the builders' invariants hold."""
import os
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
CSRC = ROOT / "alien_amd" / "csrc"


@pytest.mark.parametrize("san", ["address,undefined", "thread"])
def test_host_builders_clean_under_sanitizers(tmp_path, san):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "host_sanitize"
    cmd = [gxx, "-O1", "-g", "-std=c++17", "-pthread", f"-fsanitize={san}", "-fno-sanitize-recover=all",
           f"-I{CSRC}", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "host_sanitize_main.cpp"),
           str(CSRC / "host" / "structure.cpp"), "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if b.returncode != 0 and "cannot find" in b.stderr:
        pytest.skip(f"sanitizer runtime not installed: {b.stderr.strip()[:200]}")
    assert b.returncode == 0, b.stderr
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
