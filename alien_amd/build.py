"""Build libalien_hip.so (gfx950) in-tree with hipcc.

Usage: python -m alien_amd.build [--jobs N]

Every .hip source is compiled for gfx950 with -ffp-contract=off (separate
multiply and add, so the kernels round like the sequential reference loops)
and with packed FP32 instructions disabled; host-only .cpp sources go through
the same compiler. Objects land in alien_amd/_build/, the library in
alien_amd/libalien_hip.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libalien_hip.so"
ROOT = PKG.parent

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", f"-I{ROOT / 'include'}", f"-I{CSRC}",
          "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources() -> list[pathlib.Path]:
    out = sorted(CSRC.rglob("*.hip")) + sorted(CSRC.rglob("*.cpp"))
    return out


def _obj(src: pathlib.Path) -> pathlib.Path:
    rel = src.relative_to(CSRC).with_suffix(".o")
    return BUILD / str(rel).replace(os.sep, "__")


def _compile(src: pathlib.Path) -> tuple[pathlib.Path, str]:
    obj = _obj(src)
    cmd = [HIPCC] + COMMON + (DEVICE if src.suffix == ".hip" else []) + ["-c", str(src), "-o", str(obj)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    warn = "\n".join(l for l in (p.stderr or "").splitlines() if "packed-fp32-ops" not in l and l.strip())
    return obj, warn


def build(jobs: int | None = None, verbose: bool = False) -> pathlib.Path:
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, warn in ex.map(_compile, srcs):
            objs.append(obj)
            if verbose and warn:
                print(warn, file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp)] + [str(o) for o in objs]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
