"""Build libphb.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

Usage: python -m ph_brush_amd.build  (or ph_brush_amd.build.build()).
Objects go to _build/ at the repository root; the library to
ph_brush_amd/libphb.so.  Every call recompiles (the build check must compile).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libphb.so"
CLI = PKG / "bin" / "MC"
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")

ARCH = os.environ.get("PHB_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off (see the build notes in the docs)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-ffp-contract=fast-honor-pragmas",
    "-Wall", "-Wno-unused-result",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def _sources():
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".hip":
        cmd[1:1] = ["-x", "hip"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    if res.stderr.strip():
        # surface warnings other than the known host-pass notice about packed-fp32-ops
        lines = [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            print(f"[build] {src.name}:\n" + "\n".join(lines), file=sys.stderr)
    return obj


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = _sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    build_cli()
    if verbose:
        print(f"[build] {LIB} ({len(objs)} objects, {ARCH}); {CLI}")
    return LIB


def build_cli() -> Path:
    """The command-line program (host code over libphb.so): ph_brush_amd/bin/MC."""
    CLI.parent.mkdir(exist_ok=True)
    tmp = CLI.with_name(CLI.name + ".tmp")
    cmd = [CXX, "-std=c++17", "-O2", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", str(PKG / "cli" / "mc_main.cpp"),
           f"-L{PKG}", "-lphb", "-Wl,-rpath,$ORIGIN/..", "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"CLI build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    build()
