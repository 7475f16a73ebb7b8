"""Build recipe for libkbgpu.so (gfx950) and the CPU oracle used by the tests.

    python kitbase.jl_amd/build.py            # both
    python kitbase.jl_amd/build.py --lib      # only libkbgpu.so

The HIP sources are compiled one object per file in parallel with hipcc, then linked into
kitbase.jl_amd/lib/libkbgpu.so.  Packed FP32 instructions are disabled for every object.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "lib" / "libkbgpu.so"
ORACLE = ROOT / "oracle"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("KBG_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]
# hipFFT for the fast spectral collision's batched 3D transforms (ROCm's own library, found through
# the image's /opt/rocm)
LINK = ["-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"]
SOURCES = ["kbg_sweep_f32.hip", "kbg_sweep_f32b.hip", "kbg_sweep_f32c.hip", "kbg_sweep_f64.hip", "kbg_sweep_f64c.hip", "kbg_1d.hip", "kbg_2d.hip", "kbg_step2d.hip", "kbg_us.hip", "kbg_fsm.hip", "kbg_api.hip"]


def _run(cmds: list[list[str]]) -> None:
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)) for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            failed.append((c, "\n".join(lines)))
        elif lines:
            print("\n".join(lines), file=sys.stderr)
    if failed:
        msg = "\n\n".join(" ".join(c) + "\n" + o for c, o in failed)
        raise RuntimeError(f"build failed:\n{msg}")


def build_lib() -> Path:
    BUILD.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    objs, cmds = [], []
    for src in SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    _run(cmds)
    tmp = LIB.with_suffix(".so.tmp")
    _run([[HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), *LINK]])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    _run([["make", "-B", "-C", str(ORACLE), "-s"]])
    return ORACLE / "liboracle.so"


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="only libkbgpu.so")
    ap.add_argument("--oracle", action="store_true", help="only the oracle")
    a = ap.parse_args()
    if not a.oracle:
        print(build_lib())
    if not a.lib:
        print(build_oracle())


if __name__ == "__main__":
    main()
