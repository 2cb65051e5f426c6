"""Build recipe for libracer_hip.so (gfx950) and the racer CLI.

Every .hip file under csrc/ is compiled by hipcc for gfx950 and linked into
one shared library next to this file. Flags:
  -ffp-contract=off        float32 expressions stay in source order, so the
                           simulation matches the CPU restatement bit for bit;
  -packed-fp32-ops off     no v_pk_* float32 instructions;
  (defaults kept)          correctly rounded float32 division and sqrt.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libracer_hip.so"
CLI = HERE / "racer"
BUILD = HERE / "_build"
ARCH = os.environ.get("RACER_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-result",
         "-Wno-unused-command-line-argument", f"-I{HERE.parent / 'include'}"]


def _compile(src: Path, verbose: bool) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False, jobs: int = 4) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(lambda s: _compile(s, verbose), sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    tmp.replace(LIB)
    cli_src = CSRC / "racer_cli.cpp"
    if cli_src.exists():
        cmd = [HIPCC, "-O2", "-std=c++17", f"-I{HERE.parent / 'include'}", str(cli_src), "-o", str(CLI),
               f"-L{HERE}", "-lracer_hip", f"-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
