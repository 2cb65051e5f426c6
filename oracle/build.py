"""Build recipe for the CPU oracle (test infrastructure): compiles
racer_oracle.c with gcc into oracle/liboracle.so.

Flags: -ffp-contract=off keeps float32 expressions in source order (no fused
multiply-adds the source does not write), -fno-fast-math keeps IEEE semantics,
-mfma lets gcc inline the explicit fmaf calls (and vectorise independent fmaf
chains, each lane still one correctly rounded fused multiply-add), -fopenmp
threads the per-car loops for the CPU baseline.
"""
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build(verbose: bool = False) -> Path:
    cmd = ["gcc", "-O3", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
           "-mfma", "-Wall", "-Wno-unknown-pragmas", str(HERE / "racer_oracle.c"), "-o", str(LIB) + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    Path(str(LIB) + ".tmp").replace(LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
