"""Build recipe of the oracle library (g++, OpenMP).

Test infrastructure: compiles oracle/oracle.cpp into oracle/liboracle.so.
"""

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "oracle.cpp"
OUT = HERE / "liboracle.so"

# x86-64-v3 (AVX2 + FMA) keeps the binary portable between this container and
# the GPU hosts; no -ffast-math so the double instantiation stays IEEE.
FLAGS = ["-O3", "-std=c++20", "-fopenmp", "-fPIC", "-shared", "-march=x86-64-v3"]


def build(verbose: bool = False) -> Path:
  cmd = ["g++", *FLAGS, str(SRC), "-o", str(OUT)]
  if verbose:
    print(" ".join(cmd), flush=True)
  subprocess.run(cmd, check=True)
  return OUT


if __name__ == "__main__":
  build(verbose="-v" in sys.argv)
