"""Build recipe of libmjx_warp_amd.so (hipcc, gfx950)."""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libmjx_warp_amd.so"
SOURCES = [CSRC / "api.hip"]
HEADERS = sorted(CSRC.glob("*.hpp")) + [PKG.parent / "include" / "mjx_warp_amd.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
  "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
  # no packed FP32 instructions in device code
  "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False) -> Path:
  cmd = [HIPCC, *FLAGS, *map(str, SOURCES), "-o", str(OUT) + ".tmp"]
  if verbose:
    print(" ".join(cmd), flush=True)
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    sys.stderr.write(r.stdout + r.stderr)
    raise RuntimeError("hipcc failed")
  os.replace(str(OUT) + ".tmp", OUT)
  return OUT


if __name__ == "__main__":
  build(verbose=True)
