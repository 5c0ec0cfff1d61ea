"""Build recipe of libmjx_warp_amd.so (hipcc, gfx950).

The device code is split over several compilations that run in parallel
into _build/ and link into one shared library: api.hip (host code and the
small kernels) and kernels.hip once per instantiation list of
csrc/stages.hpp (MJH_KERNELS_<NAME>, passed as -DMJH_KERNEL_LIST).
"""

from __future__ import annotations

import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libmjx_warp_amd.so"
OBJ = PKG / "_build"
HEADERS = sorted(CSRC.glob("*.hpp")) + [PKG.parent / "include" / "mjx_warp_amd.h"]
# the register solver's lists compile longest: start them first
_SLOW = ("BIG_STEP_G", "BIG_STEP", "BIG_SOLVE_G", "BIG_SOLVE", "FUSED_SOLVE_G", "FUSED_SOLVE")


def units() -> list[tuple[str, Path, list[str]]]:
  """(object name, source, extra flags) of every compilation."""
  lists = re.findall(r"^#define MJH_KERNELS_(\w+)\(X, G\)", (CSRC / "stages.hpp").read_text(), re.M)
  lists.sort(key=lambda n: (_SLOW.index(n) if n in _SLOW else len(_SLOW), n))
  out = [(f"kernels_{n.lower()}", CSRC / "kernels.hip", [f"-DMJH_KERNEL_LIST=MJH_KERNELS_{n}"]) for n in lists]
  return out + [("api", CSRC / "api.hip", [])]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
  "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
  # no packed FP32 instructions in device code
  "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _jobs() -> int:
  n = os.environ.get("MAX_JOBS")
  if n and n.isdigit():
    return max(1, int(n))
  return max(1, len(os.sched_getaffinity(0)))


def _compile(unit, verbose: bool) -> tuple[Path, str]:
  name, src, extra = unit
  obj = OBJ / (name + ".o")
  cmd = [HIPCC, *FLAGS, *extra, "-c", str(src), "-o", str(obj)]
  if verbose:
    print(" ".join(cmd), flush=True)
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    return obj, r.stdout + r.stderr
  return obj, ""


def build(verbose: bool = False, incremental: bool = False) -> Path:
  """Compiles every unit (only those older than a source or header with
  incremental=True) and links the library."""
  OBJ.mkdir(exist_ok=True)
  newest_header = max(h.stat().st_mtime for h in HEADERS)
  todo = []
  all_units = units()
  for u in all_units:
    obj = OBJ / (u[0] + ".o")
    if incremental and obj.exists() and obj.stat().st_mtime > max(newest_header, u[1].stat().st_mtime):
      continue
    todo.append(u)
  with ThreadPoolExecutor(_jobs()) as pool:
    results = list(pool.map(lambda u: _compile(u, verbose), todo))
  errors = [msg for _, msg in results if msg]
  if errors:
    sys.stderr.write("\n".join(errors))
    raise RuntimeError("hipcc failed")
  objs = [OBJ / (u[0] + ".o") for u in all_units]
  cmd = [HIPCC, *FLAGS, "-shared", "-Wl,-z,defs", *map(str, objs), "-o", str(OUT) + ".tmp"]
  if verbose:
    print(" ".join(cmd), flush=True)
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    sys.stderr.write(r.stdout + r.stderr)
    raise RuntimeError("hipcc link failed")
  os.replace(str(OUT) + ".tmp", OUT)
  return OUT


if __name__ == "__main__":
  build(verbose=True, incremental="--incremental" in sys.argv)
