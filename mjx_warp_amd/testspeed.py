"""The testspeed command line (testspeed.py:30-219):

    python -m mjx_warp_amd.testspeed --mjcf=humanoid.xml --batch_size=8192

Same flags, defaults and report as the reference; the timing is
`mjx_warp_amd.benchmark` (a HIP graph of one call, replayed `--nstep`
times).  The host model comes from the MJCF compiler
(`mjcf.from_xml_path`); a relative `--mjcf` path is looked up in
`--base_path`, by default the package's model assets.  The reference's
`mj_forward` on the host data (testspeed.py:96-97), which sizes the default
contact / row pools, is computed by one device `forward` of a single world.
"""

from __future__ import annotations

import argparse
import inspect
import sys
from pathlib import Path
from typing import Sequence

import numpy as np

from . import mjcf
from .test_util import benchmark, forward_host


def _functions() -> list:
  """The package's (m, d) functions, the choices of --function."""
  import mjx_warp_amd

  out = []
  for name, f in inspect.getmembers(mjx_warp_amd, inspect.isfunction):
    params = list(inspect.signature(f).parameters)
    if params[:2] == ["m", "d"] and len(params) == 2:
      out.append(name)
  return sorted(out)


def _bool(v: str) -> bool:
  if v.lower() in ("1", "true", "yes", "y", "t"):
    return True
  if v.lower() in ("0", "false", "no", "n", "f"):
    return False
  raise argparse.ArgumentTypeError(f"not a boolean: {v}")


def _parser() -> argparse.ArgumentParser:
  p = argparse.ArgumentParser(prog="python -m mjx_warp_amd.testspeed", description=__doc__.split("\n\n")[0].strip())
  p.add_argument("--function", default="step", choices=_functions(), help="the function to run")
  p.add_argument("--mjcf", required=True, help="path to model `.xml`")
  p.add_argument("--base_path", default=None, help="base path of a relative --mjcf (default: the package's assets)")
  p.add_argument("--nstep", type=int, default=1000, help="number of steps per rollout")
  p.add_argument("--batch_size", type=int, default=4096, help="number of parallel rollouts")
  p.add_argument("--solver", default="newton", choices=["cg", "newton"], help="constraint solver")
  p.add_argument("--iterations", type=int, default=1, help="number of solver iterations")
  p.add_argument("--ls_iterations", type=int, default=4, help="number of linesearch iterations")
  p.add_argument("--ls_parallel", type=_bool, nargs="?", const=True, default=False,
                 help="solve with parallel linesearch")
  p.add_argument("--is_sparse", type=_bool, nargs="?", const=True, default=False,
                 help="if model should create sparse mass matrices")
  p.add_argument("--nconmax", type=int, default=-1, help="Maximum number of contacts in a batch physics step.")
  p.add_argument("--njmax", type=int, default=-1, help="Maximum number of constraints in a batch physics step.")
  p.add_argument("--keyframe", type=int, default=0, help="Keyframe to initialize simulation.")
  p.add_argument("--output", default="text", choices=["text", "tsv"], help="format to print results")
  p.add_argument("--clear_kernel_cache", type=_bool, nargs="?", const=True, default=False,
                 help="accepted for compatibility; kernels are compiled ahead of time")
  p.add_argument("--event_trace", type=_bool, nargs="?", const=True, default=False, help="Provide a full event trace")
  p.add_argument("--measure_alloc", type=_bool, nargs="?", const=True, default=False,
                 help="Measure how much of nconmax, njmax is used.")
  p.add_argument("--graph", type=_bool, nargs="?", const=True, default=True,
                 help="replay a captured HIP graph (false: call the function eagerly, e.g. under a profiler)")
  return p


def _print_trace(trace: dict, indent: int, steps: int) -> None:
  for k, (times, sub) in trace.items():
    if len(times) == 1:
      print("  " * indent + f"{k}: {1e6 * times[0] / steps:.2f}")
    else:
      print("  " * indent + f"{k}: [ " + ", ".join(f"{1e6 * t / steps:.2f}" for t in times) + " ]")
    _print_trace(sub, indent + 1, steps)


def _print_table(matrix, headers) -> None:
  n = len(headers)
  widths = [max(max(len(f"{row[i]:g}") for row in matrix), len(headers[i])) for i in range(n)]
  print("  ".join(f"{headers[i]:<{widths[i]}}" for i in range(n)))
  print("-" * sum(widths) + "--" * 3)
  for row in matrix:
    print("  ".join(f"{row[i]:{widths[i]}g}" for i in range(n)))


def main(argv: Sequence[str] | None = None) -> int:
  args = _parser().parse_args(argv)
  import mjx_warp_amd

  path = Path(args.mjcf)
  if not path.is_absolute():
    path = Path(args.base_path) / path if args.base_path else Path(__file__).resolve().parent / "assets" / path
  m = mjcf.from_xml_path(str(path))
  if args.is_sparse:
    raise NotImplementedError("sparse mass matrix (--is_sparse) is not implemented")
  m.opt.jacobian = 0
  d = mjcf.MjData(m)
  if m.nkey > 0:
    mjcf.mj_resetDataKeyframe(m, d, args.keyframe)
  # populate some constraints
  forward_host(m, d)

  print(f"Model nbody: {m.nbody} nv: {m.nv} ngeom: {m.ngeom} is_sparse: {args.is_sparse} solver: {args.solver}")
  print(f"Params nconmax: {args.nconmax} njmax: {args.njmax}")
  print(f"Data ncon: {d.ncon} nefc: {d.nefc} keyframe: {args.keyframe}")
  print(f"Linesearch: {'parallel' if args.ls_parallel else 'iterative'}")
  print(f"Rolling out {args.nstep} steps at dt = {m.opt.timestep:.3f}...")
  jit_time, run_time, trace, steps, ncon, nefc = benchmark(
    getattr(mjx_warp_amd, args.function), m, d, args.nstep, args.batch_size, args.solver, args.iterations,
    args.ls_iterations, args.ls_parallel, args.nconmax, args.njmax, args.event_trace, args.measure_alloc,
    graph=args.graph)

  if args.output == "text":
    print(f"""
Summary for {args.batch_size} parallel rollouts

 Total JIT time: {jit_time:.2f} s
 Total simulation time: {run_time:.2f} s
 Total steps per second: {steps / run_time:,.0f}
 Total realtime factor: {steps * m.opt.timestep / run_time:,.2f} x
 Total time per step: {1e9 * run_time / steps:.2f} ns""")
    if trace:
      print("\nEvent trace:\n")
      _print_trace(trace, 0, steps)
    if ncon and nefc:
      num_buckets = min(10, args.nstep)  # no empty buckets for short runs
      idx = 0
      ncon_matrix, nefc_matrix = [], []
      for i in range(num_buckets):
        size = args.nstep // num_buckets + (i < (args.nstep % num_buckets))
        c = np.array(ncon[idx:idx + size])
        e = np.array(nefc[idx:idx + size])
        ncon_matrix.append([np.mean(c), np.std(c), np.min(c), np.max(c)])
        nefc_matrix.append([np.mean(e), np.std(e), np.min(e), np.max(e)])
        idx += size
      print("\nncon alloc:\n")
      _print_table(ncon_matrix, ("mean", "std", "min", "max"))
      print("\nnefc alloc:\n")
      _print_table(nefc_matrix, ("mean", "std", "min", "max"))
  else:
    name = Path(sys.argv[0]).name.replace("testspeed_", "")
    print(f"{name}\tjit: {jit_time:.2f}s\tsteps/second: {steps / run_time:.0f}")
  return 0


if __name__ == "__main__":
  sys.exit(main())
