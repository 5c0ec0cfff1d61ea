"""Benchmark helpers (test_util.py:41-120): `benchmark` times a batched
function from a HIP graph replay, like the reference's CUDA-graph loop.

`benchmark(fn, mjm, mjd, ...)` sets the solver options on the host model,
builds the device Model / Data with `put_model` / `put_data`, calls `fn`
twice (the first call loads the library; there is no JIT), captures one
call of `fn` as a HIP graph and replays it `nstep` times, returning
`(first-call time, run time, trace, batch_size * nstep, ncon, nefc)` like
the reference.  With `event_trace` the calls run eagerly under an
`EventTracer` instead of from the graph (the per-stage events need the
per-stage entry points, which the one-call step does not have), and the
traces of the calls are summed.  With `measure_alloc` the contact and row
counts are read back after every call.
"""

from __future__ import annotations

import time
from typing import Callable, Tuple

import numpy as np

from . import api
from . import io
from . import tracing
from . import types


def _sum(stack1: dict, stack2: dict) -> dict:
  """Element-wise sum of two traces with the same keys (test_util.py:41-49)."""
  ret = {}
  for k in stack1:
    times1, sub1 = stack1[k]
    times2, sub2 = stack2[k]
    ret[k] = ([t1 + t2 for t1, t2 in zip(times1, times2)], _sum(sub1, sub2))
  return ret


def forward_host(mjm, mjd, nconmax: int = 4096, njmax: int = 16384) -> None:
  """Fills mjd (one world) with a device forward pass: the counterpart of
  mujoco.mj_forward on the host data, including ncon / nefc and the contact
  and constraint rows."""
  import torch

  m = io.put_model(mjm)
  d = io.make_data(mjm, nworld=1, nconmax=nconmax, njmax=njmax)
  for name in ("qpos", "qvel", "act", "ctrl", "qacc_warmstart", "qfrc_applied"):
    v = np.asarray(getattr(mjd, name), np.float32).reshape(1, -1)
    if v.size:
      getattr(d, name).copy_(torch.as_tensor(v))
  api.forward(m, d)
  torch.cuda.synchronize()
  io.get_data_into(mjd, mjm, d)


def benchmark(
  fn: Callable[[types.Model, types.Data], None],
  mjm,
  mjd,
  nstep: int = 1000,
  batch_size: int = 1024,
  solver: str = "cg",
  iterations: int = 1,
  ls_iterations: int = 4,
  ls_parallel: bool = False,
  nconmax: int = -1,
  njmax: int = -1,
  event_trace: bool = False,
  measure_alloc: bool = False,
  graph: bool = True,
) -> Tuple[float, float, dict, int, list, list]:
  """Times nstep calls of fn on batch_size copies of mjd (test_util.py:88-120).
  graph=False (no counterpart in the reference) calls fn eagerly instead of
  replaying a captured graph, e.g. for counter collection by a profiler."""
  import torch

  if solver == "cg":
    mjm.opt.solver = int(types.SolverType.CG)
  elif solver == "newton":
    mjm.opt.solver = int(types.SolverType.NEWTON)
  mjm.opt.iterations = iterations
  mjm.opt.ls_iterations = ls_iterations

  m = io.put_model(mjm)
  m.opt.ls_parallel = ls_parallel
  d = io.put_data(mjm, mjd, nworld=batch_size, nconmax=nconmax, njmax=njmax)

  jit_beg = time.perf_counter()
  fn(m, d)
  fn(m, d)
  torch.cuda.synchronize()
  jit_duration = time.perf_counter() - jit_beg

  trace = {}
  ncon, nefc = [], []
  if event_trace:
    run_beg = time.perf_counter()
    for _ in range(nstep):
      # one tracer per call, so that each call's events are summed once
      with tracing.EventTracer(enabled=True) as tracer:
        fn(m, d)
        step_trace = tracer.trace()
      trace = _sum(trace, step_trace) if trace else step_trace
      if measure_alloc:
        ncon.append(int(d.ncon[0].item()))
        nefc.append(int(d.nefc[0].item()))
    torch.cuda.synchronize()
    run_duration = time.perf_counter() - run_beg
  elif not graph:
    torch.cuda.synchronize()
    run_beg = time.perf_counter()
    for _ in range(nstep):
      fn(m, d)
      if measure_alloc:
        ncon.append(int(d.ncon[0].item()))
        nefc.append(int(d.nefc[0].item()))
    torch.cuda.synchronize()
    run_duration = time.perf_counter() - run_beg
  else:
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(stream):
      with torch.cuda.graph(graph, stream=stream):
        fn(m, d)
    torch.cuda.current_stream().wait_stream(stream)
    torch.cuda.synchronize()
    run_beg = time.perf_counter()
    for _ in range(nstep):
      graph.replay()
      if measure_alloc:
        torch.cuda.synchronize()
        ncon.append(int(d.ncon[0].item()))
        nefc.append(int(d.nefc[0].item()))
    torch.cuda.synchronize()
    run_duration = time.perf_counter() - run_beg

  return jit_duration, run_duration, trace, batch_size * nstep, ncon, nefc
