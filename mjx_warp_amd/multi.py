"""Heterogeneous batches: several models, each with its own worlds, stepped
together.

The reference batches worlds of one Model (types.py: every world of a Data
shares it).  `MultiStep([(m0, d0), (m1, d1), ...])` takes groups of
(Model, Data) -- a different model per group, any number of worlds each --
and advances every world of every group by one fused step per `step()`.
In the "grouped" mode the library runs one launch per stage over all groups
(`mjh_multi_step`: each workgroup finds its group in a device table), so
many small groups cost the launches of one batch; in the "per_group" mode
it issues one `mjh_step_fused` per group; the "static" mode runs the
grouped launches through `mjh_multi_step_static`, which keeps the group
table in a fixed device array of the library and allocates nothing per
call (for callers that cannot allocate, e.g. under graph capture; at most
64 groups, and calls on one GPU must not overlap).

Launch choices per group (`tunes`, the grouped mode): lanes per world (32:
two worlds per wavefront with the register solver, nv <= 32 only; 64: one
world per wavefront with the LDS solver), the LDS solver's row capacity and
the Jacobian rows staged in LDS -- (team, rmax, rj), 0 / 0 / -1 for the
library's choice; `buckets=True` launches groups of different LDS
occupancy classes separately, each launch sized for its own groups;
`queue=True` makes each launch a work queue (a grid of a few workgroups per
SIMD taking worlds, largest models first, from eight device counters).  Every
choice computes the same step.

`autotune=True` picks them by timing on the groups' own state (restored
after every trial): first each group alone over its candidate choices
(`candidates`), then the whole batch grouped with and without buckets and
work queues, and per group; the fastest is kept, remembered per group signature (each
group's sizes and world count) for the rest of the process.

The plan holds the groups' device pointers: write Data fields in place
(`d.qpos.copy_(...)`); a field replaced by a new tensor is picked up by
rebuilding the plan on the next `step()`.
"""

from __future__ import annotations

import ctypes
import time

from . import _native
from .types import Data, Model

_TUNED: dict = {}

_STATE = ("qpos", "qvel", "act", "ctrl", "qacc_warmstart")


def candidates(m) -> list:
  """The autotuner's (team, rmax, rj) choices for a model, the library's
  own first."""
  out = [(0, 0, -1)]
  if int(m.nv) <= 32:
    out += [(32, 0, 32), (32, 0, 64), (32, 0, 0)]
  out += [(64, 0, 32), (64, 0, 0), (64, 0, 64)]
  if int(m.nv) > 32:
    out.append((64, 0, 128))
  return out


def _signature(groups) -> tuple:
  return tuple((int(m.nq), int(m.nv), int(m.nbody), int(m.ngeom), int(m.nu), int(d.nworld), int(d.njmax))
               for m, d in groups)


class MultiStep:
  """One fused step of every world of every (Model, Data) group per call."""

  def __init__(self, groups, autotune: bool = False, mode: str = "grouped", tunes=None, buckets: bool = False,
               queue: bool = False):
    if not groups:
      raise ValueError("MultiStep needs at least one (Model, Data) group")
    if mode not in ("grouped", "per_group", "static"):
      raise ValueError(f"unknown mode {mode!r}")
    self.groups = [(m, d) for m, d in groups]
    if tunes is not None and len(tunes) != len(self.groups):
      raise ValueError("tunes needs one (team, rmax, rj) per group")
    self.mode = mode
    self.tunes = list(tunes) if tunes is not None else None
    self.buckets = bool(buckets)
    self.queue = bool(queue)
    self._plan = None
    self._structs = None
    self._build()
    if autotune:
      key = _signature(self.groups)
      if key not in _TUNED:
        _TUNED[key] = self._tune()
      self.mode, self.tunes, self.buckets, self.queue = _TUNED[key]
      self._build()

  def _build(self):
    lib = _native.lib()
    self.close()
    ms = [_native.model_struct(m) for m, _ in self.groups]
    ds = [_native.data_struct(m, d) for m, d in self.groups]
    n = len(self.groups)
    M = (type(ms[0]) * n)(*ms)
    D = (type(ds[0]) * n)(*ds)
    self._arrays = (M, D)
    self._structs = ds
    if self.mode == "static":
      return  # no plan: the library's own table
    tunes = None
    if self.tunes is not None:
      tunes = (_native.GroupTuneStruct * n)(*[_native.GroupTuneStruct(*map(int, t)) for t in self.tunes])
    plan = ctypes.c_void_p()
    flags = (_native.MJH_MULTI_BUCKETS if self.buckets else 0) | (_native.MJH_MULTI_QUEUE if self.queue else 0)
    rc = lib.mjh_multi_create_tuned(n, M, D, tunes, flags, ctypes.byref(plan))
    if rc != _native.MJH_OK:
      _native._raise(rc, "mjh_multi_create_tuned")
    self._plan = plan

  def _stale(self) -> bool:
    return any(d._abi is not s for (_, d), s in zip(self.groups, self._structs))

  def step(self):
    """Advances every world of every group by one step (asynchronous on the
    current stream; each group's `d.time` advances on the host)."""
    if self._stale():
      self._build()
    if self.mode == "grouped":
      rc = _native.lib().mjh_multi_step(self._plan, _native._stream())
      if rc != _native.MJH_OK:
        _native._raise(rc, "mjh_multi_step")
    elif self.mode == "static":
      M, D = self._arrays
      rc = _native.lib().mjh_multi_step_static(len(self.groups), M, D, _native._stream())
      if rc != _native.MJH_OK:
        _native._raise(rc, "mjh_multi_step_static")
    else:
      for m, d in self.groups:
        _native.call("mjh_step_fused", m, d)
    for m, d in self.groups:
      d.time = d.time + m.opt.timestep

  def _timed(self, steps: int) -> float:
    import torch

    self.step()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
      self.step()
    torch.cuda.synchronize()
    return time.perf_counter() - t0

  def _tune(self, steps: int = 5) -> tuple:
    """(mode, tunes, buckets, queue) of the fastest trial."""
    import torch

    saved = [({f: getattr(d, f).clone() for f in _STATE}, d.time) for _, d in self.groups]

    def restore(idx):
      for i in idx:
        (_, d), (fields, t) = self.groups[i], saved[i]
        for f, v in fields.items():
          getattr(d, f).copy_(v)
        d.time = t

    everything = range(len(self.groups))
    # each group alone over its candidates; another choice replaces the
    # library's only when it is at least 3% faster
    tunes = []
    for i, (m, d) in enumerate(self.groups):
      best, best_t = (0, 0, -1), float("inf")
      for cand in candidates(m):
        try:
          one = MultiStep([(m, d)], mode="grouped", tunes=[cand])
        except (ValueError, NotImplementedError):
          continue  # does not fit this device's LDS
        dt = one._timed(steps)
        one.close()
        restore([i])
        if dt < (best_t if cand == (0, 0, -1) or best_t == float("inf") else 0.97 * best_t):
          best, best_t = cand, dt
      tunes.append(best)
    # the whole batch: grouped with and without buckets, each with and
    # without work-queue launches, and per group
    trials = [("grouped", tunes, bk, qu) for bk in (False, True) for qu in (False, True)]
    trials.append(("per_group", None, False, False))
    best, best_t = trials[0], float("inf")
    for mode, tn, bk, qu in trials:
      self.mode, self.tunes, self.buckets, self.queue = mode, tn, bk, qu
      self._build()
      dt = self._timed(steps)
      restore(everything)
      if dt < best_t:
        best, best_t = (mode, tn, bk, qu), dt
    torch.cuda.synchronize()
    return best

  def close(self):
    """Frees the plan (synchronises the device first)."""
    if self._plan is not None and self._plan.value:
      import torch

      torch.cuda.synchronize()
      _native.lib().mjh_multi_destroy(self._plan)
    self._plan = None

  def __del__(self):
    try:
      self.close()
    except Exception:
      pass
