"""Functional step and unrolled rollouts for PyTorch code
(contrib/jax_unroll.py:39-90 counterpart).

The reference hands `step` to JAX as a functional callable over
(qpos, qvel) -- inputs copied into the batch Data, one step, outputs copied
out (jax_unroll.py:39-58) -- and unrolls it with `lax.scan` under `jax.jit`,
which XLA records as one CUDA graph (jax_unroll.py:63-79).  Here:

* `FunctionalStep(m, d)` is that callable: `qpos, qvel = f(qpos, qvel[,
  ctrl])` returns new tensors and leaves the caller's inputs untouched.  It
  is also registered as the operator `torch.ops.mjx_warp_amd.step`
  (`torch.library.custom_op`, with a shape-only fake implementation), so
  that traced / compiled PyTorch programs can hold it as one opaque node.
* `Unroll(m, d, nstep)` is the scan: `nstep` steps captured once as a HIP
  graph between static input and output buffers; each call copies the
  inputs in, replays the graph and returns copies of the outputs.

These are forward-only (inputs that require gradients are rejected); the
differentiable rollout, with an adjoint the reference does not have, is
`mjx_warp_amd.derivative.Rollout`.
"""

from __future__ import annotations

import itertools

import torch

from . import api
from .types import Data, Model

_REGISTRY: dict = {}
_KEYS = itertools.count()


def _check(*tensors):
  for t in tensors:
    if t is not None and t.requires_grad:
      raise RuntimeError("FunctionalStep / Unroll are forward-only: use mjx_warp_amd.derivative.Rollout "
                         "for gradients")


class FunctionalStep:
  """qpos, qvel (, ctrl) -> qpos, qvel after one step of every world of d."""

  def __init__(self, m: Model, d: Data, fused: bool = True):
    self.m, self.d, self.fused = m, d, fused
    self.key = next(_KEYS)
    _REGISTRY[self.key] = self

  def run(self, qpos, qvel, ctrl=None):
    d = self.d
    d.qpos.copy_(qpos)
    d.qvel.copy_(qvel)
    if ctrl is not None:
      d.ctrl.copy_(ctrl)
    (api.step_fused if self.fused else api.step)(self.m, d)
    return d.qpos.clone(), d.qvel.clone()

  def __call__(self, qpos, qvel, ctrl=None):
    _check(qpos, qvel, ctrl)
    if ctrl is None:
      ctrl = self.d.ctrl
    return torch.ops.mjx_warp_amd.step(self.key, qpos, qvel, ctrl)

  def close(self):
    _REGISTRY.pop(self.key, None)


@torch.library.custom_op("mjx_warp_amd::step", mutates_args=())
def _step_op(key: int, qpos: torch.Tensor, qvel: torch.Tensor, ctrl: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
  return _REGISTRY[key].run(qpos, qvel, ctrl)


@_step_op.register_fake
def _(key, qpos, qvel, ctrl):
  return torch.empty_like(qpos), torch.empty_like(qvel)


class Unroll:
  """nstep steps of every world from (qpos, qvel), replayed from one HIP
  graph; ctrl (if given) is held fixed over the rollout, as the reference's
  scan carries only qpos and qvel."""

  def __init__(self, m: Model, d: Data, nstep: int, fused: bool = True):
    self.m, self.d, self.nstep = m, d, int(nstep)
    self.qpos_in = torch.empty_like(d.qpos)
    self.qvel_in = torch.empty_like(d.qvel)
    self.qpos_out = torch.empty_like(d.qpos)
    self.qvel_out = torch.empty_like(d.qvel)
    fn = api.step_fused if fused else api.step
    self.qpos_in.copy_(d.qpos)
    self.qvel_in.copy_(d.qvel)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    self.graph = torch.cuda.CUDAGraph()
    time0 = d.time
    with torch.cuda.stream(stream):
      with torch.cuda.graph(self.graph, stream=stream):
        d.qpos.copy_(self.qpos_in)
        d.qvel.copy_(self.qvel_in)
        for _ in range(self.nstep):
          fn(m, d)
        self.qpos_out.copy_(d.qpos)
        self.qvel_out.copy_(d.qvel)
    torch.cuda.current_stream().wait_stream(stream)
    d.time = time0  # capture launched nothing

  def __call__(self, qpos, qvel, ctrl=None):
    _check(qpos, qvel, ctrl)
    self.qpos_in.copy_(qpos)
    self.qvel_in.copy_(qvel)
    if ctrl is not None:
      self.d.ctrl.copy_(ctrl)
    self.graph.replay()
    self.d.time = self.d.time + self.nstep * self.m.opt.timestep
    return self.qpos_out.clone(), self.qvel_out.clone()
