"""Transition derivatives of the step and a differentiable rollout.

The reference has no gradients: its kernels overwrite arrays in place and
use integer atomics, and collision_convex.py:22 turns Warp's backward code
generation off.  This module follows MuJoCo's own derivative interface,
mjd_transitionFD (engine_derivative_fd.c), on the GPU:

* `TransitionFD(mjm, m, nworld, ...)` -- `A, B = tfd(d)` gives, for every
  world of d, A = dx'/dx [nx, nx] and B = dx'/du [nx, nu] of one step, with
  x = (dq [nv], qvel [nv], act [na]) (dq: tangent space of qpos, local-frame
  rotation vectors for free / ball joints) and u = ctrl.  All perturbed
  copies of all worlds are stepped in one batched step of a second Data of
  nworld * (1 + 2 (nx + nu)) worlds (centered differences; + (nx + nu) for
  forward differences), the fused step's own kernels, so a Jacobian costs
  one step of that batch.
* `Rollout(mjm, m, nworld, nstep, ...)` -- `qpos, qvel = r(qpos0, qvel0,
  ctrl)` runs nstep steps and is differentiable with respect to qpos0, qvel0
  and ctrl [nstep, nworld, nu] through `torch.autograd`.  The forward pass
  keeps each step's inputs (qpos, qvel, act, qacc_warmstart: a few hundred
  bytes per world-step, so every step is a checkpoint); the backward pass
  walks the steps in reverse, computes each step's (A, B) at its saved
  inputs and runs the adjoint recursion
      lam_T = g(x_T),  lam_t = A_t' lam_{t+1} + g(x_t),
      dL/du_t = B_t' lam_{t+1},
  on the GPU (mjh_adjoint_step), with g the loss gradient the caller's
  autograd graph hands back for each returned state.

Gradients are those of the discrete FP32 step at the active set of the
nominal state; a perturbation that switches a contact or limit on or off
inside +-eps makes the corresponding column a secant, as mjd_transitionFD's
does.  The default eps = 4e-3 balances the two FP32 error sources: in
smooth dynamics the rounding of the perturbed steps dominates (it shrinks
as 1/eps; the truncation error of central differences, eps^2, is still far
below it there), while a larger eps straddles more activations.
"""

from __future__ import annotations

import ctypes

import torch

from . import _native
from . import api
from . import io
from .types import Data, Model


def _ptr(t):
  return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def nx_of(mjm) -> int:
  return 2 * int(mjm.nv) + int(mjm.na)


class TransitionFD:
  """Finite-difference transition Jacobians of the fused step for every
  world of a Data of `nworld` worlds (mjd_transitionFD counterpart)."""

  def __init__(self, mjm, m: Model, nworld: int, nconmax: int = -1, njmax: int = -1, eps: float = 4e-3,
               centered: bool = True, device=None):
    if m.opt.integrator not in (0, 3):
      raise NotImplementedError(f"integrator {m.opt.integrator} not implemented.")
    self.m = m
    self.nworld = int(nworld)
    self.eps = float(eps)
    self.centered = bool(centered)
    self.nx, self.nu = nx_of(mjm), int(mjm.nu)
    self.per_world = 1 + (2 if centered else 1) * (self.nx + self.nu)
    nconmax = 512 if nconmax == -1 else nconmax
    njmax = 512 if njmax == -1 else njmax
    # the same per-world capacity as the nominal Data
    k = self.per_world
    self.fd = io.make_data(mjm, nworld=k * self.nworld, nconmax=k * nconmax, njmax=k * njmax, device=device)
    dev = self.fd.qpos.device
    self.A = torch.zeros((self.nworld, self.nx, self.nx), dtype=torch.float32, device=dev)
    self.B = torch.zeros((self.nworld, self.nx, self.nu), dtype=torch.float32, device=dev)
    self._fd_struct = _native.data_struct(m, self.fd)

  def __call__(self, d: Data):
    """(A, B) at d's current state; d is read, not stepped.  The returned
    tensors are overwritten by the next call."""
    if d.nworld != self.nworld:
      raise ValueError(f"TransitionFD was built for {self.nworld} worlds, got {d.nworld}")
    _native.call("mjh_transition_fd", self.m, d, ctypes.byref(self._fd_struct), ctypes.c_float(self.eps),
                 ctypes.c_int(int(self.centered)), _ptr(self.A), _ptr(self.B))
    return self.A, self.B


def adjoint_step(m: Model, d: Data, A, B, lam, g_qpos=None, g_qvel=None, g_act=None, lam_next=None, grad_ctrl=None):
  """lam = A' lam_next + g(x) at d's state; grad_ctrl = B' lam_next."""
  _native.call("mjh_adjoint_step", m, d, _ptr(A), _ptr(B), _ptr(g_qpos), _ptr(g_qvel), _ptr(g_act), _ptr(lam_next),
               _ptr(lam), _ptr(grad_ctrl))


def vjp_step(m: Model, d: Data, qpos, qvel, lam_next, lam, grad_ctrl, gadd=None):
  """The analytic vector-Jacobian product of the step just taken from
  (qpos, qvel) [nworld, nq] / [nworld, nv]: d holds that step's forward
  quantities (as `step_fused` leaves them).  lam = (dx'/dx)' lam_next (+
  gadd), grad_ctrl = (dx'/du)' lam_next."""
  _native.call("mjh_adjoint_vjp", m, d, _ptr(qpos), _ptr(qvel), _ptr(lam_next), _ptr(gadd), _ptr(lam),
               _ptr(grad_ctrl))


def tape_bytes(m: Model, d: Data) -> int:
  """Bytes of one step's slot of the adjoint tape for (m, d)."""
  n = ctypes.c_ulonglong(0)
  _native.call("mjh_adjoint_tape_bytes", m, d, ctypes.byref(n))
  return int(n.value)


def tape_save(m: Model, d: Data, slot):
  """Saves the fields of the step d just took that the adjoint reads into
  `slot` (a uint8 device tensor of tape_bytes(m, d) bytes)."""
  _native.call("mjh_adjoint_tape_save", m, d, _ptr(slot))


def vjp_step_taped(m: Model, d: Data, slot, qpos, qvel, lam_next, lam, grad_ctrl, gadd=None):
  """vjp_step with the step's fields read from a tape slot (tape_save)
  instead of d."""
  _native.call("mjh_adjoint_vjp_taped", m, d, _ptr(slot), _ptr(qpos), _ptr(qvel), _ptr(lam_next), _ptr(gadd),
               _ptr(lam), _ptr(grad_ctrl))


def check_analytic(mjm):
  """Raises NotImplementedError if the analytic adjoint does not cover the
  model (mjh_adjoint_vjp: Euler / implicitfast, free / hinge / slide joints,
  no springs on free joints, no activation states, joint transmissions with
  fixed gains, no actuator force clamps at the joints)."""
  import numpy as np

  if int(mjm.opt.integrator) not in (0, 3):
    raise NotImplementedError("analytic adjoint: integrator must be Euler or implicitfast")
  if int(mjm.na) > 0:
    raise NotImplementedError("analytic adjoint: activation states are not differentiated")
  jt = np.asarray(mjm.jnt_type)
  if np.any(jt == 1):
    raise NotImplementedError("analytic adjoint: ball joints are not differentiated")
  if np.any((jt == 0) & (np.asarray(mjm.jnt_stiffness) != 0)):
    raise NotImplementedError("analytic adjoint: free-joint springs are not differentiated")
  if int(mjm.nu):
    if np.any(np.asarray(mjm.actuator_gaintype) != 0) or np.any(np.asarray(mjm.actuator_biastype) != 0):
      raise NotImplementedError("analytic adjoint: actuators must have fixed gains and no bias")
    if np.any(np.asarray(mjm.actuator_dyntype) != 0):
      raise NotImplementedError("analytic adjoint: actuator dynamics are not differentiated")
    trn = np.asarray(mjm.actuator_trntype)
    tid = np.asarray(mjm.actuator_trnid)[:, 0]
    if np.any(trn != 0) or np.any(jt[tid] == 1):
      raise NotImplementedError("analytic adjoint: actuators must act on hinge, slide or free joints")
  if hasattr(mjm, "jnt_actfrclimited") and np.any(np.asarray(mjm.jnt_actfrclimited)):
    raise NotImplementedError("analytic adjoint: joint actuator force limits are not differentiated")


def adjoint_state_grad(m: Model, d: Data, lam, g_qpos, g_qvel, g_act=None):
  """Splits lam [nworld, nx] into gradients with respect to d's raw qpos,
  qvel and act."""
  _native.call("mjh_adjoint_state_grad", m, d, _ptr(lam), _ptr(g_qpos), _ptr(g_qvel), _ptr(g_act))


class Rollout:
  """Differentiable `nstep`-step rollout of `nworld` worlds:
  `qpos, qvel = r(qpos0, qvel0, ctrl)` with qpos [nstep, nworld, nq] and
  qvel [nstep, nworld, nv] the states after each step.  Inputs other than
  (qpos, qvel, ctrl) -- qfrc_applied, mocap -- are read from `r.d`, set
  them there; act and qacc_warmstart start from `act0` / `warmstart0`
  (zeros by default).

  adjoint="analytic" (the default) runs the reverse pass with the
  hand-written step adjoint (mjh_adjoint_vjp), one vector-Jacobian product
  per step.  The fields it reads are taped during the forward pass
  (tape_save: tape_bytes(m, d) per step, the pools at full capacity) when
  nstep slots fit in `tape_budget` bytes; otherwise, or with tape=False,
  each backward step recomputes its forward quantities from its saved input
  state with one fused step.  checkpoint_every=k (recomputing only) keeps
  the input state of every k-th step (memory / k) and recomputes each
  segment of k steps from its checkpoint (one more step per step).
  adjoint="fd" uses finite-difference transition Jacobians instead
  (TransitionFD; any model the step runs)."""

  def __init__(self, mjm, m: Model, nworld: int, nstep: int, nconmax: int = -1, njmax: int = -1, eps: float = 4e-3,
               centered: bool = True, device=None, adjoint: str = "analytic", checkpoint_every: int = 1,
               tape=None, tape_budget: int = 8 << 30):
    if adjoint not in ("analytic", "fd"):
      raise ValueError(f"adjoint must be 'analytic' or 'fd', got {adjoint!r}")
    if adjoint == "analytic":
      check_analytic(mjm)
    self.m = m
    self.adjoint = adjoint
    self.nworld, self.nstep = int(nworld), int(nstep)
    self.every = max(1, int(checkpoint_every))
    self.d = io.make_data(mjm, nworld=nworld, nconmax=nconmax, njmax=njmax, device=device)
    self.tfd = TransitionFD(mjm, m, nworld, nconmax, njmax, eps, centered, device) if adjoint == "fd" else None
    self.nq, self.nv, self.na, self.nu = int(mjm.nq), int(mjm.nv), int(mjm.na), int(mjm.nu)
    dev = self.d.qpos.device
    z = lambda *s: torch.zeros(s, dtype=torch.float32, device=dev)
    T, W = self.nstep, self.nworld
    # input states of the checkpointed steps, and of one segment's steps
    C = (T + self.every - 1) // self.every
    self.c_qpos, self.c_qvel = z(C, W, self.nq), z(C, W, self.nv)
    self.c_act, self.c_warm = z(C, W, self.na), z(C, W, self.nv)
    S = self.every
    self.s_qpos, self.s_qvel = z(S, W, self.nq), z(S, W, self.nv)
    self.s_act, self.s_warm = z(S, W, self.na), z(S, W, self.nv)
    self.lam = z(W, nx_of(mjm))
    self.lam_next = z(W, nx_of(mjm))
    self.lam_g = z(W, nx_of(mjm))
    self.tape = None
    if adjoint == "analytic" and tape is not False and self.every == 1:
      nb = tape_bytes(m, self.d)
      if tape or T * nb <= tape_budget:
        self.tape = torch.empty((T, nb), dtype=torch.uint8, device=dev)

  def _save(self, buf, i):
    d = self.d
    buf[0][i].copy_(d.qpos)
    buf[1][i].copy_(d.qvel)
    buf[2][i].copy_(d.act)
    buf[3][i].copy_(d.qacc_warmstart)

  def _restore(self, buf, i):
    d = self.d
    d.qpos.copy_(buf[0][i])
    d.qvel.copy_(buf[1][i])
    d.act.copy_(buf[2][i])
    d.qacc_warmstart.copy_(buf[3][i])

  def forward(self, qpos0, qvel0, ctrl, act0=None, warmstart0=None, record: bool = True):
    """Runs the rollout without recording anything for autograd; with
    record=False the tape is not written either (the states and checkpoints
    are), so a later backward() is invalid."""
    d, T = self.d, self.nstep
    if tuple(ctrl.shape) != (T, self.nworld, self.nu):
      raise ValueError(f"ctrl must be [{T}, {self.nworld}, {self.nu}], got {tuple(ctrl.shape)}")
    d.qpos.copy_(qpos0)
    d.qvel.copy_(qvel0)
    d.act.copy_(act0) if act0 is not None else d.act.zero_()
    d.qacc_warmstart.copy_(warmstart0) if warmstart0 is not None else d.qacc_warmstart.zero_()
    qpos = torch.empty((T, self.nworld, self.nq), dtype=torch.float32, device=d.qpos.device)
    qvel = torch.empty((T, self.nworld, self.nv), dtype=torch.float32, device=d.qpos.device)
    ck = (self.c_qpos, self.c_qvel, self.c_act, self.c_warm)
    for t in range(T):
      if t % self.every == 0:
        self._save(ck, t // self.every)
      d.ctrl.copy_(ctrl[t])
      api.step_fused(self.m, d)
      if self.tape is not None and record:
        tape_save(self.m, d, self.tape[t])
      qpos[t].copy_(d.qpos)
      qvel[t].copy_(d.qvel)
    return qpos, qvel

  def backward(self, ctrl, qpos, g_qpos, g_qvel):
    """Gradients (qpos0, qvel0, ctrl) from the gradients with respect to the
    returned trajectories (the forward pass must have run with ctrl)."""
    m, d, T, K = self.m, self.d, self.nstep, self.every
    if self.adjoint == "analytic" and bool(torch.any(d.xfrc_applied != 0)):
      raise NotImplementedError("analytic adjoint: xfrc_applied is not differentiated; use adjoint='fd'")
    grad_ctrl = torch.zeros_like(ctrl)
    ck = (self.c_qpos, self.c_qvel, self.c_act, self.c_warm)
    seg = (self.s_qpos, self.s_qvel, self.s_act, self.s_warm)
    # lam_T = g(x_T), at x_T's qpos
    d.qpos.copy_(qpos[T - 1])
    adjoint_step(m, d, None, None, self.lam_next, g_qpos=g_qpos[T - 1], g_qvel=g_qvel[T - 1])
    if self.tape is not None:
      # every step's input state and fields are kept: no recomputation
      for t in range(T - 1, -1, -1):
        gadd = None
        if t > 0:
          d.qpos.copy_(self.c_qpos[t])
          adjoint_step(m, d, None, None, self.lam_g, g_qpos=g_qpos[t - 1], g_qvel=g_qvel[t - 1])
          gadd = self.lam_g
        vjp_step_taped(m, d, self.tape[t], self.c_qpos[t], self.c_qvel[t], self.lam_next, self.lam, grad_ctrl[t],
                       gadd)
        self.lam, self.lam_next = self.lam_next, self.lam
    for s0 in (range(((T - 1) // K) * K, -1, -K) if self.tape is None else ()):
      s1 = min(s0 + K, T)
      # the segment's input states, recomputed from its checkpoint
      self._restore(ck, s0 // K)
      for t in range(s0, s1):
        self._save(seg, t - s0)
        if t + 1 < s1:
          d.ctrl.copy_(ctrl[t])
          api.step_fused(m, d)
      for t in range(s1 - 1, s0 - 1, -1):
        i = t - s0
        self._restore(seg, i)
        d.ctrl.copy_(ctrl[t])
        g_q = g_qpos[t - 1] if t > 0 else None
        g_v = g_qvel[t - 1] if t > 0 else None
        if self.adjoint == "fd":
          A, B = self.tfd(d)
          adjoint_step(m, d, A, B, self.lam, g_qpos=g_q, g_qvel=g_v, lam_next=self.lam_next,
                       grad_ctrl=grad_ctrl[t])
        else:
          gadd = None
          if t > 0:
            adjoint_step(m, d, None, None, self.lam_g, g_qpos=g_q, g_qvel=g_v)
            gadd = self.lam_g
          api.step_fused(m, d)
          vjp_step(m, d, seg[0][i], seg[1][i], self.lam_next, self.lam, grad_ctrl[t], gadd)
        self.lam, self.lam_next = self.lam_next, self.lam
    # lam_next now holds lam_0; x_0's qpos turns its rotation part into
    # raw quaternion gradients
    d.qpos.copy_(self.c_qpos[0])
    gq0 = torch.empty((self.nworld, self.nq), dtype=torch.float32, device=d.qpos.device)
    gv0 = torch.empty((self.nworld, self.nv), dtype=torch.float32, device=d.qpos.device)
    ga0 = torch.empty((self.nworld, self.na), dtype=torch.float32, device=d.qpos.device)
    adjoint_state_grad(m, d, self.lam_next, gq0, gv0, ga0 if self.na else None)
    return gq0, gv0, grad_ctrl

  def __call__(self, qpos0, qvel0, ctrl):
    return _RolloutFn.apply(self, qpos0, qvel0, ctrl)


class _RolloutFn(torch.autograd.Function):

  @staticmethod
  def forward(ctx, r: Rollout, qpos0, qvel0, ctrl):
    ctrl = ctrl.detach().contiguous()
    qpos, qvel = r.forward(qpos0.detach(), qvel0.detach(), ctrl)
    ctx.r = r
    ctx.save_for_backward(ctrl, qpos)
    return qpos, qvel

  @staticmethod
  def backward(ctx, g_qpos, g_qvel):
    ctrl, qpos = ctx.saved_tensors
    gq0, gv0, gu = ctx.r.backward(ctrl, qpos, g_qpos.contiguous(), g_qvel.contiguous())
    return None, gq0, gv0, gu
