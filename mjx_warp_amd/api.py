"""The per-stage functions of the drop-in API.

Each function has the name and signature of its mujoco_warp counterpart
(mujoco_warp/__init__.py:18-47), mutates `d` in place on torch's current
stream and returns without synchronising.  All compute runs in
libmjx_warp_amd.so; there is no Python or CPU fallback.
"""

from __future__ import annotations

import ctypes

from . import _native
from .tracing import active as _tracing
from .tracing import event_scope
from .types import Data, Model


def _ptr(t):
  return ctypes.c_void_p(t.data_ptr())


@event_scope
def kinematics(m: Model, d: Data):
  """Forward kinematics (smooth.py:32-145)."""
  _native.call("mjh_kinematics", m, d)


@event_scope
def com_pos(m: Model, d: Data):
  """Subtree COM, com-based inertias and motion dofs (smooth.py:148-245)."""
  _native.call("mjh_com_pos", m, d)


@event_scope
def crb(m: Model, d: Data):
  """Composite rigid-body inertia and qM (smooth.py:248-314)."""
  _native.call("mjh_crb", m, d)


@event_scope
def factor_m(m: Model, d: Data):
  """Cholesky factor of qM into qLD (smooth.py:392-395)."""
  _native.call("mjh_factor_m", m, d)


@event_scope
def solve_m(m: Model, d: Data, x, y):
  """x = inv(M) y using qLD (smooth.py:693-696)."""
  _native.call("mjh_solve_m", m, d, _ptr(x), _ptr(y))


@event_scope
def rne(m: Model, d: Data):
  """Bias forces by recursive Newton-Euler (smooth.py:398-466)."""
  _native.call("mjh_rne", m, d)


@event_scope
def transmission(m: Model, d: Data):
  """Actuator lengths and moments (smooth.py:469-536)."""
  _native.call("mjh_transmission", m, d)
  return d


@event_scope
def com_vel(m: Model, d: Data):
  """cvel and cdof_dot (smooth.py:539-606)."""
  _native.call("mjh_com_vel", m, d)


@event_scope
def passive(m: Model, d: Data):
  """Joint springs and dof dampers (passive.py:27-108)."""
  _native.call("mjh_passive", m, d)


@event_scope
def mul_m(m: Model, d: Data, res, vec, skip):
  """res = M vec for worlds with skip == False (support.py:35-137)."""
  _native.call("mjh_mul_m", m, d, _ptr(res), _ptr(vec), _ptr(skip) if skip is not None else None)


@event_scope
def xfrc_accumulate(m: Model, d: Data, qfrc):
  """qfrc += J' xfrc_applied (support.py:140-172)."""
  _native.call("mjh_xfrc_accumulate", m, d, _ptr(qfrc))


@event_scope
def collision(m: Model, d: Data):
  """Broadphase, narrowphase and contact parameters (collision_driver.py:525-547)."""
  if _tracing():
    # the broadphase the library picks: all pairs up to 100 geoms, else
    # sweep-and-prune; each broadphase entry point clears the contact count
    (nxn_broadphase if m.ngeom <= 100 else sap_broadphase)(m, d)
    primitive_narrowphase(m, d)
    return
  _native.call("mjh_collision", m, d)


@event_scope
def nxn_broadphase(m: Model, d: Data):
  """All-pairs broadphase (collision_driver.py:458-512)."""
  _native.call("mjh_nxn_broadphase", m, d)


@event_scope
def sap_broadphase(m: Model, d: Data):
  """Sweep-and-prune broadphase (collision_driver.py:350-455)."""
  _native.call("mjh_sap_broadphase", m, d)
  return d


@event_scope
def primitive_narrowphase(m: Model, d: Data):
  """Primitive-pair narrowphase (collision_primitive.py:271-274)."""
  _native.call("mjh_primitive_narrowphase", m, d)


@event_scope
def make_constraint(m: Model, d: Data):
  """Joint-limit and pyramidal contact rows (constraint.py:213-240)."""
  _native.call("mjh_make_constraint", m, d)


@event_scope
def solve(m: Model, d: Data):
  """CG / Newton constraint solver (solver.py:872-1000)."""
  _native.call("mjh_solve", m, d)


@event_scope
def fwd_position(m: Model, d: Data):
  """Position-dependent computations (forward.py:408-420)."""
  if _tracing():
    kinematics(m, d)
    com_pos(m, d)
    crb(m, d)
    factor_m(m, d)
    collision(m, d)
    make_constraint(m, d)
    transmission(m, d)
    return
  _native.call("mjh_fwd_position", m, d)


@event_scope
def fwd_velocity(m: Model, d: Data):
  """Velocity-dependent computations (forward.py:423-500)."""
  _native.call("mjh_fwd_velocity", m, d)


@event_scope
def fwd_actuation(m: Model, d: Data):
  """Actuator forces (forward.py:503-641)."""
  _native.call("mjh_fwd_actuation", m, d)


@event_scope
def fwd_acceleration(m: Model, d: Data):
  """qfrc_smooth and qacc_smooth (forward.py:644-661)."""
  _native.call("mjh_fwd_acceleration", m, d)


@event_scope
def forward(m: Model, d: Data):
  """Forward dynamics (forward.py:664-679)."""
  if _tracing():
    fwd_position(m, d)
    fwd_velocity(m, d)
    fwd_actuation(m, d)
    fwd_acceleration(m, d)
    if d.njmax == 0:
      d.qacc.copy_(d.qacc_smooth)
    else:
      solve(m, d)
    return
  _native.call("mjh_forward", m, d)


@event_scope
def euler(m: Model, d: Data):
  """Semi-implicit Euler with implicit damping (forward.py:165-242)."""
  _native.call("mjh_euler", m, d)
  d.time = d.time + m.opt.timestep


@event_scope
def implicit(m: Model, d: Data):
  """implicitfast integrator (forward.py:245-405)."""
  if m.opt.is_sparse:
    raise AssertionError("implicit integration requires a dense mass matrix")
  _native.call("mjh_implicit", m, d)
  d.time = d.time + m.opt.timestep


@event_scope
def step(m: Model, d: Data):
  """Advance simulation by one step (forward.py:682-695), one launch per stage."""
  if m.opt.integrator not in (0, 3):
    raise NotImplementedError(f"integrator {m.opt.integrator} not implemented.")
  if _tracing():
    forward(m, d)
    (euler if m.opt.integrator == 0 else implicit)(m, d)
    return
  _native.call("mjh_step", m, d)
  d.time = d.time + m.opt.timestep


@event_scope
def step_fused(m: Model, d: Data):
  """The same step computed by a single fused launch per step."""
  if m.opt.integrator not in (0, 3):
    raise NotImplementedError(f"integrator {m.opt.integrator} not implemented.")
  _native.call("mjh_step_fused", m, d)
  d.time = d.time + m.opt.timestep


@event_scope
def rollout(m: Model, d: Data, nstep: int, fused: bool = True):
  """nstep consecutive steps issued by one library call."""
  _native.call("mjh_rollout", m, d, ctypes.c_int(int(nstep)), ctypes.c_int(int(bool(fused))))
  d.time = d.time + nstep * m.opt.timestep


class StepGraph:
  """`nstep` consecutive steps captured once as a HIP graph and replayed.

  The counterpart of the reference benchmark's captured step
  (test_util.py:99-118, wp.capture_begin / capture_launch): the step's
  launches are recorded on a capture stream (the library allocates no
  memory and records no memset nodes, so a step is capturable; its
  big-world solve forks onto a second stream and joins back) and `replay()`
  launches them on the current stream.  Replays read and write the same
  Model / Data tensors that were captured; `d.time` is advanced on the
  host by each replay, as `step` does.
  """

  def __init__(self, m: Model, d: Data, nstep: int = 1, fused: bool = True):
    import torch

    if m.opt.integrator not in (0, 3):
      raise NotImplementedError(f"integrator {m.opt.integrator} not implemented.")
    self.m, self.d, self.nstep = m, d, int(nstep)
    name = "mjh_step_fused" if fused else "mjh_step"
    _native.check(m, d)
    object.__setattr__(d, "_checked", m)
    self.graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(self.graph, stream=side):
      for _ in range(self.nstep):
        _native.call(name, m, d)
    torch.cuda.current_stream().wait_stream(side)

  def replay(self):
    self.graph.replay()
    self.d.time = self.d.time + self.nstep * self.m.opt.timestep
