"""EventTracer / event_scope (warp_util.py:29-85 counterpart).

CPU: the decorator is a pass-through without a tracer, a disabled tracer
traces nothing, and the nesting bookkeeping (names, per-call tuples,
sub-traces) is checked with stand-in timing events.  GPU: a traced step has
the reference's nesting (step > forward > fwd_position > kinematics, ...),
and gives the same state, bit for bit, as an untraced step through the same
per-stage kernels.
"""

import numpy as np
import pytest
import torch

import mjx_warp_amd as mjw
from mjx_warp_amd import models, tracing

import helpers


def test_scope_is_passthrough_without_tracer():
  calls = []

  @mjw.event_scope
  def f(x):
    calls.append(x)
    return 2 * x

  assert not tracing.active()
  assert f(3) == 6 and calls == [3]
  assert f.__name__ == "f"


def test_disabled_tracer_traces_nothing():
  @mjw.event_scope
  def f():
    return 1

  with mjw.EventTracer(enabled=False) as tracer:
    assert not tracing.active()
    assert f() == 1
    assert tracer.trace() == {}
  assert not tracing.active()


class _Clock:
  t = 0.0


class _Event:
  """Timing-event stand-in: record() reads a clock that each record advances
  by 1 ms."""

  def __init__(self, enable_timing=False):
    self.t = None

  def record(self, stream=None):
    _Clock.t += 1.0
    self.t = _Clock.t

  def synchronize(self):
    pass

  def elapsed_time(self, other):
    return other.t - self.t


def test_nesting_and_per_call_times(monkeypatch):
  monkeypatch.setattr(torch.cuda, "Event", _Event)
  monkeypatch.setattr(torch.cuda, "current_stream", lambda: None)
  _Clock.t = 0.0

  @mjw.event_scope
  def leaf():
    pass

  @mjw.event_scope
  def other():
    pass

  def named():
    leaf()

  named_scope = mjw.event_scope(named, "renamed")

  @mjw.event_scope
  def outer():
    leaf()
    leaf()
    other()
    named_scope()

  with mjw.EventTracer() as tracer:
    outer()
    outer()
    trace = tracer.trace()
  assert not tracing.active()

  assert set(trace) == {"outer"}
  times, sub = trace["outer"]
  # each outer call records: begin, 2 x leaf (2 each), other (2), renamed
  # (2, around leaf's 2), end = 12 records, 11 ms apart begin to end
  assert times == (11.0, 11.0)
  # as in the reference, a sub-trace holds the events of the parent's last
  # call (each call of the parent starts a fresh level)
  assert set(sub) == {"leaf", "other", "renamed"}
  assert sub["leaf"][0] == (1.0, 1.0) and sub["leaf"][1] == {}
  assert sub["other"][0] == (1.0,)
  assert sub["renamed"][0] == (3.0,)
  assert set(sub["renamed"][1]) == {"leaf"}
  assert sub["renamed"][1]["leaf"][0] == (1.0,)


def _state(mjm, nworld, seed):
  rng = np.random.default_rng(seed)
  d = mjw.make_data(mjm, nworld=nworld, nconmax=64 * nworld, njmax=277 * nworld)
  q = np.tile(mjm.key_qpos[0], (nworld, 1))
  q[:, 7:] += rng.normal(0, 0.05, (nworld, mjm.nq - 7))
  helpers.set_field(d, "qpos", q)
  helpers.set_field(d, "qvel", rng.normal(0, 0.5, (nworld, mjm.nv)))
  helpers.set_field(d, "ctrl", rng.uniform(-1, 1, (nworld, mjm.nu)))
  return d


@pytest.mark.gpu
def test_traced_step_nesting_and_state():
  mjm = models.humanoid(iterations=10, ls_iterations=20)
  m = mjw.put_model(mjm)
  d_plain = _state(mjm, 64, 5)
  d_trace = _state(mjm, 64, 5)
  nstep = 2
  for _ in range(nstep):
    mjw.step(m, d_plain)
  with mjw.EventTracer() as tracer:
    for _ in range(nstep):
      mjw.step(m, d_trace)
    trace = tracer.trace()
  torch.cuda.synchronize()

  for f in ("qpos", "qvel", "qacc", "qacc_warmstart", "qfrc_constraint"):
    a = helpers.data_numpy(d_trace, f)
    b = helpers.data_numpy(d_plain, f)
    assert np.array_equal(a, b), (f, np.abs(a - b).max())
  assert abs(d_trace.time - d_plain.time) < 1e-12

  assert set(trace) == {"step"}
  t_step, sub = trace["step"]
  assert len(t_step) == nstep
  assert set(sub) == {"forward", "implicit"}
  t_fwd, fwd = sub["forward"]
  assert set(fwd) == {"fwd_position", "fwd_velocity", "fwd_actuation", "fwd_acceleration", "solve"}
  pos = fwd["fwd_position"][1]
  assert set(pos) == {"kinematics", "com_pos", "crb", "factor_m", "collision", "make_constraint",
                      "transmission"}
  assert set(pos["collision"][1]) == {"nxn_broadphase", "primitive_narrowphase"}

  # the top level has one entry per step; a sub-trace holds the events of
  # its parent's last call, each stage called once per parent call
  def walk(node, parent_last):
    for name, (times, child) in node.items():
      assert all(t > 0.0 for t in times), name
      if parent_last is not None:
        assert len(times) == 1, name
      # children run back to back inside the parent's last call's events
      inner = sum(c[0][0] for c in child.values())
      assert inner <= times[-1] * 1.001 + 1e-3, (name, inner, times[-1])
      walk(child, times[-1])

  walk(trace, None)
  assert t_fwd[0] <= t_step[-1]
