"""Nested event timing of the step's stages (warp_util.py:29-85 counterpart).

`EventTracer` and `event_scope` keep the reference's interface: inside
`with EventTracer() as tracer:`, every function decorated with
`event_scope` records a pair of timing events on the current HIP stream
around its body, nested by call; `tracer.trace()` returns
`{name: (elapsed_ms_per_call, sub_trace)}`.  As in the reference, each call
of a scope opens a fresh level below it, so a sub-trace holds the events of
its parent's last call.  Outside a tracer the decorator adds nothing.

`step` / `forward` normally issue the whole step from one library call;
while a tracer is active they run through the per-stage entry points
instead (the same kernels in the same order), so that the trace has the
reference's nesting: step > forward > fwd_position > kinematics, ...
"""

from __future__ import annotations

import functools

_STACK = None


def active() -> bool:
  return _STACK is not None


class EventTracer:
  def __init__(self, enabled: bool = True):
    global _STACK
    if enabled:
      _STACK = {}

  def __enter__(self):
    return self

  def trace(self) -> dict:
    """{name: (tuple of elapsed ms, one per call; nested trace)}; waits for
    the recorded events."""
    global _STACK
    if _STACK is None:
      return {}
    ret = {}
    for name, (events, sub) in _STACK.items():
      saved, _STACK = _STACK, sub
      sub_trace = self.trace()
      _STACK = saved
      for beg, end in events:
        end.synchronize()
      ret[name] = (tuple(beg.elapsed_time(end) for beg, end in events), sub_trace)
    return ret

  def __exit__(self, exc_type, exc, tb):
    global _STACK
    _STACK = None


def event_scope(fn, name: str = ""):
  name = name or fn.__name__

  @functools.wraps(fn)
  def wrapper(*args, **kwargs):
    global _STACK
    if _STACK is None:
      return fn(*args, **kwargs)
    import torch

    saved, _STACK = _STACK, {}
    stream = torch.cuda.current_stream()
    beg = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    beg.record(stream)
    try:
      res = fn(*args, **kwargs)
    finally:
      end.record(stream)
      sub, _STACK = _STACK, saved
      events, _ = _STACK.get(name, ((), None))
      _STACK[name] = (events + ((beg, end),), sub)
    return res

  return wrapper
