"""mjh_multi_step_static: the heterogeneous step without a plan and without
allocation (the group table in the library's fixed device array), called
through the C ABI and checked against the planned entry points
(mjh_multi_create + mjh_multi_step) on identical groups: the same kernels
on the same inputs, so every world's state must agree bit for bit.

Two calls in a row: first five groups (three 32-lane box-world models, a
humanoid, a 64-lane 40-link chain), then a smaller input of two groups
with fewer worlds, whose table overwrites the start of the first call's.
"""

import ctypes

import numpy as np
import pytest
import torch

import mjx_warp_amd as mjw
from mjx_warp_amd import _native

import helpers
from test_multi import _groups

pytestmark = pytest.mark.gpu

_FIELDS = ("qpos", "qvel", "qacc", "qacc_warmstart")


def _abi(groups):
  ms = [_native.model_struct(m) for m, _ in groups]
  ds = [_native.data_struct(m, d) for m, d in groups]
  n = len(groups)
  return (type(ms[0]) * n)(*ms), (type(ds[0]) * n)(*ds)


def _static_step(groups):
  M, D = _abi(groups)
  rc = _native.lib().mjh_multi_step_static(len(groups), M, D, _native._stream())
  assert rc == _native.MJH_OK, _native.lib().mjh_last_error().decode()


def _planned_step(groups):
  lib = _native.lib()
  M, D = _abi(groups)
  plan = ctypes.c_void_p()
  assert lib.mjh_multi_create(len(groups), M, D, ctypes.byref(plan)) == _native.MJH_OK
  assert lib.mjh_multi_step(plan, _native._stream()) == _native.MJH_OK
  torch.cuda.synchronize()
  lib.mjh_multi_destroy(plan)


def _same(ga, gb, what):
  for i, ((_, da), (_, db)) in enumerate(zip(ga, gb)):
    for f in _FIELDS:
      a, b = helpers.data_numpy(da, f), helpers.data_numpy(db, f)
      assert np.isfinite(a).all(), (what, i, f)
      assert np.array_equal(a, b), (what, i, f, np.abs(a - b).max())


def test_static_multi_step_equals_planned_step():
  _, make = _groups(seed=4)
  ga, gb = make(), make()
  for _ in range(2):
    _static_step(ga)
    _planned_step(gb)
  torch.cuda.synchronize()
  _same(ga, gb, "five groups")

  # a smaller second input: two box-world groups of fewer worlds
  _, make_small = _groups(seed=5)
  sa, sb = make_small()[:2], make_small()[:2]
  for (m, d) in sa + sb:
    assert d.nworld <= 7
  _static_step(sa)
  _planned_step(sb)
  torch.cuda.synchronize()
  _same(sa, sb, "two groups")


def test_static_multi_step_rejects_too_many_groups():
  _, make = _groups(seed=6)
  g = make()[:1]
  M, D = _abi(g * 65)
  rc = _native.lib().mjh_multi_step_static(65, M, D, _native._stream())
  assert rc == _native.MJH_ERR_CAPACITY
  assert "at most 64" in _native.lib().mjh_last_error().decode()
