"""Functional step operator and graph unroll (contrib/jax_unroll.py counterpart).

CPU: the operator is registered with a shape-only fake implementation, and
inputs that require gradients are rejected.  GPU: the functional step and
the unroll give bit for bit the state of eager steps on a Data with the
same inputs, leave their inputs untouched, and an unroll called twice on
the same inputs returns the same outputs.
"""

import numpy as np
import pytest
import torch

import mjx_warp_amd as mjw
from mjx_warp_amd import models

import helpers


def test_operator_registered_with_fake_shapes():
  from torch._subclasses.fake_tensor import FakeTensorMode

  with FakeTensorMode():
    q = torch.empty(4, 28)
    v = torch.empty(4, 27)
    u = torch.empty(4, 21)
    qo, vo = torch.ops.mjx_warp_amd.step(0, q, v, u)
  assert qo.shape == (4, 28) and vo.shape == (4, 27)


def test_inputs_requiring_grad_are_rejected():
  mjm = models.humanoid()
  m = mjw.put_model(mjm, device="cpu")
  d = mjw.make_data(mjm, nworld=2, device="cpu")
  f = mjw.FunctionalStep(m, d)
  q = d.qpos.clone().requires_grad_(True)
  with pytest.raises(RuntimeError, match="forward-only"):
    f(q, d.qvel)
  f.close()


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
@pytest.mark.parametrize("fused", [True, False])
def test_functional_step_matches_eager(fused):
  mjm = models.humanoid(iterations=10, ls_iterations=20)
  m = mjw.put_model(mjm)
  d_ref = _state(mjm, 32, 7)
  d_fn = _state(mjm, 32, 7)
  f = mjw.FunctionalStep(m, d_fn, fused=fused)
  q0, v0 = d_ref.qpos.clone(), d_ref.qvel.clone()
  qin, vin = q0.clone(), v0.clone()
  for _ in range(3):
    (mjw.step_fused if fused else mjw.step)(m, d_ref)
    qin, vin = f(qin, vin)
  torch.cuda.synchronize()
  assert torch.equal(qin, d_ref.qpos) and torch.equal(vin, d_ref.qvel)
  # the caller's tensors are new objects, not views of d
  assert qin.data_ptr() != d_fn.qpos.data_ptr()
  f.close()


@pytest.mark.gpu
def test_unroll_matches_eager_and_repeats():
  mjm = models.humanoid(iterations=10, ls_iterations=20)
  m = mjw.put_model(mjm)
  d_ref = _state(mjm, 32, 9)
  d_un = _state(mjm, 32, 9)
  q0, v0 = d_ref.qpos.clone(), d_ref.qvel.clone()
  qc, vc = q0.clone(), v0.clone()
  un = mjw.Unroll(m, d_un, nstep=5)
  for _ in range(5):
    mjw.step_fused(m, d_ref)
  w0 = d_un.qacc_warmstart.clone()
  q1, v1 = un(q0, v0)
  # the solver's warm start is state carried in d, as in the reference; with
  # it restored, a second call reproduces the first
  d_un.qacc_warmstart.copy_(w0)
  q2, v2 = un(q0, v0)
  torch.cuda.synchronize()
  assert torch.equal(q1, d_ref.qpos) and torch.equal(v1, d_ref.qvel)
  assert torch.equal(q1, q2) and torch.equal(v1, v2)
  assert torch.equal(q0, qc) and torch.equal(v0, vc)
  assert abs(d_un.time - 10 * mjm.opt.timestep) < 1e-9
