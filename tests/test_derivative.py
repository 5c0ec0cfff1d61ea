"""Transition derivatives (mjd_transitionFD counterpart) and the adjoint of a
rollout.

The reference has no gradients, so the oracle is finite differences of the
FP64 oracle step (oracle/derivative.py):

CPU, pinning the restatement:
* a free sphere in zero gravity has a closed-form A (translation: identity
  plus dt coupling; rotation dofs: rotated by exp(-dt w); angular velocity
  into rotation: dt times the right Jacobian of exp);
* the adjoint recursion over the oracle's (A, B) reproduces the gradient of
  a rollout loss taken by central differences of whole rollouts, for the
  humanoid (a free joint with a quaternion in the loss, contacts, 21
  actuators), with respect to ctrl and qvel0.
GPU: the device Jacobians against the oracle's, and the device rollout's
autograd gradients against the loss's central differences on the FP64
oracle.
"""

import numpy as np
import pytest
import torch

import mjx_warp_amd as mjw
from mjx_warp_amd import derivative, mjcf, models
from oracle import derivative as od

import helpers

SPHERE = """<mujoco><option timestep="0.01" gravity="0 0 0"/><worldbody>
  <body pos="0 0 1"><freejoint/><geom type="sphere" size="0.1" mass="1"/></body>
</worldbody></mujoco>"""


def _skew(v):
  x, y, z = v
  return np.array([[0, -z, y], [z, 0, -x], [-y, x, 0]])


def _rot(phi):
  """Rotation matrix of the rotation vector phi."""
  th = np.linalg.norm(phi)
  if th == 0:
    return np.eye(3)
  K = _skew(phi / th)
  return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K


def _right_jac(phi):
  th = np.linalg.norm(phi)
  K = _skew(phi)
  return np.eye(3) - (1 - np.cos(th)) / th**2 * K + (th - np.sin(th)) / th**3 * K @ K


def _sphere_A(dt, w):
  A = np.eye(12)
  A[0:3, 6:9] = dt * np.eye(3)
  A[3:6, 3:6] = _rot(-dt * w)
  A[3:6, 9:12] = dt * _right_jac(dt * w)
  return A


def _sphere_case():
  mjm = mjcf.from_xml_string(SPHERE)
  rng = np.random.default_rng(0)
  q = np.tile(mjm.qpos0, (2, 1)).astype(np.float64)
  q[:, 3:7] = rng.normal(size=(2, 4))
  q[:, 3:7] /= np.linalg.norm(q[:, 3:7], axis=1, keepdims=True)
  v = rng.normal(0, 1.0, (2, 6))
  return mjm, q, v


def test_oracle_fd_free_sphere_closed_form():
  mjm, q, v = _sphere_case()
  m = mjw.put_model(mjm, device="cpu")
  step = helpers.oracle_stepper(mjm, m, nconmax_w=4, njmax_w=8)
  z = np.zeros((2, 0))
  A, B = od.transition_fd(step, mjm, q, v, z, z, np.zeros((2, 6)), 1e-6)
  assert B.shape == (2, 12, 0)
  for w in range(2):
    np.testing.assert_allclose(A[w], _sphere_A(mjm.opt.timestep, v[w, 3:]), atol=1e-8)


def _humanoid():
  return models.humanoid(iterations=10, ls_iterations=20)


def _loss(traj, ctrl):
  """Per-world sum over steps of (torso z - 1.3)^2 + 0.1 torso quaternion x
  + 1e-3 |ctrl|^2 (the quaternion term exercises the rotation tangent)."""
  L = 0.0
  for t, (q, _, _) in enumerate(traj):
    L = L + (q[..., 2] - 1.3) ** 2 + 0.1 * q[..., 4] + 1e-3 * np.sum(ctrl[t] ** 2, -1)
  return L


def _loss_grads(traj):
  g_qpos = []
  g_qvel = []
  for q, v, _ in traj:
    g = np.zeros_like(q)
    g[:, 2] = 2 * (q[:, 2] - 1.3)
    g[:, 4] = 0.1
    g_qpos.append(g)
    g_qvel.append(np.zeros_like(v))
  return g_qpos, g_qvel


def _rollout_case(mjm, n, T, seed=3):
  q0, v0 = models.humanoid_rollout_state(mjm, n, seed)
  rng = np.random.default_rng(seed + 100)
  v0 = v0 + rng.normal(0, 0.1, v0.shape)
  ctrl = rng.uniform(-1, 1, (T, n, mjm.nu))
  return q0, v0, ctrl


def test_oracle_adjoint_matches_rollout_loss_fd():
  mjm = _humanoid()
  m = mjw.put_model(mjm, device="cpu")
  step = helpers.oracle_stepper(mjm, m)
  n, T = 2, 4
  q0, v0, ctrl = _rollout_case(mjm, n, T)
  z0, w0 = np.zeros((n, 0)), np.zeros((n, mjm.nv))
  traj, inputs = od.rollout(step, q0, v0, z0, ctrl, w0)
  jac = [od.transition_fd(step, mjm, *inp, 1e-6) for inp in inputs]
  gq, gv = _loss_grads(traj)
  lam0, gu = od.adjoint(mjm, jac, inputs, traj, gq, gv)
  gu = gu + 2e-3 * ctrl
  fd_u, fd_v0 = od.loss_grad_fd(step, _loss, q0, v0, z0, ctrl, w0, 1e-6)
  scale_u = np.max(np.abs(fd_u))
  assert scale_u > 1e-3
  np.testing.assert_allclose(gu, fd_u, atol=2e-4 * scale_u)
  np.testing.assert_allclose(lam0[:, mjm.nv:2 * mjm.nv], fd_v0, atol=2e-4 * np.max(np.abs(fd_v0)))


def test_oracle_tangent_gradient_round_trip():
  """tangent_grad is the adjoint of integrate_pos's first-order change."""
  mjm = _humanoid()
  rng = np.random.default_rng(5)
  q0, _ = models.humanoid_rollout_state(mjm, 3)
  g_raw = rng.normal(size=q0.shape)
  dq = rng.normal(size=(3, mjm.nv)) * 1e-6
  lhs = np.sum(g_raw * (od.integrate_pos(mjm, q0, dq) - q0), -1)
  rhs = np.sum(od.tangent_grad(mjm, q0, g_raw) * dq, -1)
  np.testing.assert_allclose(lhs, rhs, rtol=1e-5, atol=1e-14)
  # differentiate_pos inverts integrate_pos
  np.testing.assert_allclose(od.differentiate_pos(mjm, od.integrate_pos(mjm, q0, dq * 1e4), q0), dq * 1e4,
                             atol=1e-10)


@pytest.mark.gpu
def test_transition_fd_free_sphere_closed_form():
  mjm, q, v = _sphere_case()
  m = mjw.put_model(mjm)
  d = mjw.make_data(mjm, nworld=2, nconmax=8, njmax=16)
  helpers.set_field(d, "qpos", q)
  helpers.set_field(d, "qvel", v)
  tfd = derivative.TransitionFD(mjm, m, 2, nconmax=8, njmax=16, eps=1e-3)
  A, B = tfd(d)
  torch.cuda.synchronize()
  A = A.cpu().numpy()
  for w in range(2):
    np.testing.assert_allclose(A[w], _sphere_A(mjm.opt.timestep, v[w, 3:]), atol=1e-4)
  # d itself was not stepped
  np.testing.assert_array_equal(helpers.data_numpy(d, "qpos"), q.astype(np.float32))


def _jac_errors(A, B, Ar, Br):
  n = A.shape[0]
  ea = [np.linalg.norm(A[w] - Ar[w]) / np.linalg.norm(Ar[w]) for w in range(n)]
  eb = [np.linalg.norm(B[w] - Br[w]) / np.linalg.norm(Br[w]) for w in range(n)]
  return np.array(ea), np.array(eb)


@pytest.mark.gpu
def test_transition_fd_humanoid_matches_oracle():
  """Device A, B (FP32, centered, eps 1e-3) against the FP64 oracle's at the
  same eps, for humanoids in contact with the floor: at the initial state
  (no warmstart) and one step later (warmstarted).  Equal eps on both sides:
  a contact or limit switching inside +-eps enters both the same way."""
  mjm = _humanoid()
  n = 4
  q0, v0, ctrl = _rollout_case(mjm, n, 2)
  step = helpers.oracle_stepper(mjm, mjw.put_model(mjm, device="cpu"))
  traj, inputs = od.rollout(step, q0, v0, np.zeros((n, 0)), ctrl, np.zeros((n, mjm.nv)))
  m = mjw.put_model(mjm)
  d = mjw.make_data(mjm, nworld=n, nconmax=64 * n, njmax=277 * n)
  tfd = derivative.TransitionFD(mjm, m, n, nconmax=64 * n, njmax=277 * n, eps=1e-3)
  for t, (q, v, _, u, w0) in enumerate(inputs):
    for f, x in (("qpos", q), ("qvel", v), ("ctrl", u), ("qacc_warmstart", w0)):
      helpers.set_field(d, f, x)
    A, B = tfd(d)
    torch.cuda.synchronize()
    A, B = A.cpu().numpy().astype(np.float64), B.cpu().numpy().astype(np.float64)
    Ar, Br = od.transition_fd(step, mjm, q, v, np.zeros((n, 0)), u, w0, 1e-3)
    ea, eb = _jac_errors(A, B, Ar, Br)
    fmt = {"float_kind": lambda x: f"{x:.1e}"}
    print(f"state {t}: |dA|/|A| {np.array2string(ea, formatter=fmt)}  |dB|/|B| {np.array2string(eb, formatter=fmt)}")
    assert np.all(ea < 1e-2) and np.all(eb < 1e-2), (t, ea, eb)


def _gpu_rollout_grads(mjm, q0, v0, ctrl, eps=4e-3):
  n, T = q0.shape[0], ctrl.shape[0]
  m = mjw.put_model(mjm)
  r = derivative.Rollout(mjm, m, n, T, nconmax=64 * n, njmax=277 * n, eps=eps, adjoint="fd")
  tq0 = torch.tensor(q0, dtype=torch.float32, device="cuda", requires_grad=True)
  tv0 = torch.tensor(v0, dtype=torch.float32, device="cuda", requires_grad=True)
  tu = torch.tensor(ctrl, dtype=torch.float32, device="cuda", requires_grad=True)
  qpos, qvel = r(tq0, tv0, tu)
  loss = ((qpos[..., 2] - 1.3) ** 2 + 0.1 * qpos[..., 4]).sum() + 1e-3 * (tu**2).sum()
  loss.backward()
  torch.cuda.synchronize()
  return tu.grad.cpu().numpy().astype(np.float64), tv0.grad.cpu().numpy().astype(np.float64)


def _rel(a, b):
  return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.gpu
def test_rollout_gradient_smooth_matches_oracle_fd():
  """Contacts and joint limits disabled (smooth dynamics: actuators, gravity,
  Coriolis, implicitfast): the device adjoint's gradients (FP32, eps 4e-3)
  with respect to ctrl and qvel0 against central differences (eps 1e-6) of
  whole rollouts of the FP64 oracle, to 1e-3 relative error."""
  mjm = _humanoid()
  mjm.opt.disableflags |= int(mjw.DisableBit.CONTACT) | int(mjw.DisableBit.LIMIT)
  n, T = 4, 5
  q0, v0, ctrl = _rollout_case(mjm, n, T)
  gu, gv0 = _gpu_rollout_grads(mjm, q0, v0, ctrl)
  step = helpers.oracle_stepper(mjm, mjw.put_model(mjm, device="cpu"))
  z0, w0 = np.zeros((n, 0)), np.zeros((n, mjm.nv))
  fd_u, fd_v0 = od.loss_grad_fd(step, _loss, q0, v0, z0, ctrl, w0, 1e-6)
  eu, ev = _rel(gu, fd_u), _rel(gv0, fd_v0)
  print(f"smooth: ctrl gradient rel. error {eu:.2e}, qvel0 gradient rel. error {ev:.2e}")
  assert eu < 1e-3 and ev < 1e-3, (eu, ev)


@pytest.mark.gpu
def test_rollout_gradient_smooth_8_worlds_50_steps():
  """The differentiable-rollout check at its stated size: 8 worlds of the
  rollout workload, 50 steps, contacts and joint limits disabled: the
  device adjoint's gradients (FP32, eps 4e-3) with respect to all 1050
  ctrl entries and qvel0 against central differences (eps 1e-6) of whole
  rollouts of the FP64 oracle.  qvel0 holds 1e-3; the ctrl gradient is
  held to 1e-2 (measured 4.4e-3): the FP32 finite-difference Jacobians
  carry ~1e-3 relative error in their qvel rows per step (the same error
  the float32 oracle's own Jacobians show against float64: the rounding of
  M^-1 in FP32, not the differencing), and the adjoint sums 50 of them."""
  mjm = _humanoid()
  mjm.opt.disableflags |= int(mjw.DisableBit.CONTACT) | int(mjw.DisableBit.LIMIT)
  n, T = 8, 50
  q0, v0, ctrl = _rollout_case(mjm, n, T)
  gu, gv0 = _gpu_rollout_grads(mjm, q0, v0, ctrl)
  step = helpers.oracle_stepper(mjm, mjw.put_model(mjm, device="cpu"), nconmax_w=4, njmax_w=4, nthreads=16)
  z0, w0 = np.zeros((n, 0)), np.zeros((n, mjm.nv))
  fd_u, fd_v0 = od.loss_grad_fd(step, _loss, q0, v0, z0, ctrl, w0, 1e-6)
  eu, ev = _rel(gu, fd_u), _rel(gv0, fd_v0)
  worst = max(_rel(gu[:, w], fd_u[:, w]) for w in range(n))
  print(f"smooth, 8 worlds x 50 steps: ctrl gradient rel. error {eu:.2e} (worst world {worst:.2e}), "
        f"qvel0 {ev:.2e}")
  assert eu < 1e-2 and ev < 1e-3, (eu, ev)


@pytest.mark.gpu
def test_rollout_gradient_in_contact_matches_oracle_adjoint():
  """Humanoids in contact with the floor: the device adjoint's gradients
  against the oracle's adjoint over FP64 Jacobians at the same eps (4e-3);
  the distance to the eps 1e-6 loss differences, which columns straddling a
  contact or limit switch inside +-eps set, is printed, not held."""
  mjm = _humanoid()
  n, T = 4, 5
  q0, v0, ctrl = _rollout_case(mjm, n, T)
  gu, gv0 = _gpu_rollout_grads(mjm, q0, v0, ctrl)
  step = helpers.oracle_stepper(mjm, mjw.put_model(mjm, device="cpu"))
  z0, w0 = np.zeros((n, 0)), np.zeros((n, mjm.nv))
  traj, inputs = od.rollout(step, q0, v0, z0, ctrl, w0)
  jac = [od.transition_fd(step, mjm, *inp, 4e-3) for inp in inputs]
  lam0, gu_r = od.adjoint(mjm, jac, inputs, traj, *_loss_grads(traj))
  gu_r = gu_r + 2e-3 * ctrl
  gv_r = lam0[:, mjm.nv:2 * mjm.nv]
  fd_u, fd_v0 = od.loss_grad_fd(step, _loss, q0, v0, z0, ctrl, w0, 1e-6)
  eu, ev = _rel(gu, gu_r), _rel(gv0, gv_r)
  print(f"in contact: vs oracle adjoint (eps 4e-3) ctrl {eu:.2e} qvel0 {ev:.2e}; "
        f"vs eps 1e-6 loss differences ctrl {_rel(gu, fd_u):.2e} qvel0 {_rel(gv0, fd_v0):.2e}")
  assert eu < 5e-2 and ev < 5e-2, (eu, ev)
