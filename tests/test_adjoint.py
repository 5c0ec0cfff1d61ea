"""The analytic step adjoint (mjh_adjoint_vjp) against the transition
Jacobians of the FP64 oracle taken by central differences (eps 1e-6,
oracle/derivative.py), and the analytic rollout gradients against central
differences of whole FP64 rollouts.

The vector-Jacobian products are checked on random adjoint vectors:
lam' A and lam' B of the oracle's (A, B) against the device's lam and
grad_ctrl.  The solver runs to convergence (tolerance 1e-15, 100
iterations): the analytic adjoint differentiates the constraint optimum,
and a truncated solve's output moves with the iteration count, not only
with the state (finite differences of a 50-iteration oracle at eps 1e-6 and
1e-5 differ by 3e-4 on deep limit violations).  The step's derivative at a
held active set is then exact up to FP32 rounding, in contact too.
"""

import numpy as np
import pytest
import torch

import mjx_warp_amd as mjw
from mjx_warp_amd import derivative, models
from oracle import derivative as od

import helpers

pytestmark = pytest.mark.gpu


def _humanoid(disable=0):
  mjm = models.humanoid(iterations=100, ls_iterations=50)
  mjm.opt.tolerance = 1e-15
  mjm.opt.ls_tolerance = 1e-15
  mjm.opt.disableflags |= disable
  return mjm


def _case(mjm, n, seed, qnoise=0.05):
  q0, v0 = models.humanoid_rollout_state(mjm, n, seed)
  rng = np.random.default_rng(seed + 100)
  q0[:, 7:] += rng.normal(0.0, qnoise, q0[:, 7:].shape)
  v0 = v0 + rng.normal(0, 0.1, v0.shape)
  u = rng.uniform(-1, 1, (n, mjm.nu))
  return q0, v0, u


def _device_vjp(mjm, q0, v0, u, lam):
  n, nv = q0.shape[0], mjm.nv
  m = mjw.put_model(mjm)
  d = mjw.make_data(mjm, nworld=n, nconmax=64 * n, njmax=277 * n)
  f = lambda a: torch.tensor(a, dtype=torch.float32, device="cuda")
  tq, tv = f(q0), f(v0)
  d.qpos.copy_(tq)
  d.qvel.copy_(tv)
  d.ctrl.copy_(f(u))
  d.qacc_warmstart.zero_()
  mjw.step_fused(m, d)
  lam_next = f(lam)
  out = torch.zeros((n, 2 * nv), dtype=torch.float32, device="cuda")
  gu = torch.zeros((n, mjm.nu), dtype=torch.float32, device="cuda")
  derivative.vjp_step(m, d, tq, tv, lam_next, out, gu)
  torch.cuda.synchronize()
  rows = helpers.data_numpy(d, "world_efc").reshape(n, 2)[:, 1]
  return out.cpu().numpy().astype(np.float64), gu.cpu().numpy().astype(np.float64), rows


def _oracle_jac(mjm, q0, v0, u):
  n = q0.shape[0]
  step = helpers.oracle_stepper(mjm, mjw.put_model(mjm, device="cpu"))
  return od.transition_fd(step, mjm, q0, v0, np.zeros((n, 0)), u, np.zeros((n, mjm.nv)), 1e-6)


def _errors(mjm, q0, v0, u, seed):
  n, nv = q0.shape[0], mjm.nv
  rng = np.random.default_rng(seed)
  A, B = _oracle_jac(mjm, q0, v0, u)
  out = []
  for trial in range(3):
    lam = rng.normal(size=(n, 2 * nv))
    got, gu, rows = _device_vjp(mjm, q0, v0, u, lam)
    ref = np.einsum("wi,wij->wj", lam, A)
    refu = np.einsum("wi,wij->wj", lam, B)
    for w in range(n):
      eq = np.linalg.norm(got[w, :nv] - ref[w, :nv]) / np.linalg.norm(ref[w, :nv])
      ev = np.linalg.norm(got[w, nv:] - ref[w, nv:]) / np.linalg.norm(ref[w, nv:])
      eu = np.linalg.norm(gu[w] - refu[w]) / np.linalg.norm(refu[w])
      out.append((w, int(rows[w]), eq, ev, eu))
  return out


def test_vjp_smooth_matches_oracle_fd():
  """Contacts and limits off: gravity, Coriolis, springs, dampers, 21
  motors, implicitfast: lam' A and lam' B to 1e-4 relative."""
  mjm = _humanoid(int(mjw.DisableBit.CONTACT) | int(mjw.DisableBit.LIMIT))
  q0, v0, u = _case(mjm, 4, 3)
  errs = _errors(mjm, q0, v0, u, 11)
  worst = np.max([e[2:] for e in errs], axis=0)
  print(f"smooth: dq columns {worst[0]:.2e}, qvel columns {worst[1]:.2e}, ctrl {worst[2]:.2e}")
  assert np.all(worst < 1e-4), worst


def test_vjp_with_joint_limits_matches_oracle_fd():
  """Contacts off, joint limits on, joints pushed past their limits: the
  limit rows' impedance and reference acceleration move with the joint
  angle, their forces are differentiated at the constraint optimum."""
  mjm = _humanoid(int(mjw.DisableBit.CONTACT))
  q0, v0, u = _case(mjm, 4, 5, qnoise=0.6)
  errs = _errors(mjm, q0, v0, u, 12)
  assert min(e[1] for e in errs) > 0, "the case needs active limit rows"
  worst = np.max([e[2:] for e in errs], axis=0)
  print(f"limits (rows {[e[1] for e in errs[:4]]}): dq columns {worst[0]:.2e}, qvel columns {worst[1]:.2e}, "
        f"ctrl {worst[2]:.2e}")
  assert np.all(worst < 1e-3), worst


def test_vjp_in_contact_against_oracle_fd():
  """Humanoids on the floor (plane-capsule feet, capsule-capsule self
  contacts): all columns to 1e-3."""
  mjm = _humanoid()
  q0, v0, u = _case(mjm, 4, 7)
  errs = _errors(mjm, q0, v0, u, 13)
  assert max(e[1] for e in errs) > 0, "the case needs contact rows"
  for w, rows, eq, ev, eu in errs[:4]:
    print(f"in contact world {w} ({rows} rows): dq columns {eq:.2e}, qvel columns {ev:.2e}, ctrl {eu:.2e}")
  worst = np.max([e[2:] for e in errs], axis=0)
  assert np.all(worst < 1e-3), worst


def _rollout_grads(mjm, q0, v0, ctrl, **kw):
  n, T = q0.shape[0], ctrl.shape[0]
  m = mjw.put_model(mjm)
  r = derivative.Rollout(mjm, m, n, T, nconmax=64 * n, njmax=277 * n, **kw)
  tq0 = torch.tensor(q0, dtype=torch.float32, device="cuda", requires_grad=True)
  tv0 = torch.tensor(v0, dtype=torch.float32, device="cuda", requires_grad=True)
  tu = torch.tensor(ctrl, dtype=torch.float32, device="cuda", requires_grad=True)
  qpos, qvel = r(tq0, tv0, tu)
  loss = ((qpos[..., 2] - 1.3) ** 2 + 0.1 * qpos[..., 4]).sum() + 1e-3 * (tu**2).sum()
  loss.backward()
  torch.cuda.synchronize()
  return (tu.grad.cpu().numpy().astype(np.float64), tv0.grad.cpu().numpy().astype(np.float64),
          tq0.grad.cpu().numpy().astype(np.float64))


def _loss(traj, ctrl):
  L = 0.0
  for t, (q, _, _) in enumerate(traj):
    L = L + (q[..., 2] - 1.3) ** 2 + 0.1 * q[..., 4] + 1e-3 * np.sum(ctrl[t] ** 2, -1)
  return L


def _rel(a, b):
  return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_analytic_rollout_smooth_8_worlds_50_steps():
  """The differentiable rollout at its stated size (8 worlds, 50 steps),
  contacts and limits off: the analytic adjoint's gradients with respect to
  all ctrl entries and qvel0 against central differences (eps 1e-6) of
  whole FP64 oracle rollouts, to 1e-3."""
  mjm = models.humanoid(iterations=10, ls_iterations=20)
  mjm.opt.disableflags |= int(mjw.DisableBit.CONTACT) | int(mjw.DisableBit.LIMIT)
  n, T = 8, 50
  q0, v0 = models.humanoid_rollout_state(mjm, n, 3)
  rng = np.random.default_rng(103)
  v0 = v0 + rng.normal(0, 0.1, v0.shape)
  ctrl = rng.uniform(-1, 1, (T, n, mjm.nu))
  gu, gv0, _ = _rollout_grads(mjm, q0, v0, ctrl)
  step = helpers.oracle_stepper(mjm, mjw.put_model(mjm, device="cpu"), nconmax_w=4, njmax_w=4, nthreads=16)
  z0, w0 = np.zeros((n, 0)), np.zeros((n, mjm.nv))
  fd_u, fd_v0 = od.loss_grad_fd(step, _loss, q0, v0, z0, ctrl, w0, 1e-6)
  eu, ev = _rel(gu, fd_u), _rel(gv0, fd_v0)
  print(f"analytic, smooth, 8 worlds x 50 steps: ctrl gradient rel. error {eu:.2e}, qvel0 {ev:.2e}")
  assert eu < 1e-3 and ev < 1e-3, (eu, ev)


def test_analytic_rollout_in_contact_reported():
  """Humanoids on the floor, 4 worlds x 5 steps: the analytic gradients
  against central differences (eps 1e-6) of whole FP64 oracle rollouts,
  printed beside the finite-difference adjoint's (eps 4e-3), which
  straddles contact activations; the analytic error is held to 1e-3
  (measured 1.1e-5 for ctrl against 0.22 for the finite-difference
  adjoint)."""
  mjm = models.humanoid(iterations=10, ls_iterations=20)
  n, T = 4, 5
  q0, v0 = models.humanoid_rollout_state(mjm, n, 3)
  rng = np.random.default_rng(103)
  v0 = v0 + rng.normal(0, 0.1, v0.shape)
  ctrl = rng.uniform(-1, 1, (T, n, mjm.nu))
  gu, gv0, _ = _rollout_grads(mjm, q0, v0, ctrl)
  gu_fd, gv0_fd, _ = _rollout_grads(mjm, q0, v0, ctrl, adjoint="fd")
  step = helpers.oracle_stepper(mjm, mjw.put_model(mjm, device="cpu"))
  z0, w0 = np.zeros((n, 0)), np.zeros((n, mjm.nv))
  fd_u, fd_v0 = od.loss_grad_fd(step, _loss, q0, v0, z0, ctrl, w0, 1e-6)
  eu, ev = _rel(gu, fd_u), _rel(gv0, fd_v0)
  print(f"in contact: analytic ctrl {eu:.2e} qvel0 {ev:.2e}; finite-difference adjoint ctrl "
        f"{_rel(gu_fd, fd_u):.2e} qvel0 {_rel(gv0_fd, fd_v0):.2e}")
  assert eu < 1e-3 and ev < 1e-3, (eu, ev)


def test_taped_and_recomputed_rollout_gradients_are_identical():
  """The taped reverse pass (the default), recomputation from every step's
  input (tape=False) and from a checkpoint every 4 steps give the same
  gradients bit for bit: the steps replay exactly, and the adjoint reads
  the same per-world fields."""
  mjm = models.humanoid(iterations=10, ls_iterations=20)
  n, T = 4, 10
  q0, v0 = models.humanoid_rollout_state(mjm, n, 3)
  ctrl = np.random.default_rng(5).uniform(-1, 1, (T, n, mjm.nu))
  a = _rollout_grads(mjm, q0, v0, ctrl)
  b = _rollout_grads(mjm, q0, v0, ctrl, tape=False)
  c = _rollout_grads(mjm, q0, v0, ctrl, checkpoint_every=4)
  for x, y, z in zip(a, b, c):
    assert np.array_equal(x, y)
    assert np.array_equal(x, z)


def test_vjp_chain_euler_matches_oracle_fd():
  """The 3-link hinge chain (Euler, joint limits, capsule links against the
  floor plane), 8 worlds with the chain hanging into the floor, to 1e-3.
  The adjoint differentiates the constraint optimum, whichever solver
  reached it; the reference is taken with Newton: the chain's CG at 200
  iterations leaves a world whose finite differences at eps 1e-6 and 1e-5
  differ by 1.2e-2 (truncation, not the state), while Newton's agree with
  the analytic derivative to 1e-8 in FP64."""
  mjm = models.chain(3)
  mjm.opt.solver = int(mjw.SolverType.NEWTON)
  mjm.opt.iterations = 100
  mjm.opt.tolerance = 1e-15
  mjm.opt.ls_tolerance = 1e-15
  n = 8
  q0, v0 = models.chain_initial_state(n, seed=4)
  q0 *= 0.3
  u = np.zeros((n, 0))
  errs = []
  A, B = _oracle_jac(mjm, q0, v0, u)
  rng = np.random.default_rng(21)
  lam = rng.normal(size=(n, 2 * mjm.nv))
  got, _, rows = _device_vjp(mjm, q0, v0, u, lam)
  ref = np.einsum("wi,wij->wj", lam, A)
  assert rows.max() > 0, "the case needs constraint rows"
  e = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
  print(f"chain (rows {rows.tolist()}): worst rel. error {e.max():.2e}")
  assert e.max() < 1e-3, e


def test_vjp_large_model_lds_path_matches_oracle_fd():
  """n_humanoids.xml (three attached humanoids, nv 81: past the register
  Cholesky's 32 rows, so both solves go through the LDS factor; a sparse
  model, so every actuator acts), lowered into the floor with 82 and 119
  constraint rows: the state columns to 5e-3, the control columns to 1e-3.
  The analytic derivative is exact here in FP64 (2e-8 against the FP64
  oracle's differences); the device's error is that of its FP32 forward
  state: with 119 stiff rows the FP32 optimum (forces, active set) sits
  ~1e-3 from the FP64 one (measured 1.7e-3), while the device product
  agrees with the same formula in FP64 on the device's own forward fields
  to 1.6e-4."""
  from mjx_warp_amd import mjcf

  mjm = mjcf.from_xml_string((models.ASSETS / "n_humanoids.xml").read_text(), base_dir=str(models.ASSETS))
  mjm.opt.iterations = 100
  mjm.opt.ls_iterations = 50
  mjm.opt.tolerance = 1e-15
  mjm.opt.ls_tolerance = 1e-15
  derivative.check_analytic(mjm)
  n = 2
  rng = np.random.default_rng(41)
  q0 = np.tile(mjm.qpos0, (n, 1))
  for k in range(3):
    q0[:, 28 * k + 2] -= 0.05
    q0[:, 28 * k + 7:28 * (k + 1)] += rng.normal(0, 0.1, (n, 21))
  v0 = rng.normal(0, 0.3, (n, mjm.nv))
  u = rng.uniform(-1, 1, (n, mjm.nu))
  step = helpers.oracle_stepper(mjm, mjw.put_model(mjm, device="cpu"), nthreads=16)
  A, B = od.transition_fd(step, mjm, q0, v0, np.zeros((n, 0)), u, np.zeros((n, mjm.nv)), 1e-6)
  lam = rng.normal(size=(n, 2 * mjm.nv))
  got, gu, rows = _device_vjp(mjm, q0, v0, u, lam)
  ref = np.einsum("wi,wij->wj", lam, A)
  refu = np.einsum("wi,wij->wj", lam, B)
  assert rows.min() > 0, "the case needs constraint rows"
  e = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
  eu = np.linalg.norm(gu - refu, axis=1) / np.linalg.norm(refu, axis=1)
  print(f"n_humanoids (rows {rows.tolist()}): state columns {e.max():.2e}, ctrl {eu.max():.2e}")
  assert e.max() < 5e-3 and eu.max() < 1e-3, (e, eu)
