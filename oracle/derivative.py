"""CPU restatement of the transition derivatives and the adjoint recursion --
test infrastructure only (the product computes them in
mjx_warp_amd/csrc/derivative.hpp).

The reference has no gradients; the interface is MuJoCo's mjd_transitionFD
(engine_derivative_fd.c): state x = (dq [nv], qvel [nv], act [na]) with dq
in the tangent space of qpos (mj_integratePos / mj_differentiatePos: a
free / ball joint's rotation dof is a local-frame rotation vector, q ->
q * exp(v)), input u = ctrl.  Everything here works on batches of worlds
and takes the step as a callable
    step(qpos, qvel, act, ctrl, warm) -> (qpos', qvel', act', warm')
over [n, .] arrays (the tests pass the FP64 oracle's step), so that
the finite differences, the recursion and the loss can be checked against
each other:

* `transition_fd`: A, B by central (or forward) differences, control
  columns one-sided at the edge of ctrlrange;
* `adjoint`: the reverse recursion lam_t = A_t' lam_{t+1} + g(x_t),
  dL/du_t = B_t' lam_{t+1}, with raw qpos gradients turned into the
  tangent space (a rotation dof moves q by 0.5 q (0, e) per unit);
* `loss_grad_fd`: the loss gradient by central differences of whole
  rollouts, the check that pins the other two.
"""

from __future__ import annotations

import numpy as np

FREE, BALL, SLIDE, HINGE = 0, 1, 2, 3


def quat_mul(a, b):
  aw, ax, ay, az = np.moveaxis(a, -1, 0)
  bw, bx, by, bz = np.moveaxis(b, -1, 0)
  return np.stack([aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
                   aw * by - ax * bz + ay * bw + az * bx, aw * bz + ax * by - ay * bx + az * bw], -1)


def quat_exp(v):
  """Unit quaternion of the rotation vector v."""
  n = np.linalg.norm(v, axis=-1, keepdims=True)
  s = np.where(n > 0, np.sin(0.5 * n) / np.where(n > 0, n, 1.0), 0.5)
  return np.concatenate([np.cos(0.5 * n), v * s], -1)


def quat_log(q):
  """Rotation vector of the unit quaternion q (angle in (-pi, pi])."""
  v = q[..., 1:]
  s = np.linalg.norm(v, axis=-1, keepdims=True)
  ang = 2.0 * np.arctan2(s, q[..., :1])
  ang = np.where(ang > np.pi, ang - 2.0 * np.pi, ang)
  return np.where(s > 0, v * ang / np.where(s > 0, s, 1.0), 0.0)


def _joints(mjm):
  return [(int(mjm.jnt_type[j]), int(mjm.jnt_qposadr[j]), int(mjm.jnt_dofadr[j])) for j in range(mjm.njnt)]


def integrate_pos(mjm, qpos, dq):
  """qpos (+) dq over [n, nq] / [n, nv] (mj_integratePos with dt = 1)."""
  out = np.array(qpos, dtype=np.float64, copy=True)
  for jt, qa, da in _joints(mjm):
    if jt in (FREE, BALL):
      if jt == FREE:
        out[:, qa:qa + 3] += dq[:, da:da + 3]
        qa, da = qa + 3, da + 3
      q = quat_mul(out[:, qa:qa + 4], quat_exp(dq[:, da:da + 3]))
      out[:, qa:qa + 4] = q / np.linalg.norm(q, axis=-1, keepdims=True)
    else:
      out[:, qa] += dq[:, da]
  return out


def differentiate_pos(mjm, qa_, qb_):
  """qa (-) qb: the tangent vector dq with qb (+) dq = qa."""
  out = np.zeros((qa_.shape[0], mjm.nv))
  for jt, qa, da in _joints(mjm):
    if jt in (FREE, BALL):
      if jt == FREE:
        out[:, da:da + 3] = qa_[:, qa:qa + 3] - qb_[:, qa:qa + 3]
        qa, da = qa + 3, da + 3
      inv = qb_[:, qa:qa + 4] * np.array([1.0, -1.0, -1.0, -1.0])
      out[:, da:da + 3] = quat_log(quat_mul(inv, qa_[:, qa:qa + 4]))
    else:
      out[:, da] = qa_[:, qa] - qb_[:, qa]
  return out


def tangent_grad(mjm, qpos, g_raw):
  """Raw qpos gradient [n, nq] -> tangent gradient [n, nv] at qpos."""
  out = np.zeros((qpos.shape[0], mjm.nv))
  for jt, qa, da in _joints(mjm):
    if jt in (FREE, BALL):
      if jt == FREE:
        out[:, da:da + 3] = g_raw[:, qa:qa + 3]
        qa, da = qa + 3, da + 3
      for r in range(3):
        e = np.zeros(4)
        e[1 + r] = 1.0
        dq = 0.5 * quat_mul(qpos[:, qa:qa + 4], np.broadcast_to(e, (qpos.shape[0], 4)))
        out[:, da + r] = np.sum(dq * g_raw[:, qa:qa + 4], -1)
    else:
      out[:, da] = g_raw[:, qa]
  return out


def _ctrl_sides(mjm, u, eps):
  """(up, down) masks [n, nu]: control may move by +eps / -eps in range."""
  clamp_on = not (int(mjm.opt.disableflags) & (1 << 7))
  lim = np.asarray(mjm.actuator_ctrllimited, bool) & clamp_on
  lo, hi = np.asarray(mjm.actuator_ctrlrange, np.float64).T
  up = ~lim | (u + eps <= hi)
  down = ~lim | (u - eps >= lo)
  return up, down


def transition_fd(step, mjm, qpos, qvel, act, ctrl, warm, eps, centered=True):
  """A [n, nx, nx], B [n, nx, nu] of one step at each world's state."""
  n, nv, na, nu = qpos.shape[0], mjm.nv, mjm.na, mjm.nu
  nx = 2 * nv + na
  nc = nx + nu
  up, down = _ctrl_sides(mjm, ctrl, eps)
  # perturbed inputs: [n, 1 + 2 nc] worlds (nominal, then +eps / -eps per column)
  K = 1 + 2 * nc
  rep = lambda a: np.repeat(np.asarray(a, np.float64)[:, None], K, 1)
  Q, V, Ac, U, Wm = rep(qpos), rep(qvel), rep(act), rep(ctrl), rep(warm)
  for c in range(nc):
    for s, h in ((0, eps), (1, -eps)):
      k = 1 + 2 * c + s
      if c < nv:
        dq = np.zeros((n, nv))
        dq[:, c] = h
        Q[:, k] = integrate_pos(mjm, Q[:, k], dq)
      elif c < 2 * nv:
        V[:, k, c - nv] += h
      elif c < nx:
        Ac[:, k, c - 2 * nv] += h
      else:
        U[:, k, c - nx] += h
  flat = lambda a: a.reshape(n * K, -1)
  q1, v1, a1, _ = step(flat(Q), flat(V), flat(Ac), flat(U), flat(Wm))
  q1, v1, a1 = (x.reshape(n, K, -1) for x in (q1, v1, a1))

  def succ_diff(kp, km):
    dq = differentiate_pos(mjm, q1[:, kp], q1[:, km])
    return np.concatenate([dq, v1[:, kp] - v1[:, km], a1[:, kp] - a1[:, km]], -1)

  A = np.zeros((n, nx, nx))
  B = np.zeros((n, nx, nu))
  for c in range(nc):
    kp, km = 1 + 2 * c, 2 + 2 * c
    if c < nx:
      A[:, :, c] = succ_diff(kp, km) / (2 * eps) if centered else succ_diff(kp, 0) / eps
      continue
    a = c - nx
    for w in range(n):
      u_, d_ = up[w, a], down[w, a]
      if centered and u_ and d_:
        col = (succ_diff(kp, km)[w]) / (2 * eps)
      elif u_:
        col = succ_diff(kp, 0)[w] / eps
      elif d_:
        col = succ_diff(0, km)[w] / eps
      else:
        col = np.zeros(nx)
      B[w, :, a] = col
  return A, B


def rollout(step, qpos0, qvel0, act0, ctrl, warm0):
  """States after each of the T = len(ctrl) steps and each step's inputs."""
  q, v, a, w = (np.array(x, np.float64) for x in (qpos0, qvel0, act0, warm0))
  traj, inputs = [], []
  for t in range(ctrl.shape[0]):
    inputs.append((q, v, a, ctrl[t], w))
    q, v, a, w = step(q, v, a, ctrl[t], w)
    traj.append((q, v, a))
  return traj, inputs


def adjoint(mjm, jac, inputs, traj, g_qpos, g_qvel):
  """Reverse recursion.  jac[t] = (A_t, B_t) at inputs[t]; g_qpos[t] /
  g_qvel[t] are the loss gradients with respect to traj[t] (the state after
  step t).  Returns (lam_0 [n, nx], dL/du [T, n, nu]) of the dynamics part."""
  T = len(jac)
  nv = mjm.nv

  def g(t_state, qpos):
    gq = tangent_grad(mjm, qpos, g_qpos[t_state])
    z = np.zeros((qpos.shape[0], mjm.na))
    return np.concatenate([gq, g_qvel[t_state], z], -1)

  lam = g(T - 1, traj[T - 1][0])
  gu = np.zeros((T,) + inputs[0][3].shape)
  for t in range(T - 1, -1, -1):
    A, B = jac[t]
    gu[t] = np.einsum("wij,wi->wj", B, lam)
    nxt = np.einsum("wij,wi->wj", A, lam)
    if t > 0:
      nxt = nxt + g(t - 1, inputs[t][0])
    lam = nxt
  del nv
  return lam, gu


def loss_grad_fd(step, loss, qpos0, qvel0, act0, ctrl, warm0, eps):
  """Central differences of loss(traj, ctrl) -> [n] with respect to ctrl
  [T, n, nu] and qvel0 [n, nv], all perturbations stepped as one batch."""
  T, n, nu = ctrl.shape
  nv = qvel0.shape[1]
  P = 2 * (T * nu + nv)
  rep = lambda a: np.repeat(np.asarray(a, np.float64)[None], P, 0)  # [P, n, ...]
  Q, V, Ac, Wm = rep(qpos0), rep(qvel0), rep(act0), rep(warm0)
  U = np.repeat(np.asarray(ctrl, np.float64)[:, None], P, 1)  # [T, P, n, nu]
  p = 0
  for t in range(T):
    for a in range(nu):
      U[t, p, :, a] += eps
      U[t, p + 1, :, a] -= eps
      p += 2
  for i in range(nv):
    V[p, :, i] += eps
    V[p + 1, :, i] -= eps
    p += 2
  flat = lambda a: a.reshape(P * n, *a.shape[2:])
  traj, _ = rollout(step, flat(Q), flat(V), flat(Ac), U.reshape(T, P * n, nu), flat(Wm))
  traj = [tuple(x.reshape(P, n, -1) for x in s) for s in traj]
  L = np.stack([loss([tuple(x[j] for x in s) for s in traj], U[:, j]) for j in range(P)])  # [P, n]
  diff = (L[0::2] - L[1::2]) / (2 * eps)  # [P/2, n]
  g_ctrl = diff[:T * nu].reshape(T, nu, n).transpose(0, 2, 1)
  g_qvel0 = diff[T * nu:].T
  return g_ctrl, g_qvel0
