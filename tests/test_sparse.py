"""Sparse mass matrix (`is_sparse`: nv >= 60 or jacobian="sparse";
support.py:29-32): qM / qLD in MuJoCo's sparse layout [nworld, 1, nM],
qLD the L'DL factor with qLDiagInv = 1 / D (smooth.py:311-350), and the
sparse Euler outputs (forward.py:171-193).

CPU (oracle): the sparse qM holds the dense matrix's entries along each
dof's ancestor chain; L' D L rebuilt from qLD equals M; the sparse Euler
matrix is M + dt diag(damping) and its factor rebuilds it; a sparse model
steps as its dense twin; implicitfast refuses a sparse model (forward.py:268).
GPU: parity of every sparse field and of the step with the oracle, for the
humanoid with jacobian="sparse" (32-lane teams) and for a 61-link chain
that is sparse by size (64-lane teams).
"""

import numpy as np
import pytest
import torch

import mjx_warp_amd as mjw
from mjx_warp_amd import io, models
from oracle import oracle

import helpers


def _sparse_humanoid(integrator=0):
  mjm = models.humanoid(iterations=10, ls_iterations=20, integrator=integrator)
  mjm.opt.jacobian = 1  # sparse
  return mjm


def _chain(n):
  mjm = models.chain(n)
  mjm.opt.iterations = 20
  return mjm


def _unpack(mjm, S):
  """Dense lower-triangular matrix from the sparse layout (entries along each
  dof's ancestor chain)."""
  nv = mjm.nv
  A = np.zeros((nv, nv))
  for i in range(nv):
    a, j = mjm.dof_Madr[i], i
    while j >= 0:
      A[i, j] = S[a]
      a += 1
      j = mjm.dof_parentid[j]
  return A


def _ldl_rebuild(mjm, qLD):
  low = _unpack(mjm, qLD)
  D = np.diag(np.diag(low))
  L = np.tril(low, -1) + np.eye(mjm.nv)
  return L.T @ D @ L, np.diag(D)


def _state(mjm, nworld, seed):
  rng = np.random.default_rng(seed)
  q = np.tile(mjm.key_qpos[0] if mjm.nkey else mjm.qpos0, (nworld, 1)).astype(np.float64)
  for j in range(mjm.njnt):
    if mjm.jnt_type[j] in (2, 3):
      q[:, mjm.jnt_qposadr[j]] += rng.normal(0, 0.1, nworld)
  v = rng.normal(0, 0.3, (nworld, mjm.nv))
  u = rng.uniform(-1, 1, (nworld, mjm.nu))
  return q, v, u


def _case(mjm, nworld, seed, device, real, nconmax_w=64, njmax_w=277):
  q, v, u = _state(mjm, nworld, seed)
  m = mjw.put_model(mjm, device=device)
  d = mjw.make_data(mjm, nworld=nworld, nconmax=nconmax_w * nworld, njmax=njmax_w * nworld, device=device)
  helpers.set_field(d, "qpos", q)
  helpers.set_field(d, "qvel", v)
  helpers.set_field(d, "ctrl", u)
  if device == "cpu":
    return m, d, helpers.host_model(m, real), helpers.host_data(m, d, real)
  return m, d


def test_sparse_layout_and_ldl_factor():
  mjm = _sparse_humanoid()
  assert io.is_sparse(mjm)
  m, d, hm, hd = _case(mjm, 3, 1, "cpu", np.float64)
  assert tuple(d.qM.shape) == (3, 1, mjm.nM) and tuple(d.qMd.shape) == (3, mjm.nv, mjm.nv)
  oracle.run("forward", hm, hd)
  for w in range(3):
    Md = hd["qMd"][w]
    np.testing.assert_allclose(_unpack(mjm, hd["qM"][w, 0]), np.tril(Md), rtol=0, atol=0)
    M, D = _ldl_rebuild(mjm, hd["qLD"][w, 0])
    np.testing.assert_allclose(M, Md, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(hd["qLDiagInv"][w], 1.0 / D, rtol=1e-12)


def test_sparse_euler_outputs():
  mjm = _sparse_humanoid(integrator=0)
  mjm.opt.disableflags &= ~int(mjw.DisableBit.EULERDAMP)  # the model file disables it
  m, d, hm, hd = _case(mjm, 2, 2, "cpu", np.float64)
  oracle.run("step", hm, hd)
  dt = mjm.opt.timestep
  for w in range(2):
    # qM_integration from the step's own (pre-integration) mass matrix
    Mi = _unpack(mjm, hd["qM_integration"][w, 0])
    Mi = Mi + np.tril(Mi, -1).T
    np.testing.assert_allclose(Mi, hd["qMd"][w] + dt * np.diag(mjm.dof_damping), rtol=1e-10, atol=1e-10)
    M, D = _ldl_rebuild(mjm, hd["qLD_integration"][w, 0])
    np.testing.assert_allclose(M, Mi, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(hd["qLDiagInv_integration"][w], 1.0 / D, rtol=1e-12)


def test_sparse_model_steps_like_its_dense_twin():
  sparse = _sparse_humanoid()
  dense = models.humanoid(iterations=10, ls_iterations=20, integrator=0)
  out = []
  for mjm in (sparse, dense):
    m, d, hm, hd = _case(mjm, 4, 3, "cpu", np.float64)
    for _ in range(3):
      oracle.run("step", hm, hd)
    out.append((hd["qpos"].copy(), hd["qvel"].copy()))
  np.testing.assert_allclose(out[0][0], out[1][0], rtol=0, atol=1e-12)
  np.testing.assert_allclose(out[0][1], out[1][1], rtol=0, atol=1e-10)


def test_implicit_refuses_a_sparse_model():
  mjm = _sparse_humanoid(integrator=3)
  m = mjw.put_model(mjm, device="cpu")
  d = mjw.make_data(mjm, nworld=1, device="cpu")
  with pytest.raises(AssertionError):
    mjw.implicit(m, d)


def _close(f, a, b, tol):
  a = np.asarray(a, np.float64)
  b = np.asarray(b, np.float64)
  assert a.shape == b.shape, f
  err = np.max(np.abs(a - b))
  assert err <= tol * (1 + np.max(np.abs(b))), (f, err)


def _rebuilt(mjm, d, f_ld, f_dinv):
  """L' D L from a device sparse factor, and the factor's 1 / D, per world."""
  out = []
  for w in range(d.nworld):
    M, D = _ldl_rebuild(mjm, helpers.data_numpy(d, f_ld)[w, 0].astype(np.float64))
    np.testing.assert_allclose(helpers.data_numpy(d, f_dinv)[w], 1.0 / D, rtol=1e-5)
    out.append(M)
  return np.stack(out)


@pytest.mark.gpu
def test_sparse_humanoid_step_parity():
  """Sparse humanoid (jacobian="sparse", Euler with its damping term,
  32-lane teams): every sparse field and the step against the oracle."""
  mjm = _sparse_humanoid(integrator=0)
  mjm.opt.disableflags &= ~int(mjw.DisableBit.EULERDAMP)
  nworld, capacity = 16, dict(nconmax_w=64, njmax_w=277)
  m, d = _case(mjm, nworld, 5, "cuda", np.float32, **capacity)
  mc, dc, hm, hd = _case(mjm, nworld, 5, "cpu", np.float32, **capacity)
  mjw.forward(m, d)
  torch.cuda.synchronize()
  oracle.run("forward", hm, hd)
  for f in ("qMd", "qM", "qLD", "qLDiagInv", "qacc_smooth"):
    _close(f, helpers.data_numpy(d, f), hd[f], 2e-4)
  mjw.step(m, d)
  torch.cuda.synchronize()
  oracle.run("step", hm, hd)
  for f, tol in (("qpos", 1e-4), ("qvel", 2e-3), ("qM_integration", 2e-4), ("qLD_integration", 2e-4),
                 ("qLDiagInv_integration", 2e-4)):
    _close(f, helpers.data_numpy(d, f), hd[f], tol)
  # the fused step from the same inputs
  m2, d2 = _case(mjm, nworld, 5, "cuda", np.float32, **capacity)
  mjw.step_fused(m2, d2)
  torch.cuda.synchronize()
  for f in ("qpos", "qvel", "qLD", "qLDiagInv_integration"):
    _close(f, helpers.data_numpy(d2, f), helpers.data_numpy(d, f), 2e-3)


@pytest.mark.gpu
def test_sparse_by_size_chain():
  """A 61-link hinge chain (nv >= 60: sparse by size, 64-lane teams), hung
  clear of the floor.  Its mass matrix is ill-conditioned (61 serial
  links), so two FP32 factorizations of slightly different inputs differ in
  their forward error; the checks are the matrix against the oracle and the
  factors' backward error: L' D L from the device qLD (and from the Euler
  step's qLD_integration) equals the device's own matrix, L L' from the
  dense factor qLDd equals qMd, and M qacc_smooth = qfrc_smooth."""
  from mjx_warp_amd import mjcf

  mjm = mjcf.from_xml_string(models.chain_xml(61, anchor_z=40.0))
  mjm.opt.iterations = 20
  assert mjm.nv == 61 and io.is_sparse(mjm)
  nworld, capacity = 4, dict(nconmax_w=8, njmax_w=128)
  m, d = _case(mjm, nworld, 5, "cuda", np.float32, **capacity)
  mc, dc, hm, hd = _case(mjm, nworld, 5, "cpu", np.float32, **capacity)
  mjw.forward(m, d)
  torch.cuda.synchronize()
  oracle.run("forward", hm, hd)
  for f in ("qMd", "qM"):
    _close(f, helpers.data_numpy(d, f), hd[f], 2e-4)
  Md = helpers.data_numpy(d, "qMd").astype(np.float64)
  scale = np.max(np.abs(Md))
  assert np.max(np.abs(_rebuilt(mjm, d, "qLD", "qLDiagInv") - Md)) <= 1e-5 * scale
  Ld = helpers.data_numpy(d, "qLDd").astype(np.float64)
  assert np.max(np.abs(Ld @ np.transpose(Ld, (0, 2, 1)) - Md)) <= 1e-5 * scale
  qacc = helpers.data_numpy(d, "qacc_smooth").astype(np.float64)
  f = helpers.data_numpy(d, "qfrc_smooth").astype(np.float64)
  r = np.einsum("wij,wj->wi", Md, qacc) - f
  assert np.max(np.abs(r)) <= 1e-3 * (1 + np.max(np.abs(f)))
  mjw.step(m, d)
  torch.cuda.synchronize()
  assert np.all(np.isfinite(helpers.data_numpy(d, "qvel")))
  Mi = _unpack(mjm, helpers.data_numpy(d, "qM_integration")[0, 0].astype(np.float64))
  Mi = Mi + np.tril(Mi, -1).T
  assert np.max(np.abs(_rebuilt(mjm, d, "qLD_integration", "qLDiagInv_integration")[0] - Mi)) <= 1e-5 * scale
