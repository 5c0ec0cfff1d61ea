"""Parity of the HIP step with the oracle on identical seeded inputs.

Every field `forward` writes is compared after one forward pass (humanoid with
Newton, chain with CG), then qpos / qvel / qacc after single steps of the
plain and the fused step.  Contacts and constraint rows are compared per
world: each world's block of the pool (world_con / world_efc) holds its
entries in a deterministic order on both sides.

Tolerances (FP32 kernels vs the FP32 oracle, differences come from summation
order and fused multiply-adds only):
  kinematics / inertia / smooth forces: max |a - b| <= 2e-4 * (1 + max |b|)
  solver outputs (qacc, efc_force, qfrc_constraint): 2e-3 * (1 + max |b|)
"""

import numpy as np
import pytest
import torch

import mjx_warp_amd as mjw
from mjx_warp_amd import models
from oracle import oracle

import helpers

pytestmark = pytest.mark.gpu

SMOOTH_FIELDS = ["xpos", "xquat", "xmat", "xipos", "ximat", "xanchor", "xaxis", "geom_xpos", "geom_xmat",
                 "subtree_com", "cinert", "cdof", "crb", "qM", "qLD", "actuator_length", "actuator_moment",
                 "actuator_velocity", "cvel", "cdof_dot", "qfrc_bias", "qfrc_spring", "qfrc_damper",
                 "qfrc_passive", "actuator_force", "qfrc_actuator", "qfrc_smooth", "qacc_smooth"]
CONTACT_FIELDS = ["contact_dist", "contact_pos", "contact_frame", "contact_includemargin", "contact_friction",
                  "contact_solref", "contact_solimp", "contact_dim", "contact_geom"]
EFC_FIELDS = ["efc_J", "efc_D", "efc_aref", "efc_pos", "efc_margin"]


def _close(name, a, b, tol):
  a = np.asarray(a, np.float64)
  b = np.asarray(b, np.float64)
  assert a.shape == b.shape, (name, a.shape, b.shape)
  if a.size == 0:
    return
  err = np.max(np.abs(a - b))
  scale = 1.0 + np.max(np.abs(b))
  assert err <= tol * scale, f"{name}: max err {err:.3e} > {tol:.1e} * {scale:.3e}"


def make_case(mjm, nworld, seed, nconmax_w, njmax_w, key=None, qpos_noise=0.05, qvel_noise=0.5, ctrl=True):
  rng = np.random.default_rng(seed)
  qpos = np.tile(mjm.key_qpos[key] if key is not None else mjm.qpos0, (nworld, 1)).astype(np.float64)
  for j in range(mjm.njnt):
    qa = mjm.jnt_qposadr[j]
    if mjm.jnt_type[j] in (2, 3):
      qpos[:, qa] += rng.normal(0, qpos_noise, nworld)
  qvel = rng.normal(0, qvel_noise, (nworld, mjm.nv))
  u = rng.uniform(-1, 1, (nworld, mjm.nu)) if ctrl else np.zeros((nworld, mjm.nu))
  out = []
  for dev in ("cuda", "cpu"):
    m = mjw.put_model(mjm, device=dev)
    d = mjw.make_data(mjm, nworld=nworld, nconmax=nconmax_w * nworld, njmax=njmax_w * nworld, device=dev)
    helpers.set_field(d, "qpos", qpos)
    helpers.set_field(d, "qvel", qvel)
    helpers.set_field(d, "ctrl", u)
    out += [m, d]
  m, d, mc, dc = out
  hm = helpers.host_model(mc, np.float32)
  hd = helpers.host_data(mc, dc, np.float32)
  return m, d, hm, hd


def compare_forward(m, d, hd, tol_smooth=2e-4, tol_solver=2e-3):
  nworld = d.nworld
  for f in SMOOTH_FIELDS:
    _close(f, helpers.data_numpy(d, f), hd[f], tol_smooth)
  # collision: per-world contact blocks
  gcon = helpers.world_blocks(helpers.data_numpy(d, "world_con"), nworld)
  ocon = helpers.world_blocks(hd["world_con"], nworld)
  assert [n for _, n in gcon] == [n for _, n in ocon], "contacts per world differ"
  assert int(helpers.data_numpy(d, "ncon")[0]) == int(hd["ncon"][0])
  for f in CONTACT_FIELDS:
    _close(f, helpers.gather_world_rows(helpers.data_numpy(d, f), gcon),
           helpers.gather_world_rows(hd[f], ocon), tol_smooth)
  gefc = helpers.world_blocks(helpers.data_numpy(d, "world_efc"), nworld)
  oefc = helpers.world_blocks(hd["world_efc"], nworld)
  assert [n for _, n in gefc] == [n for _, n in oefc], "rows per world differ"
  assert int(helpers.data_numpy(d, "nefc")[0]) == int(hd["nefc"][0])
  for f in EFC_FIELDS:
    _close(f, helpers.gather_world_rows(helpers.data_numpy(d, f), gefc),
           helpers.gather_world_rows(hd[f], oefc), tol_smooth)
  ew = helpers.gather_world_rows(helpers.data_numpy(d, "efc_worldid"), gefc)
  assert np.array_equal(ew, np.repeat(np.arange(nworld), [n for _, n in gefc]))
  for f in ("qacc", "qfrc_constraint", "qacc_warmstart"):
    _close(f, helpers.data_numpy(d, f), hd[f], tol_solver)
  _close("efc_force", helpers.gather_world_rows(helpers.data_numpy(d, "efc_force"), gefc),
         helpers.gather_world_rows(hd["efc_force"], oefc), tol_solver)


@pytest.mark.parametrize("iterations,ls_iterations", [(10, 20), (1, 4)])
def test_forward_humanoid(iterations, ls_iterations):
  mjm = models.humanoid(iterations=iterations, ls_iterations=ls_iterations)
  m, d, hm, hd = make_case(mjm, 16, seed=1, nconmax_w=64, njmax_w=277, key=0)
  mjw.forward(m, d)
  torch.cuda.synchronize()
  oracle.run("forward", hm, hd)
  compare_forward(m, d, hd)


def test_forward_chain():
  mjm = models.chain(3)
  m, d, hm, hd = make_case(mjm, 64, seed=0, nconmax_w=8, njmax_w=35, qpos_noise=0.5, qvel_noise=0.1, ctrl=False)
  mjw.forward(m, d)
  torch.cuda.synchronize()
  oracle.run("forward", hm, hd)
  compare_forward(m, d, hd)


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("model", ["humanoid", "chain"])
def test_step(model, fused):
  if model == "humanoid":
    mjm = models.humanoid(iterations=10, ls_iterations=20)
    m, d, hm, hd = make_case(mjm, 16, seed=2, nconmax_w=64, njmax_w=277, key=0)
  else:
    mjm = models.chain(3)
    m, d, hm, hd = make_case(mjm, 64, seed=0, nconmax_w=8, njmax_w=35, qpos_noise=0.5, qvel_noise=0.1,
                             ctrl=False)
  (mjw.step_fused if fused else mjw.step)(m, d)
  torch.cuda.synchronize()
  oracle.run("step", hm, hd)
  _close("qpos", helpers.data_numpy(d, "qpos"), hd["qpos"], 1e-4)
  _close("qvel", helpers.data_numpy(d, "qvel"), hd["qvel"], 2e-3)
  _close("qacc", helpers.data_numpy(d, "qacc"), hd["qacc"], 2e-3)
  assert abs(d.time - mjm.opt.timestep) < 1e-12
