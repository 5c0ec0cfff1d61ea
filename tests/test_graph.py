"""Graph-captured steps (mjx_warp_amd.StepGraph) against eager steps.

Two Data objects start from the same seeded state; one advances with eager
`step_fused` calls, the other by replaying a captured graph of the same
steps.  Every replay is checked against the eager result, with eager work
and synchronisations between the replays.  Each world's computation does
not depend on where its block lands in the contact / row pools, so the two
must agree bit for bit.  The lying-humanoid case has worlds with more than
64 constraint rows, which a second launch on a side stream solves beside the
main solve launch: the graph then holds a fork and a join.
"""

import numpy as np
import pytest
import torch

import mjx_warp_amd as mjw
from mjx_warp_amd import models

import helpers

pytestmark = pytest.mark.gpu


def _state(mjm, nworld, seed, noise=0.05):
  rng = np.random.default_rng(seed)
  d = mjw.make_data(mjm, nworld=nworld, nconmax=64 * nworld, njmax=277 * nworld)
  q = np.tile(mjm.key_qpos[0], (nworld, 1))
  q[:, 7:] += rng.normal(0, noise, (nworld, mjm.nq - 7))
  helpers.set_field(d, "qpos", q)
  helpers.set_field(d, "qvel", rng.normal(0, 0.5, (nworld, mjm.nv)))
  helpers.set_field(d, "ctrl", rng.uniform(-1, 1, (nworld, mjm.nu)))
  return d


@pytest.mark.parametrize("fused", [True, False])
def test_graph_replays_match_eager_steps(fused):
  mjm = models.humanoid(iterations=10, ls_iterations=20)
  m = mjw.put_model(mjm)
  d_eager = _state(mjm, 64, 3)
  d_graph = _state(mjm, 64, 3)
  g = mjw.StepGraph(m, d_graph, nstep=2, fused=fused)
  step = mjw.step_fused if fused else mjw.step
  for rep in range(3):
    step(m, d_eager)
    step(m, d_eager)
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    for f in ("qpos", "qvel", "qacc", "qacc_warmstart"):
      a = helpers.data_numpy(d_graph, f)
      b = helpers.data_numpy(d_eager, f)
      assert np.array_equal(a, b), (rep, f, np.abs(a - b).max())
    assert abs(d_graph.time - d_eager.time) < 1e-12


@pytest.mark.parametrize("fused", [True, False])
def test_graph_with_big_worlds_matches_eager(fused):
  mjm = models.humanoid(iterations=10, ls_iterations=20)
  q0 = mjm.key_qpos[0].copy()
  q0[:3] = [0.0, 0.0, 0.05]
  q0[3:7] = [np.cos(np.pi / 4), 0.0, np.sin(np.pi / 4), 0.0]
  q0[7:] = 0.0
  mjm.key_qpos[0] = q0
  m = mjw.put_model(mjm)
  d_eager = _state(mjm, 64, 5, noise=0.2)
  d_graph = _state(mjm, 64, 5, noise=0.2)
  step = mjw.step_fused if fused else mjw.step
  g = mjw.StepGraph(m, d_graph, nstep=1, fused=fused)
  for rep in range(3):
    step(m, d_eager)
    torch.cuda.synchronize()
    if rep == 0:
      rows = np.array([n for _, n in helpers.world_blocks(helpers.data_numpy(d_eager, "world_efc"), 64)])
      assert rows.max() > 64 and rows.min() <= 64, rows
    g.replay()
    torch.cuda.synchronize()
    for f in ("qpos", "qvel", "qacc", "qacc_warmstart"):
      a = helpers.data_numpy(d_graph, f)
      b = helpers.data_numpy(d_eager, f)
      assert np.array_equal(a, b), (rep, f, np.abs(a - b).max())


@pytest.mark.parametrize("fused", [True, False])
def test_fused_step_description_path_equals_argument_path(fused):
  """The fused step and the plain step's solve read the model from
  Data.step_desc (the default) or from their kernel arguments (step_desc
  left empty): the same state bit for bit, humanoids standing and lying
  (with worlds over 64 rows)."""
  step = mjw.step_fused if fused else mjw.step
  for lying in (False, True):
    mjm = models.humanoid(iterations=10, ls_iterations=20)
    if lying:
      q0 = mjm.key_qpos[0].copy()
      q0[:3] = [0.0, 0.0, 0.05]
      q0[3:7] = [np.cos(np.pi / 4), 0.0, np.sin(np.pi / 4), 0.0]
      q0[7:] = 0.0
      mjm.key_qpos[0] = q0
    m = mjw.put_model(mjm)
    d_desc = _state(mjm, 64, 7, noise=0.2)
    d_arg = _state(mjm, 64, 7, noise=0.2)
    d_arg.ndesc = 0
    d_arg.step_desc = torch.zeros(0, dtype=torch.int32, device="cuda")
    for _ in range(3):
      step(m, d_desc)
      step(m, d_arg)
    torch.cuda.synchronize()
    for f in ("qpos", "qvel", "qacc", "qacc_warmstart"):
      a = helpers.data_numpy(d_desc, f)
      b = helpers.data_numpy(d_arg, f)
      assert np.array_equal(a, b), (lying, f, np.abs(a - b).max())


def test_fused_step_description_path_equals_argument_path_64_lane():
  """The same for a model over 32 dofs (a 40-link chain hung so that its
  lower links rest on the floor): the one-launch fused step of 64-lane
  teams, from Data.step_desc and from its kernel arguments."""
  from mjx_warp_amd import mjcf

  mjm = mjcf.from_xml_string(models.chain_xml(40, anchor_z=19.0))
  mjm.opt.iterations = 20
  m = mjw.put_model(mjm)
  rng = np.random.default_rng(8)
  q = rng.uniform(-0.4, 0.4, (16, mjm.nq))
  v = rng.normal(0, 0.3, (16, mjm.nv))
  out = []
  for desc in (True, False):
    d = mjw.make_data(mjm, nworld=16, nconmax=64 * 16, njmax=256 * 16)
    helpers.set_field(d, "qpos", q)
    helpers.set_field(d, "qvel", v)
    if not desc:
      d.ndesc = 0
      d.step_desc = torch.zeros(0, dtype=torch.int32, device="cuda")
    for _ in range(3):
      mjw.step_fused(m, d)
    out.append(d)
  torch.cuda.synchronize()
  assert int(helpers.data_numpy(out[0], "nefc")[0]) > 0
  for f in ("qpos", "qvel", "qacc", "qacc_warmstart"):
    a, b = helpers.data_numpy(out[0], f), helpers.data_numpy(out[1], f)
    assert np.isfinite(a).all() and np.array_equal(a, b), (f, np.abs(a - b).max())
