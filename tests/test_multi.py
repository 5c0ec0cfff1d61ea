"""Heterogeneous batches (mjx_warp_amd.MultiStep / mjh_multi_*): several
models, each with its own worlds, advanced together.

CPU: the heterogeneous workload's models (chain + spheres in a walled box)
compile with inward-facing walls and spheres placed inside, and a sphere
thrown at a wall stays inside the box under the oracle's step.
GPU: one grouped step per stage over groups of different models (32-lane
box worlds and a humanoid, a 64-lane 40-link chain) equals each group's own
fused step bit for bit; a captured graph of grouped steps replays the eager
result; the box-world group matches the oracle's step; the autotuner picks a
mode and leaves the state as it found it.
"""

import numpy as np
import pytest
import torch

import mjx_warp_amd as mjw
from mjx_warp_amd import models, mjcf
from oracle import oracle

import helpers


def test_box_world_walls_and_placement():
  radii = (0.05, 0.08, 0.03)
  mjm = models.box_world(4, radii)
  assert mjm.nv == 4 + 6 * len(radii)
  m = mjw.put_model(mjm, device="cpu")
  d = mjw.make_data(mjm, nworld=1, device="cpu")
  hm, hd = helpers.host_model_exact(mjm, m), helpers.host_data(m, d)
  oracle.run("kinematics", hm, hd)
  xmat = hd["geom_xmat"][0].reshape(-1, 3, 3)
  xpos = hd["geom_xpos"][0]
  for g in range(1, 5):
    n = xmat[g][:, 2]  # plane normal
    inward = -np.array([xpos[g][0], xpos[g][1], 0.0])
    assert np.dot(n, inward) > 0, g  # faces the box centre
  q, _ = models.box_world_states(mjm, 4, radii, 50, np.random.default_rng(1))
  for k, r in enumerate(radii):
    p = q[:, 4 + 7 * k:4 + 7 * k + 3]
    assert np.all(np.abs(p[:, :2]) <= 0.5 - r + 1e-12) and np.all(p[:, 2] >= r)


def test_sphere_stays_in_box_under_oracle():
  mjm = models.box_world(2, (0.05,))
  m = mjw.put_model(mjm, device="cpu")
  d = mjw.make_data(mjm, nworld=1, nconmax=64, njmax=256, device="cpu")
  hm, hd = helpers.host_model_exact(mjm, m), helpers.host_data(m, d)
  hd["qpos"][0, 2:5] = [0.3, 0.3, 0.3]
  hd["qpos"][0, 5] = 1.0
  hd["qvel"][0, 2:5] = [4.0, 3.0, 0.0]  # towards the +x / +y walls
  for _ in range(300):
    oracle.run("step", hm, hd)
  p = hd["qpos"][0, 2:5]
  assert np.all(np.abs(p[:2]) < 0.5) and p[2] > 0.0, p


def _groups(device="cuda", seed=0):
  rng = np.random.default_rng(seed)
  specs = []
  for links, radii, n in ((2, (0.05,), 5), (8, (0.04, 0.06, 0.05, 0.09), 7), (3, (0.07, 0.03), 4)):
    mjm = models.box_world(links, radii)
    q, v = models.box_world_states(mjm, links, radii, n, rng)
    v = rng.normal(0, 0.5, v.shape)
    specs.append((mjm, n, q, v, 8 * (links + len(radii)), None))
  hum = models.humanoid(iterations=10, ls_iterations=20)
  qh = np.tile(hum.key_qpos[0], (3, 1))
  qh[:, 7:] += rng.normal(0, 0.05, (3, hum.nq - 7))
  specs.append((hum, 3, qh, rng.normal(0, 0.3, (3, hum.nv)), 64, rng.uniform(-1, 1, (3, hum.nu))))
  ch = mjcf.from_xml_string(models.chain_xml(40, anchor_z=25.0))
  ch.opt.iterations = 20
  qc = rng.uniform(-0.3, 0.3, (2, ch.nq))
  specs.append((ch, 2, qc, rng.normal(0, 0.1, (2, ch.nv)), 8, None))

  def make():
    out = []
    for mjm, n, q, v, ncon_w, u in specs:
      m = mjw.put_model(mjm, device=device)
      d = mjw.make_data(mjm, nworld=n, nconmax=ncon_w * n, njmax=(4 * ncon_w + mjm.njnt) * n, device=device)
      helpers.set_field(d, "qpos", q)
      helpers.set_field(d, "qvel", v)
      if u is not None:
        helpers.set_field(d, "ctrl", u)
      out.append((m, d))
    return out

  return specs, make


_FIELDS = ("qpos", "qvel", "qacc", "qacc_warmstart")


@pytest.mark.gpu
def test_grouped_step_equals_each_groups_fused_step():
  specs, make = _groups()
  ga, gb = make(), make()
  ms = mjw.MultiStep(ga, mode="grouped")
  for _ in range(3):
    ms.step()
    for m, d in gb:
      mjw.step_fused(m, d)
  torch.cuda.synchronize()
  for i, ((_, da), (_, db)) in enumerate(zip(ga, gb)):
    for f in _FIELDS:
      a, b = helpers.data_numpy(da, f), helpers.data_numpy(db, f)
      assert np.array_equal(a, b), (i, f, np.abs(a - b).max())
    assert da.time == db.time
  ms.close()


@pytest.mark.gpu
def test_grouped_step_graph_replays_match_eager():
  _, make = _groups(seed=1)
  ga, gb = make(), make()
  eager = mjw.MultiStep(ga)
  cap = mjw.MultiStep(gb)
  graph = torch.cuda.CUDAGraph()
  side = torch.cuda.Stream()
  side.wait_stream(torch.cuda.current_stream())
  with torch.cuda.graph(graph, stream=side):
    cap.step()
    cap.step()
  torch.cuda.current_stream().wait_stream(side)
  for rep in range(3):
    eager.step()
    eager.step()
    torch.cuda.synchronize()
    graph.replay()
    torch.cuda.synchronize()
    for i, ((_, da), (_, db)) in enumerate(zip(ga, gb)):
      for f in _FIELDS:
        a, b = helpers.data_numpy(da, f), helpers.data_numpy(db, f)
        assert np.array_equal(a, b), (rep, i, f, np.abs(a - b).max())


@pytest.mark.gpu
def test_grouped_step_matches_oracle():
  """Every group of one grouped step -- the 32-lane box worlds and
  humanoid, the 64-lane 40-link chain -- against the oracle's own step."""
  specs, make = _groups(seed=2)
  groups = make()
  host = []
  for (mjm, *_), (m, d) in zip(specs, groups):
    mc = mjw.put_model(mjm, device="cpu")
    host.append((helpers.host_model(mc, np.float32), helpers.host_data(m, d, np.float32)))
  mjw.MultiStep(groups).step()
  torch.cuda.synchronize()
  for i, ((hm, hd), (m, d)) in enumerate(zip(host, groups)):
    oracle.run("step", hm, hd)
    for f, tol in (("qpos", 1e-4), ("qvel", 2e-3)):
      a, b = helpers.data_numpy(d, f).astype(np.float64), hd[f].astype(np.float64)
      assert np.max(np.abs(a - b)) <= tol * (1 + np.max(np.abs(b))), (i, f, np.max(np.abs(a - b)))


@pytest.mark.gpu
def test_autotune_picks_a_mode_and_restores_state():
  _, make = _groups(seed=3)
  groups = make()
  before = [{f: helpers.data_numpy(d, f).copy() for f in ("qpos", "qvel", "qacc_warmstart")} for _, d in groups]
  ms = mjw.MultiStep(groups, autotune=True)
  assert ms.mode in ("grouped", "per_group")
  torch.cuda.synchronize()
  for (_, d), b in zip(groups, before):
    for f, v in b.items():
      assert np.array_equal(helpers.data_numpy(d, f), v), f
    assert d.time == 0.0
