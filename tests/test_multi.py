"""Heterogeneous batches (mjx_warp_amd.MultiStep / mjh_multi_*): several
models, each with its own worlds, advanced together.

CPU: the heterogeneous workload's models (chain + spheres in a walled box)
compile with inward-facing walls and spheres placed inside, and a sphere
thrown at a wall stays inside the box under the oracle's step.
GPU: one grouped step per stage over groups of different models (32-lane
box worlds and a humanoid, a 64-lane 40-link chain) equals each group's own
fused step bit for bit; a captured graph of grouped steps replays the eager
result; the box-world group matches the oracle's step; plans with caller
choices (64-lane teams for the small models, no staged Jacobian rows,
buckets by LDS class) match the oracle too, buckets with the library's
choices change nothing, and invalid choices are refused; the autotuner
picks a mode and per-group choices and leaves the state as it found it.
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


def _oracle_check(specs, groups, host):
  for i, ((hm, hd), (m, d)) in enumerate(zip(host, groups)):
    oracle.run("step", hm, hd)
    for f, tol in (("qpos", 1e-4), ("qvel", 2e-3)):
      a, b = helpers.data_numpy(d, f).astype(np.float64), hd[f].astype(np.float64)
      assert np.max(np.abs(a - b)) <= tol * (1 + np.max(np.abs(b))), (i, f, np.max(np.abs(a - b)))


@pytest.mark.gpu
@pytest.mark.parametrize("choice", ["team64", "rj0", "rmax_small", "buckets"])
def test_tuned_plan_matches_oracle(choice):
  """A grouped step with caller choices, every group against the oracle:
  every group on 64-lane teams (the LDS solver for the nv <= 32 models),
  no Jacobian rows staged in LDS, an LDS row capacity of 8 (worlds past it
  go to the global-memory tail), or the library's choices in buckets."""
  specs, make = _groups(seed=4)
  groups = make()
  host = []
  for (mjm, *_), (m, d) in zip(specs, groups):
    mc = mjw.put_model(mjm, device="cpu")
    host.append((helpers.host_model(mc, np.float32), helpers.host_data(m, d, np.float32)))
  tunes = {"team64": [(64, 0, -1)] * len(groups), "rj0": [(0, 0, 0)] * len(groups),
           "rmax_small": [(64, 8, -1)] * len(groups), "buckets": None}[choice]
  ms = mjw.MultiStep(groups, tunes=tunes, buckets=choice in ("buckets", "team64"))
  ms.step()
  torch.cuda.synchronize()
  assert not any(np.any(helpers.data_numpy(d, "world_flags")) for _, d in groups)
  _oracle_check(specs, groups, host)
  ms.close()


@pytest.mark.gpu
def test_buckets_with_library_choices_change_nothing():
  _, make = _groups(seed=5)
  ga, gb = make(), make()
  a, b = mjw.MultiStep(ga), mjw.MultiStep(gb, buckets=True)
  for _ in range(3):
    a.step()
    b.step()
  torch.cuda.synchronize()
  for i, ((_, da), (_, db)) in enumerate(zip(ga, gb)):
    for f in _FIELDS:
      x, y = helpers.data_numpy(da, f), helpers.data_numpy(db, f)
      assert np.array_equal(x, y), (i, f, np.abs(x - y).max())


@pytest.mark.gpu
@pytest.mark.parametrize("buckets", [False, True])
def test_work_queue_launches_change_nothing(buckets):
  """Work-queue launches (a grid of a few workgroups per SIMD taking worlds
  from a device counter) compute the same step bit for bit as one
  workgroup per world, 32- and 64-lane groups, with and without buckets;
  also when the plan's step is captured in a graph and replayed."""
  _, make = _groups(seed=7)
  ga, gb = make(), make()
  a, b = mjw.MultiStep(ga, buckets=buckets), mjw.MultiStep(gb, buckets=buckets, queue=True)
  for _ in range(3):
    a.step()
    b.step()
  torch.cuda.synchronize()
  g = torch.cuda.CUDAGraph()
  side = torch.cuda.Stream()
  side.wait_stream(torch.cuda.current_stream())
  with torch.cuda.graph(g, stream=side):
    b.step()
  torch.cuda.current_stream().wait_stream(side)
  for _ in range(2):
    a.step()
    g.replay()
    torch.cuda.synchronize()
    for i, ((_, da), (_, db)) in enumerate(zip(ga, gb)):
      for f in _FIELDS:
        x, y = helpers.data_numpy(da, f), helpers.data_numpy(db, f)
        assert np.array_equal(x, y), (i, f, np.abs(x - y).max())


@pytest.mark.gpu
def test_invalid_choices_are_refused():
  _, make = _groups(seed=6)
  groups = make()
  n = len(groups)
  with pytest.raises(ValueError, match="team must be"):
    mjw.MultiStep(groups, tunes=[(48, 0, -1)] * n)
  # the humanoid (nv 27) may take 32 lanes, the 40-link chain may not
  with pytest.raises(ValueError, match="nv <= 32"):
    mjw.MultiStep(groups, tunes=[(32, 0, -1)] * n)
  with pytest.raises(ValueError, match="one"):
    mjw.MultiStep(groups, tunes=[(0, 0, -1)])


def _oracle_runs(mjm, n, q, v, ncon_w, njmax_w, rng):
  """The oracle's step of one group in float32, float64 and twice in float32
  from hinge angles / qvel moved by 1e-6: (qvel, cost) per run."""
  hinge = [mjm.jnt_qposadr[j] for j in range(mjm.njnt) if mjm.jnt_type[j] in (2, 3)]
  out = []
  for i, real in enumerate((np.float32, np.float64, np.float32, np.float32)):
    mc = mjw.put_model(mjm, device="cpu")
    dc = mjw.make_data(mjm, nworld=n, nconmax=ncon_w * n, njmax=njmax_w * n, device="cpu")
    qq, vv = q.copy(), v.copy()
    if i >= 2:
      qq[:, hinge] += 1e-6 * rng.normal(size=(n, len(hinge)))
      vv = vv * (1 + 1e-6 * rng.normal(size=vv.shape))
    helpers.set_field(dc, "qpos", qq)
    helpers.set_field(dc, "qvel", vv)
    hm = helpers.host_model(mc, real) if real == np.float32 else helpers.host_model_exact(mjm, mc, real)
    hd = helpers.host_data(mc, dc, real)
    oracle.run("step", hm, hd, nthreads=8)
    out.append(dict(qpos=hd["qpos"].astype(np.float64), qvel=hd["qvel"].astype(np.float64),
                    cost=hd["efc_cost"].astype(np.float64), nefc=int(hd["nefc"][0])))
  return out


@pytest.mark.gpu
def test_stated_size_heterogeneous_worlds_match_oracle():
  """Worlds of the heterogeneous workload at its stated size: chains up to
  32 links and up to 64 spheres (nv up to 416: the Newton Hessian of the
  largest does not fit a workgroup's LDS and lives in efc_h), spheres
  resting on the floor, stepped together in one grouped step and each
  group against the oracle.  qpos to 1e-4; qvel to 2e-3 or within twice
  the oracle's own spread (its float64 run and two float32 runs from inputs
  moved by 1e-6) -- the 416-dof worlds' 20 Newton iterations over ~670
  rows end in a flat cost valley where the float32 and float64 oracle
  runs are 0.17 apart in qvel at costs equal to 3e-7 -- or, failing both,
  with the cost within 2.5% of the oracle's (solver_test.py:127)."""
  rng = np.random.default_rng(7)
  specs, groups = [], []
  for links, nsph, n in ((32, 64, 2), (20, 30, 2), (12, 9, 3), (5, 3, 3)):
    radii = tuple(float(r) for r in rng.uniform(0.03, 0.1, nsph))
    mjm = models.box_world(links, radii)
    q, _ = models.box_world_states(mjm, links, radii, n, rng)
    for k, r in enumerate(radii):
      q[:, links + 7 * k + 2] = r - 0.002  # on the floor
    v = rng.normal(0, 0.3, (n, mjm.nv))
    ncon_w = 3 * (links + nsph)
    njmax_w = 4 * ncon_w + mjm.njnt
    m = mjw.put_model(mjm)
    d = mjw.make_data(mjm, nworld=n, nconmax=ncon_w * n, njmax=njmax_w * n)
    helpers.set_field(d, "qpos", q)
    helpers.set_field(d, "qvel", v)
    specs.append((mjm, n, q, v, ncon_w, njmax_w))
    groups.append((m, d))
  assert max(int(m.nv) for m, _ in groups) == 416
  ms = mjw.MultiStep(groups)
  ms.step()
  torch.cuda.synchronize()
  by_cost = 0
  for i, ((mjm, n, q, v, ncon_w, njmax_w), (m, d)) in enumerate(zip(specs, groups)):
    runs = _oracle_runs(mjm, n, q, v, ncon_w, njmax_w, rng)
    ref = runs[0]
    assert ref["nefc"] > 0, i
    assert int(helpers.data_numpy(d, "nefc")[0]) == ref["nefc"], i
    qp = helpers.data_numpy(d, "qpos").astype(np.float64)
    assert np.max(np.abs(qp - ref["qpos"])) <= 1e-4 * (1 + np.max(np.abs(ref["qpos"]))), (i, "qpos")
    qv = helpers.data_numpy(d, "qvel").astype(np.float64)
    cost = helpers.data_numpy(d, "efc_cost").astype(np.float64)
    err = np.abs(qv - ref["qvel"]).max(axis=1)
    spread = np.max([np.abs(r["qvel"] - ref["qvel"]).max(axis=1) for r in runs[1:]], axis=0)
    ok = (err <= 2e-3 * (1 + np.abs(ref["qvel"]).max())) | (err <= 2.0 * spread)
    costs = np.stack([r["cost"] for r in runs])
    held = ~ok & (cost <= 1.025 * costs.max(0) + 1e-6) & (cost >= costs.min(0) / 1.025 - 1e-6)
    assert np.all(ok | held), (i, err, spread)
    by_cost += int(held.sum())
  print(f"stated-size worlds: {by_cost} held by the cost criterion")
  ms.close()


@pytest.mark.gpu
def test_autotune_picks_a_mode_and_restores_state():
  _, make = _groups(seed=3)
  groups = make()
  before = [{f: helpers.data_numpy(d, f).copy() for f in ("qpos", "qvel", "qacc_warmstart")} for _, d in groups]
  ms = mjw.MultiStep(groups, autotune=True)
  assert ms.mode in ("grouped", "per_group")
  if ms.mode == "grouped":
    assert len(ms.tunes) == len(groups)
    for (m, _), t in zip(groups, ms.tunes):
      assert t in mjw.multi.candidates(m), t
  torch.cuda.synchronize()
  for (_, d), b in zip(groups, before):
    for f, v in b.items():
      assert np.array_equal(helpers.data_numpy(d, f), v), f
    assert d.time == 0.0
