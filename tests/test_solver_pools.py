"""solve on constraint-row pools laid out the reference's way
(solver_test.py:134-326): rows tagged by efc_worldid anywhere in [0, nefc),
unused rows tagged -1, the library's own world_efc blocks stale.

Three humanoid worlds (keyframes 0, 2, 1 as in the reference's batch test;
keyframe 2 floats with no rows) go through forward, which builds the rows
and solves them from the library's per-world blocks.  Their rows are then written into a second Data in three
layouts -- world after world with the unused capacity after them (the
reference test's pool), merged so that the worlds' rows interleave while
each world's own rows keep their order, and shuffled at random -- and
solve() runs from the same warm start.  The first two must reproduce the
library's own solve bit for bit (same rows in the same order per world);
the shuffled pool sums in another order: Newton must agree to FP32
rounding, truncated CG by the reference's cost criterion.
efc_force is compared at each row's own position in the caller's pool.
"""

import numpy as np
import pytest
import torch

import mjx_warp_amd as mjw
from mjx_warp_amd import models

import helpers

pytestmark = pytest.mark.gpu

ROW_FIELDS = ("J", "D", "aref", "pos", "margin")


def _forwarded(solver, iterations, ls_iterations):
  mjm = models.humanoid(iterations=iterations, ls_iterations=ls_iterations, solver=solver)
  m = mjw.put_model(mjm)
  d = mjw.make_data(mjm, nworld=3, nconmax=64 * 3, njmax=277 * 3)
  keys = [0, 2, 1]
  q = np.stack([mjm.key_qpos[k] for k in keys])
  q[:, 2] -= 0.005  # into the floor: contact rows in the worlds that touch it
  helpers.set_field(d, "qpos", q)
  helpers.set_field(d, "qvel", np.stack([mjm.key_qvel[k] for k in keys]))
  mjw.forward(m, d)
  torch.cuda.synchronize()
  return mjm, m, d


def _pool(mjm, d, layout, rng):
  """A Data holding d's smooth quantities and its rows in `layout`;
  returns it and, per world, the pool positions of that world's rows."""
  nworld = d.nworld
  blocks = helpers.world_blocks(helpers.data_numpy(d, "world_efc"), nworld)
  nactive = sum(n for _, n in blocks)
  tag = np.concatenate([np.full(n, w) for w, (_, n) in enumerate(blocks)])
  local = np.concatenate([np.arange(n) for _, n in blocks])
  if layout == "blocks":
    order = np.arange(nactive)
  elif layout == "interleaved":
    # a random merge of the three worlds' row sequences
    seq = rng.permutation(tag)
    order = np.empty(nactive, int)
    nxt = {w: 0 for w in range(nworld)}
    starts = np.cumsum([0] + [n for _, n in blocks])
    for i, w in enumerate(seq):
      order[i] = starts[w] + nxt[w]
      nxt[w] += 1
  else:
    order = rng.permutation(nactive)
  njmax = 2 * nactive
  d2 = mjw.make_data(mjm, nworld=nworld, nconmax=d.nconmax, njmax=njmax)
  # forward's own solve started from a zero warm start, as make_data leaves it
  for f in ("qM", "qLD", "qacc_smooth", "qfrc_smooth"):
    getattr(d2, f).copy_(getattr(d, f))
  src = np.concatenate([np.arange(b, b + n) for b, n in blocks])[order]
  for f in ROW_FIELDS:
    full = helpers.data_numpy(d, "efc_" + f)[src]
    pad = np.zeros((njmax - nactive,) + full.shape[1:], full.dtype)
    helpers.set_field(d2, "efc_" + f, np.concatenate([full, pad]))
  helpers.set_field(d2, "efc_worldid", np.concatenate([tag[order], np.full(njmax - nactive, -1)]).astype(np.int32))
  helpers.set_field(d2, "nefc", np.array([nactive], np.int32))
  # stale blocks: what a caller that knows nothing of them leaves
  helpers.set_field(d2, "world_efc", rng.integers(0, nactive, (nworld, 2)).astype(np.int32))
  pos = [np.where(tag[order] == w)[0][np.argsort(local[order][tag[order] == w])] for w in range(nworld)]
  return d2, pos, blocks


@pytest.mark.parametrize("layout", ["blocks", "interleaved", "shuffled"])
@pytest.mark.parametrize("solver,iterations,ls_iterations", [(1, 25, 5), (2, 2, 4)])
def test_solve_reads_reference_layout_pools(layout, solver, iterations, ls_iterations):
  mjm, m, d = _forwarded(solver, iterations, ls_iterations)
  rng = np.random.default_rng(solver * 10 + len(layout))
  d2, pos, blocks = _pool(mjm, d, layout, rng)
  # keyframe 2 hangs in the air with no rows: an empty block, as in the reference's test
  assert sum(n > 0 for _, n in blocks) >= 2
  mjw.solve(m, d2)
  torch.cuda.synchronize()
  # the caller's row tags are as they were and no world is left flagged
  assert not np.any(helpers.data_numpy(d2, "world_flags"))
  wid = helpers.data_numpy(d2, "efc_worldid")[: sum(n for _, n in blocks)]
  assert sorted(wid.tolist()) == sorted(np.repeat(np.arange(3), [n for _, n in blocks]).tolist())
  exact = layout != "shuffled"
  force_ref = helpers.data_numpy(d, "efc_force")
  force = helpers.data_numpy(d2, "efc_force")
  if not exact and solver == 1:
    # truncated CG from rows summed in another order ends elsewhere in a
    # flat cost valley: the reference's criterion, the cost within 2.5%
    # (solver_test.py:127)
    c2, c1 = helpers.data_numpy(d2, "efc_cost"), helpers.data_numpy(d, "efc_cost")
    assert np.all(c2 <= 1.025 * c1 + 1e-6) and np.all(c1 <= 1.025 * c2 + 1e-6), (c2, c1)
    return
  for f in ("qacc", "qfrc_constraint"):
    a, b = helpers.data_numpy(d2, f), helpers.data_numpy(d, f)
    if exact:
      assert np.array_equal(a, b), (f, np.abs(a - b).max())
    else:
      np.testing.assert_allclose(a, b, atol=2e-3 * (1 + np.abs(b).max()), err_msg=f)
  for w, (b0, n) in enumerate(blocks):
    if n == 0:
      continue
    got, ref = force[pos[w]], force_ref[b0:b0 + n]
    if exact:
      assert np.array_equal(got, ref), (w, np.abs(got - ref).max())
    else:
      np.testing.assert_allclose(got, ref, atol=2e-3 * (1 + np.abs(ref).max()))
