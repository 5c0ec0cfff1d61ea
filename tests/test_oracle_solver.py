"""Pins the oracle's truncated solver iterates: the line searches of
solver.py:339-734 against what the mathematics of one search fixes.

Along a search direction s from qacc0, the primal cost of the constrained
problem (solver.py:75-173),

  f(a) = 1/2 (qacc - qacc_smooth)' M (qacc - qacc_smooth)
         + sum_r 1/2 D_r min(J_r qacc - aref_r, 0)^2,  qacc = qacc0 + a s,

is a convex piecewise quadratic in a whose breakpoints are the roots of
J_r qacc - aref_r.  The checks evaluate f in float64 with numpy from the
oracle's own rows (efc_J, efc_D, efc_aref, qM, qacc_smooth):

* the parallel line search (solver.py:667-734) returns the candidate of
  alpha_candidate = linspace(0, 1, nlsp) with the least f (exact index), and
  its per-candidate costs are f at the candidates;
* the iterative line search (solver.py:339-664) with many rounds returns
  the exact minimiser of f (found by enumerating the pieces between the
  breakpoints) when its first Newton point overshoots the minimiser, and
  otherwise a point between that Newton point and the minimiser (the
  reference's bracket then holds two points of negative slope);
* a truncated iterative search never raises f above f(0), and the solver's
  cost never increases from one iteration to the next;
* with no constraint rows, one Newton or CG iteration takes a = 1 and lands
  on qacc_smooth (the quadratic's minimiser).

The search direction of a solve is read from a zero-iteration run (the
context only: warm start, constraint update, gradient, first direction),
the step length from efc_alpha after one iteration.
"""

import copy

import numpy as np
import pytest

from mjx_warp_amd import models
from oracle import oracle

import helpers

NWORLD = 16


def _prepared(mjm, nworld=NWORLD, seed=0, lift=0.0, qvel_scale=0.5, warm_scale=5.0, njmax_w=277):
  """FP64 oracle data after the forward stages that precede the solver,
  for seeded humanoid states around the squat keyframe (pressed into the
  floor so that contacts and limit rows exist, or lifted clear of it)."""
  rng = np.random.default_rng(seed)
  mjm, m, d, hm, hd = helpers.load(mjm, nworld=nworld, njmax=njmax_w * nworld, nconmax=64 * nworld)
  q = np.tile(mjm.key_qpos[0], (nworld, 1))
  q[:, 2] += lift - 0.01
  q[:, 7:] += rng.normal(0, 0.15, (nworld, mjm.nq - 7))
  hd.arrays["qpos"][:] = q
  hd.arrays["qvel"][:] = rng.normal(0, qvel_scale, (nworld, mjm.nv))
  hd.arrays["ctrl"][:] = rng.uniform(-1, 1, (nworld, mjm.nu))
  hd.arrays["qacc_warmstart"][:] = rng.normal(0, warm_scale, (nworld, mjm.nv))
  for s in ("fwd_position", "fwd_velocity", "fwd_actuation", "fwd_acceleration"):
    oracle.run(s, hm, hd)
  return hm, hd


def _solve(hm, hd, iterations, ls_iterations=None, ls_parallel=None):
  hm = copy.deepcopy(hm)
  hd = copy.deepcopy(hd)
  hm.opt["iterations"] = iterations
  if ls_iterations is not None:
    hm.opt["ls_iterations"] = ls_iterations
    hm.sizes["nlsp"] = ls_iterations
    hm.arrays["alpha_candidate"] = np.linspace(0.0, 1.0, ls_iterations)
  if ls_parallel is not None:
    hm.opt["ls_parallel"] = int(ls_parallel)
  if hm.sizes["nlsp"] * hd.sizes["nworld"] != hd.arrays["efc_cost_candidate"].size:
    n = hd.sizes["nworld"]
    hd.arrays["efc_cost_candidate"] = np.zeros((n, hm.sizes["nlsp"]))
    hd.arrays["efc_quad_total_candidate"] = np.zeros((n, hm.sizes["nlsp"], 3))
  oracle.run("solve", hm, hd)
  return hm, hd


class Line:
  """f(a) of one world along s from qacc0, in float64."""

  def __init__(self, hd, w, qacc0, s):
    r0, nr = (int(x) for x in hd["world_efc"].reshape(-1, 2)[w])
    self.J = hd["efc_J"][r0:r0 + nr].astype(np.float64)
    self.D = hd["efc_D"][r0:r0 + nr].astype(np.float64)
    self.aref = hd["efc_aref"][r0:r0 + nr].astype(np.float64)
    self.M = hd["qM"][w].astype(np.float64)
    self.a_s = hd["qacc_smooth"][w].astype(np.float64)
    self.q0 = np.asarray(qacc0, np.float64)
    self.s = np.asarray(s, np.float64)
    self.jar0 = self.J @ self.q0 - self.aref
    self.jv = self.J @ self.s

  def f(self, a):
    q = self.q0 + a * self.s
    jar = self.jar0 + a * self.jv
    e = q - self.a_s
    return 0.5 * e @ self.M @ e + 0.5 * np.sum(self.D * np.minimum(jar, 0.0) ** 2)

  def slope(self, a):
    q = self.q0 + a * self.s
    jar = self.jar0 + a * self.jv
    return self.s @ self.M @ (q - self.a_s) + np.sum(self.D * self.jv * np.minimum(jar, 0.0))

  def argmin(self):
    """The exact minimiser: on each piece between consecutive breakpoints
    the active set is fixed and f a quadratic; f is convex, so the piece
    whose quadratic's vertex lies inside it holds the minimum."""
    with np.errstate(divide="ignore", invalid="ignore"):
      bp = np.where(self.jv != 0, -self.jar0 / self.jv, np.nan)
    bp = np.sort(bp[np.isfinite(bp)])
    edges = np.concatenate([[-np.inf], bp, [np.inf]])
    sMs = self.s @ self.M @ self.s
    sMe = self.s @ self.M @ (self.q0 - self.a_s)
    for lo, hi in zip(edges[:-1], edges[1:]):
      if hi <= lo:
        continue
      mid = 0.5 * (lo + hi) if np.isfinite(lo) and np.isfinite(hi) else (hi - 1.0 if np.isfinite(hi) else lo + 1.0)
      if not np.isfinite(mid):
        mid = 0.0
      act = (self.jar0 + mid * self.jv) < 0
      A = sMs + np.sum(self.D[act] * self.jv[act] ** 2)
      B = sMe + np.sum(self.D[act] * self.jv[act] * self.jar0[act])
      a = -B / A
      if lo - 1e-12 * (1 + abs(lo)) <= a <= hi + 1e-12 * (1 + abs(hi)):
        return a
    raise AssertionError("no piece holds the minimum")


def _first_search(hm, hd):
  """(qacc0, s) per world: the warm start and the first search direction."""
  _, h0 = _solve(hm, hd, 0)
  return h0["qacc"].astype(np.float64).copy(), h0["efc_search"].astype(np.float64).copy()


@pytest.mark.parametrize("solver", [1, 2])
@pytest.mark.parametrize("nlsp", [8, 17])
def test_parallel_line_search_is_the_candidate_argmin(solver, nlsp):
  mjm = models.humanoid(solver=solver)
  hm, hd = _prepared(mjm, seed=solver * 10 + nlsp)
  q0, s = _first_search(hm, hd)
  hm1, h1 = _solve(hm, hd, 1, ls_iterations=nlsp, ls_parallel=True)
  cand = np.linspace(0.0, 1.0, nlsp)
  checked = 0
  for w in range(NWORLD):
    line = Line(hd, w, q0[w], s[w])
    costs = np.array([line.f(a) for a in cand])
    # the per-candidate costs the search compared (solver.py:690-716)
    np.testing.assert_allclose(h1["efc_cost_candidate"][w], costs, rtol=1e-9, atol=1e-9 * (1 + np.abs(costs).max()))
    best = int(np.argmin(costs))
    order = np.sort(costs)
    if order[1] - order[0] <= 1e-9 * (1 + abs(order[0])):
      continue  # a tie within rounding: either index is right
    assert h1["efc_alpha"][w] == cand[best], (w, h1["efc_alpha"][w], cand[best], costs)
    np.testing.assert_allclose(h1["qacc"][w], q0[w] + cand[best] * s[w], rtol=1e-12, atol=1e-10)
    checked += 1
  assert checked >= NWORLD // 2


@pytest.mark.parametrize("solver,lifting", [(1, False), (2, False), (2, True)])
def test_iterative_line_search_minimises_inside_a_bracket(solver, lifting):
  """The search (solver.py:339-664) starts from p0 at a = 0 and the Newton
  point a_N = -f'(0)/f''(0+) of the quadratic active at 0.  Its initial
  "lo" is the point of smaller slope (solver.py:404-431); f is convex, so
  that is always p0, and "hi" is a_N.
  * a_N overshoots (f'(a_N) >= 0): [0, a_N] brackets the minimiser, every
    round moves an end onto a Newton point or the midpoint inside the
    bracket, and with enough rounds the search returns the exact minimiser
    (to gtol in slope);
  * a_N undershoots (f'(a_N) < 0: rows active at 0 turn inactive before
    a_N, so f'' drops): hi carries a negative slope, the pair brackets
    nothing, and the search only advances by Newton steps while the slope
    stays negative: it returns a point in [a_N, a*] -- short of the
    minimiser, never past it, never costlier than a_N.  MuJoCo C brackets
    differently; this is the reference's behaviour, kept for parity."""
  mjm = models.humanoid(solver=solver)
  hm, hd = _prepared(mjm, seed=solver, warm_scale=0.0 if lifting else 5.0)
  if lifting:
    # a large upward torso acceleration as the warm start: the contact rows
    # are inactive at a = 0 and turn active along the line, so a_N overshoots
    hd.arrays["qacc_warmstart"][:, 2] = 200.0
  q0, s = _first_search(hm, hd)
  _, h1 = _solve(hm, hd, 1, ls_iterations=60, ls_parallel=False)
  seen = {"bracket": 0, "undershoot": 0}
  for w in range(NWORLD):
    line = Line(hd, w, q0[w], s[w])
    a_star = line.argmin()
    a = float(h1["efc_alpha"][w])
    assert a_star > 0, "the first direction descends"
    act0 = line.jar0 < 0
    curv0 = line.s @ line.M @ line.s + np.sum(line.D[act0] * line.jv[act0] ** 2)
    a_n = -line.slope(0.0) / curv0
    gtol = float(h1["efc_gtol"][w])
    f_star = line.f(a_star)
    if line.slope(a_n) >= 0:
      seen["bracket"] += 1
      assert abs(a - a_star) <= 1e-6 * (1 + a_star) or abs(line.slope(a)) <= 2 * gtol, (w, a, a_star)
      assert line.f(a) <= f_star + 1e-9 * (line.f(0.0) - f_star) + 1e-12 * (1 + abs(f_star))
    else:
      seen["undershoot"] += 1
      tol = 1e-9 * (1 + a_star)
      assert a_n - tol <= a <= a_star + tol, (w, a_n, a, a_star)
      assert line.f(a) <= line.f(a_n) * (1 + 1e-12)
    np.testing.assert_allclose(h1["qacc"][w], q0[w] + a * s[w], rtol=1e-12, atol=1e-10)
  assert seen["bracket" if lifting else "undershoot"] >= NWORLD // 2, seen


@pytest.mark.parametrize("solver", [1, 2])
@pytest.mark.parametrize("ls_iterations", [1, 2, 4])
def test_truncated_iterative_search_never_raises_the_cost(solver, ls_iterations):
  mjm = models.humanoid(solver=solver)
  hm, hd = _prepared(mjm, seed=7 + ls_iterations)
  q0, s = _first_search(hm, hd)
  _, h1 = _solve(hm, hd, 1, ls_iterations=ls_iterations, ls_parallel=False)
  for w in range(NWORLD):
    line = Line(hd, w, q0[w], s[w])
    a = float(h1["efc_alpha"][w])
    assert line.f(a) <= line.f(0.0) * (1 + 1e-12) + 1e-12
    # the returned cost is the cost of the returned iterate (solver.py:75-173)
    np.testing.assert_allclose(h1["efc_cost"][w], line.f(a), rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("solver", [1, 2])
def test_solver_cost_never_increases_across_iterations(solver):
  mjm = models.humanoid(solver=solver)
  hm, hd = _prepared(mjm, seed=20 + solver)
  costs = []
  for it in range(1, 9):
    _, h = _solve(hm, hd, it, ls_iterations=4, ls_parallel=False)
    costs.append(h["efc_cost"].astype(np.float64).copy())
  costs = np.stack(costs)
  scale = 1e-12 * (1 + np.abs(costs).max(axis=0))
  assert np.all(np.diff(costs, axis=0) <= scale[None]), np.diff(costs, axis=0).max()


@pytest.mark.parametrize("solver", [1, 2])
@pytest.mark.parametrize("ls_parallel", [False, True])
def test_one_iteration_without_rows_lands_on_qacc_smooth(solver, ls_parallel):
  # lifted clear of the floor with joints inside their ranges: no rows, so
  # f is the quadratic with minimiser qacc_smooth; Newton (H = M) and CG
  # (M-preconditioned) both point at it from the warm start, and both line
  # searches take a = 1 (a candidate of linspace(0, 1, nlsp))
  mjm = models.humanoid(solver=solver)
  rng = np.random.default_rng(5)
  mjm, m, d, hm, hd = helpers.load(mjm, nworld=NWORLD, njmax=277 * NWORLD, nconmax=64 * NWORLD)
  q = np.tile(mjm.qpos0, (NWORLD, 1))
  q[:, 2] += 1.0
  hd.arrays["qpos"][:] = q
  hd.arrays["qvel"][:] = rng.normal(0, 0.5, (NWORLD, mjm.nv))
  hd.arrays["ctrl"][:] = rng.uniform(-1, 1, (NWORLD, mjm.nu))
  hd.arrays["qacc_warmstart"][:] = rng.normal(0, 5.0, (NWORLD, mjm.nv))
  for st in ("fwd_position", "fwd_velocity", "fwd_actuation", "fwd_acceleration"):
    oracle.run(st, hm, hd)
  assert int(hd["nefc"][0]) == 0
  _, h1 = _solve(hm, hd, 1, ls_iterations=5, ls_parallel=ls_parallel)
  np.testing.assert_allclose(h1["efc_alpha"], 1.0, rtol=1e-12)
  np.testing.assert_allclose(h1["qacc"], hd["qacc_smooth"], rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("real", [np.float32, np.float64])
def test_bench_workload_diverges_at_one_newton_iteration(real):
  """The headline workload (humanoid from the squat keyframe, gears 100,
  ctrl ~ U(-1, 1) resampled every step, implicitfast, dt 0.005) stepped by
  the oracle: with the reference benchmark's solver setting (Newton, 1
  iteration, 4 line-search rounds) nearly every world reaches NaN within
  60 steps, in float32 and in float64 alike -- the algorithm, not the
  rounding -- while the model file's 10 / 20 keeps every world finite."""
  nworld, nstep = 64, 60
  out = {}
  for its, lsi in ((1, 4), (10, 20)):
    mjm = models.humanoid(iterations=its, ls_iterations=lsi)
    mjm, m, d, hm, hd = helpers.load(mjm, nworld=nworld, nconmax=64 * nworld, njmax=277 * nworld, real=real)
    if real == np.float32:
      hm = helpers.host_model(m, real)
    hd.arrays["qpos"][:] = mjm.key_qpos[0][None]
    rng = np.random.default_rng(1)
    for _ in range(nstep):
      hd.arrays["ctrl"][:] = rng.uniform(-1, 1, (nworld, mjm.nu))
      oracle.run("step", hm, hd, nthreads=4)
    out[its] = int((~np.isfinite(hd["qpos"]).all(axis=1)).sum())
  assert out[1] >= 0.9 * nworld, out
  assert out[10] == 0, out
