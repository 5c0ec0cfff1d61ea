"""Pins the oracle (the CPU restatement of the reference step) against the
reference's own known answers and against what the mathematics fixes.

Reference-held answers: math_test.py:25-88 (closest segment points),
broad_phase_test.py:45-242 (SAP and nxn pair sets), forward_test.py:130-145
(Euler without implicit damping).  Mathematical checks: the mass matrix
against a Jacobian-sum formula, the Cholesky factor and solve against
numpy, bias forces against gravity and finite-difference Coriolis terms,
contacts against closed forms, the converged solver against the unique
minimum of the convex primal cost, free fall against the discrete closed
form, and the SAP pair set against brute force.
"""

import numpy as np
import pytest
import scipy.optimize

from mjx_warp_amd import mjcf, models
from oracle import oracle

import helpers


# ---------------------------------------------------------------------------
# math_test.py:25-88 known answers
@pytest.mark.parametrize("real", [np.float32, np.float64])
def test_closest_segment_points_known_answers(real):
  cases = [
    (([0.73432405, 0.12372768, 0.20272314], [1.10600128, 0.88555209, 0.65209485],
      [0.85599262, 0.61736299, 0.9843583], [1.84270939, 0.92891793, 1.36343326]),
     [1.09063, 0.85404, 0.63351], [0.99596, 0.66156, 1.03813], 5),
    (([0, 0, -1], [0, 0, 1], [-1, 0, 0], [1, 0, 0]), [0, 0, 0], [0, 0, 0], 5),
    (([0.2, 0.2, 0], [1, 1, 0], [0.2, 0.4, 0], [1, 2, 0]), [0.3, 0.3, 0], [0.2, 0.4, 0], 2),
    (([0, 0, -1], [0, 0, 1], [1, 0, -1], [1, 0, 1]), [0, 0, 0], [1, 0, 0], 5),
    (([0, 0, -1], [0, 0, 1], [1, 0, 1], [1, 0, 3]), [0, 0, 1], [1, 0, 1], 5),
    (([0, 0, -1], [0, 0, -1], [1, 0, 0.1], [1, 0, 0.1]), [0, 0, -1], [1, 0, 0.1], 5),
    (([0, 0, -1], [0, 0, 1], [0, 0, -1], [0, 0, 1]), [0, 0, 0], [0, 0, 0], 5),
  ]
  for args, ea, eb, places in cases:
    a, b = oracle.closest_segments(*args, real=real)
    np.testing.assert_allclose(a, ea, atol=1.5 * 10.0 ** -places)
    np.testing.assert_allclose(b, eb, atol=1.5 * 10.0 ** -places)


# ---------------------------------------------------------------------------
# broad_phase_test.py:45-120: sweep and prune finds exactly these 19 pairs
SAP_SCENE = """
<mujoco>
  <worldbody>
    <geom size="40 40 40" type="plane"/>
    <body pos="0 0 0.7"><freejoint/><geom size="0.5 0.5 0.5" type="box"/></body>
    <body pos="0.1 0 0.7"><freejoint/><geom size="0.5 0.5 0.5" type="box"/></body>
    <body pos="1.8 0 0.7"><freejoint/><geom size="0.5 0.5 0.5" type="box"/></body>
    <body pos="1.6 0 0.7"><freejoint/><geom size="0.5 0.5 0.5" type="box"/></body>
    <body pos="0 0 1.8"><freejoint/><geom size="0.5 0.5 0.5" type="box"/>
      <geom size="0.5 0.5 0.5" type="box" pos="0 0 -1"/></body>
    <body pos="0 0.5 1.2"><freejoint/><geom size="0.5 0.5 0.5" type="box"/></body>
  </worldbody>
</mujoco>
"""
SAP_PAIRS = [(0, 1), (0, 2), (0, 3), (0, 4), (0, 6), (1, 2), (1, 4), (1, 5), (1, 6), (1, 7), (2, 3), (2, 4),
             (2, 5), (2, 6), (2, 7), (3, 4), (4, 6), (5, 7), (6, 7)]


def test_sap_known_pairs():
  mjm, m, d, hm, hd = helpers.load(SAP_SCENE)
  oracle.run("kinematics", hm, hd)
  oracle.run("sap_broadphase", hm, hd)
  assert int(hd["ncollision"][0]) == 19
  assert helpers.pairs_of(hd) == SAP_PAIRS


# broad_phase_test.py:124-242
NXN_SCENE = """
<mujoco>
  <worldbody>
    <body><freejoint/><geom type="sphere" size="0.1"/></body>
    <body><freejoint/><geom type="sphere" size="0.1"/></body>
    <body><freejoint/><geom type="capsule" size="0.1 0.1"/></body>
    <body><freejoint/><geom type="sphere" size="0.1"/></body>
    <body><freejoint/><geom type="sphere" size="0.1"/><geom type="sphere" size="0.1"/>
      <body><geom type="sphere" size="0.1"/><joint type="hinge"/></body></body>
  </worldbody>
  <keyframe>
    <key qpos='0 0 0 1 0 0 0 1 0 0 1 0 0 0 2 0 0 1 0 0 0 3 0 0 1 0 0 0 4 0 0 1 0 0 0 0'/>
    <key qpos='0 0 0 1 0 0 0 .05 0 0 1 0 0 0 2 0 0 1 0 0 0 3 0 0 1 0 0 0 4 0 0 1 0 0 0 0'/>
    <key qpos='0 0 0 1 0 0 0 .01 0 0 1 0 0 0 .02 0 0 1 0 0 0 3 0 0 1 0 0 0 4 0 0 1 0 0 0 0'/>
    <key qpos='0 0 0 1 0 0 0 1 0 0 1 0 0 0 2 0 0 1 0 0 0 2 0 0 1 0 0 0 4 0 0 1 0 0 0 0'/>
  </keyframe>
</mujoco>
"""


def _nxn(key, nworld=1, contype=None):
  mjm = mjcf.from_xml_string(NXN_SCENE)
  if contype is not None:
    mjm.geom_contype[:] = contype
  mjm, m, d, hm, hd = helpers.load(mjm, nworld=nworld, keyframe=key)
  oracle.run("kinematics", hm, hd)
  oracle.run("nxn_broadphase", hm, hd)
  return hd


def test_nxn_known_pairs():
  assert int(_nxn(0)["ncollision"][0]) == 0
  hd = _nxn(1)
  assert int(hd["ncollision"][0]) == 1
  assert hd["collision_pair"][0].tolist() == [0, 1]
  hd = _nxn(2)
  assert hd["collision_pair"][:3].tolist() == [[0, 1], [0, 2], [1, 2]]
  # contype / conaffinity mismatch: no pairs
  assert int(_nxn(1, contype=0)["ncollision"][0]) == 0
  # pairs are stored with geom_type[g1] <= geom_type[g2]: capsule(2) after sphere(3)
  hd = _nxn(3)
  assert int(hd["ncollision"][0]) == 1
  assert hd["collision_pair"][0].tolist() == [3, 2]


def test_nxn_two_worlds():
  mjm = mjcf.from_xml_string(NXN_SCENE)
  mjm, m, d, hm, hd = helpers.load(mjm, nworld=2)
  hd.arrays["qpos"][0] = mjm.key_qpos[1]
  hd.arrays["qpos"][1] = mjm.key_qpos[2]
  oracle.run("kinematics", hm, hd)
  oracle.run("nxn_broadphase", hm, hd)
  assert int(hd["ncollision"][0]) == 4
  assert hd["collision_pair"][:4].tolist() == [[0, 1], [0, 1], [0, 2], [1, 2]]


def _brute_pairs(mjm, hd, w, sap_bounds):
  """All filtered pairs whose bounding volumes overlap, by brute force."""
  out = []
  gx = hd["geom_xpos"][w]
  gm = hd["geom_xmat"][w].reshape(-1, 3, 3)
  for a in range(mjm.ngeom):
    for b in range(a + 1, mjm.ngeom):
      wa, wb = mjm.body_weldid[mjm.geom_bodyid[a]], mjm.body_weldid[mjm.geom_bodyid[b]]
      if wa == wb:
        continue
      pa, pb = mjm.body_weldid[mjm.body_parentid[wa]], mjm.body_weldid[mjm.body_parentid[wb]]
      if wa != 0 and wb != 0 and (wa == pb or wb == pa):
        continue
      if not ((mjm.geom_contype[a] & mjm.geom_conaffinity[b]) or (mjm.geom_contype[b] & mjm.geom_conaffinity[a])):
        continue
      ra, rb = mjm.geom_rbound[a], mjm.geom_rbound[b]
      ma, mb = mjm.geom_margin[a], mjm.geom_margin[b]
      if ra == 0 and rb == 0:
        continue
      if ra == 0 or rb == 0:
        p, s = (a, b) if ra == 0 else (b, a)
        n = gm[p][:, 2]
        hit = np.dot(gx[s] - gx[p], n) <= mjm.geom_rbound[s] + ma + mb
      else:
        rs = ra + rb + ((ma + mb) if sap_bounds else max(ma, mb))
        hit = np.sum((gx[a] - gx[b]) ** 2) <= rs * rs
      if hit:
        out.append((a, b))
  return sorted(out)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_sap_and_nxn_match_brute_force(seed):
  rng = np.random.default_rng(seed)
  n = 40
  bodies = "".join(
    f'<body pos="{x:.3f} {y:.3f} {z:.3f}"><freejoint/><geom type="{"sphere" if i % 2 else "capsule"}" '
    f'size="{r:.3f} {h:.3f}" margin="{mg:.3f}"/></body>'
    for i, (x, y, z, r, h, mg) in enumerate(zip(rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(0, 1.5, n),
                                                rng.uniform(0.05, 0.2, n), rng.uniform(0.05, 0.2, n),
                                                rng.uniform(0, 0.05, n))))
  xml = (f'<mujoco><option jacobian="dense"/><worldbody><geom type="plane" size="5 5 .1"/>{bodies}'
         '</worldbody></mujoco>')
  mjm, m, d, hm, hd = helpers.load(xml, nconmax=4096)
  oracle.run("kinematics", hm, hd)
  oracle.run("sap_broadphase", hm, hd)
  assert helpers.pairs_of(hd) == _brute_pairs(mjm, hd, 0, sap_bounds=True)
  oracle.run("nxn_broadphase", hm, hd)
  assert helpers.pairs_of(hd) == _brute_pairs(mjm, hd, 0, sap_bounds=False)


# ---------------------------------------------------------------------------
# forward_test.py:130-145: Euler with implicit damping disabled
def test_euler_without_damping_closed_form():
  mjm = mjcf.from_xml_path(str(helpers_path("pendula.xml")))
  mjm.opt.disableflags |= 1 << 14
  mjm, m, d, hm, hd = helpers.load(mjm)
  hd.arrays["qvel"][:] = 1.0
  hd.arrays["qacc"][:] = 1.0
  oracle.run("euler", hm, hd)
  np.testing.assert_allclose(hd["qvel"][0], 1 + mjm.opt.timestep)


def helpers_path(name):
  from pathlib import Path
  return Path(__file__).resolve().parent / "golden" / name


# ---------------------------------------------------------------------------
# smooth dynamics against independent formulas
def _humanoid_state(seed=0):
  rng = np.random.default_rng(seed)
  mjm = models.humanoid()
  q = mjm.key_qpos[1].copy()
  q[7:] += rng.normal(0, 0.1, mjm.nq - 7)
  return mjm, q


def test_mass_matrix_matches_jacobian_sum():
  mjm, q = _humanoid_state()
  mjm, m, d, hm, hd = helpers.load(mjm)
  hd.arrays["qpos"][0] = q
  for s in ("kinematics", "com_pos", "crb", "factor_m"):
    oracle.run(s, hm, hd)
  M = hd["qM"][0]
  np.testing.assert_allclose(M, mjcf.mass_matrix(mjm, q), atol=1e-10, rtol=1e-9)
  # factor: the unique lower Cholesky factor
  np.testing.assert_allclose(hd["qLD"][0], np.linalg.cholesky(M), atol=1e-10)


def test_solve_m_and_mul_m():
  mjm, q = _humanoid_state(1)
  mjm, m, d, hm, hd = helpers.load(mjm)
  hd.arrays["qpos"][0] = q
  for s in ("kinematics", "com_pos", "crb", "factor_m", "fwd_velocity"):
    oracle.run(s, hm, hd)
  oracle.run("fwd_acceleration", hm, hd)
  M = hd["qM"][0]
  np.testing.assert_allclose(hd["qacc_smooth"][0], np.linalg.solve(M, hd["qfrc_smooth"][0]), rtol=1e-9, atol=1e-9)


def _bias(mjm, q, v, gravity=True):
  mjm = mjm.copy()
  if not gravity:
    mjm.opt.disableflags |= 1 << 6
  mjm, m, d, hm, hd = helpers.load(mjm)
  hd.arrays["qpos"][0] = q
  hd.arrays["qvel"][0] = v
  for s in ("kinematics", "com_pos", "crb", "fwd_velocity"):
    oracle.run(s, hm, hd)
  return hd["qfrc_bias"][0].copy()


def test_bias_is_gravity_at_rest():
  mjm, q = _humanoid_state(2)
  b = _bias(mjm, q, np.zeros(mjm.nv))
  _, _, xipos, sc, cdof = mjcf._fk_com(mjm, q)
  g = np.zeros(mjm.nv)
  for body in range(1, mjm.nbody):
    jp, _ = mjcf._body_jacobian(mjm, xipos[body], body, sc, cdof)
    g += mjm.body_mass[body] * jp.T @ -mjm.opt.gravity
  np.testing.assert_allclose(b, g, atol=1e-8)


def test_coriolis_matches_lagrangian_finite_difference():
  # hinge chain (no quaternion coordinates): c(q, v) = Mdot v - 1/2 d/dq (v' M v)
  mjm = models.chain(3)
  rng = np.random.default_rng(3)
  q = rng.uniform(-0.5, 0.5, 3)
  v = rng.normal(0, 1.0, 3)
  c = _bias(mjm, q, v, gravity=False)
  eps = 1e-6
  Mdot = (mjcf.mass_matrix(mjm, q + eps * v) - mjcf.mass_matrix(mjm, q - eps * v)) / (2 * eps)
  dT = np.zeros(3)
  for i in range(3):
    e = np.zeros(3)
    e[i] = eps
    dT[i] = 0.5 * (v @ mjcf.mass_matrix(mjm, q + e) @ v - v @ mjcf.mass_matrix(mjm, q - e) @ v) / (2 * eps)
  np.testing.assert_allclose(c, Mdot @ v - dT, atol=1e-6)


# ---------------------------------------------------------------------------
# contacts and constraint rows against closed forms
def test_plane_sphere_contact_closed_form():
  xml = """<mujoco><worldbody><geom type="plane" size="5 5 .1"/>
    <body pos="0.3 -0.2 0.08"><freejoint/><geom type="sphere" size="0.1"/></body></worldbody></mujoco>"""
  mjm, m, d, hm, hd = helpers.load(xml)
  oracle.run("kinematics", hm, hd)
  oracle.run("collision", hm, hd)
  assert int(hd["ncon"][0]) == 1
  np.testing.assert_allclose(hd["contact_dist"][0], 0.08 - 0.1)
  np.testing.assert_allclose(hd["contact_pos"][0], [0.3, -0.2, 0.08 - 0.1 - 0.5 * (0.08 - 0.1)])
  np.testing.assert_allclose(hd["contact_frame"][0].reshape(-1)[:3], [0, 0, 1])
  np.testing.assert_allclose(hd["contact_friction"][0], [1, 1, 0.005, 0.0001, 0.0001])


def test_limit_row_closed_form():
  mjm = models.chain(3)
  mjm, m, d, hm, hd = helpers.load(mjm)
  hd.arrays["qpos"][0] = [1.7, 0.0, -1.65]
  hd.arrays["qvel"][0] = [0.3, 0.0, -0.2]
  oracle.run("fwd_position", hm, hd)
  rows = [r for r in range(int(hd["nefc"][0])) if np.count_nonzero(hd["efc_J"][r]) == 1]
  assert len(rows) == 2
  dt = mjm.opt.timestep
  for r, (j, jsign) in zip(rows, [(0, -1.0), (2, 1.0)]):
    q, qd = hd["qpos"][0][j], hd["qvel"][0][j]
    pos = min(q + 1.57, 1.57 - q)
    assert hd["efc_J"][r][j] == jsign
    np.testing.assert_allclose(hd["efc_pos"][r], pos)
    # solref (0.02, 1), solimp (0.9, 0.95, 0.001, 0.5, 2): |pos| > width -> imp = dmax
    tc = max(0.02, 2 * dt)
    imp = 0.95
    k = 1 / (imp * imp * tc * tc)
    b = 2 / (imp * tc)
    np.testing.assert_allclose(hd["efc_aref"][r], -k * imp * pos - b * jsign * qd, rtol=1e-12)
    np.testing.assert_allclose(hd["efc_D"][r], 1 / (mjm.dof_invweight0[j] * (1 - imp) / imp), rtol=1e-12)


# ---------------------------------------------------------------------------
# solver: the converged iterate is the unique minimiser of the primal cost
@pytest.mark.parametrize("solver", [1, 2])
def test_solver_reaches_primal_minimum(solver):
  mjm = models.humanoid(iterations=100, ls_iterations=50, solver=solver)
  mjm.opt.tolerance = 1e-12
  mjm, m, d, hm, hd = helpers.load(mjm, njmax=512)
  q = mjm.key_qpos[0].copy()
  q[2] -= 0.01  # press the feet into the floor
  hd.arrays["qpos"][0] = q
  hd.arrays["qvel"][0] = np.random.default_rng(4).normal(0, 0.3, mjm.nv)
  oracle.run("forward", hm, hd)
  n = int(hd["nefc"][0])
  assert n > 0
  M = hd["qM"][0]
  a0 = hd["qacc_smooth"][0]
  J, D, aref = hd["efc_J"][:n], hd["efc_D"][:n], hd["efc_aref"][:n]

  def cost(a):
    jar = J @ a - aref
    act = jar < 0
    return 0.5 * (a - a0) @ M @ (a - a0) + 0.5 * np.sum(D[act] * jar[act] ** 2)

  def grad(a):
    jar = J @ a - aref
    act = jar < 0
    return M @ (a - a0) + J[act].T @ (D[act] * jar[act])

  res = scipy.optimize.minimize(cost, a0, jac=grad, method="L-BFGS-B", options=dict(maxiter=20000, gtol=1e-12,
                                                                                      ftol=1e-15))
  qacc = hd["qacc"][0]
  assert cost(qacc) <= res.fun * (1 + 1e-9) + 1e-9
  np.testing.assert_allclose(qacc, res.x, rtol=1e-4, atol=1e-4 * (1 + np.abs(res.x).max()))
  # efc_force is the constraint force of that minimiser
  jar = J @ qacc - aref
  np.testing.assert_allclose(hd["efc_force"][:n], np.where(jar < 0, -D * jar, 0.0), rtol=1e-9, atol=1e-9)
  np.testing.assert_allclose(hd["qfrc_constraint"][0], J.T @ hd["efc_force"][:n], atol=1e-8)


def test_free_fall_closed_form():
  xml = """<mujoco><option timestep="0.01"/><worldbody>
    <body pos="0 0 10"><freejoint/><geom type="sphere" size="0.1"/></body></worldbody></mujoco>"""
  mjm, m, d, hm, hd = helpers.load(xml)
  mjm.opt.disableflags |= 1 << 14
  n = 50
  for _ in range(n):
    oracle.run("step", hm, hd)
  g, dt = 9.81, 0.01
  np.testing.assert_allclose(hd["qpos"][0][2], 10 - g * dt * dt * n * (n + 1) / 2, rtol=1e-12)
  np.testing.assert_allclose(hd["qvel"][0][2], -g * dt * n, rtol=1e-12)


def test_cg_and_newton_converge_to_the_same_acceleration():
  out = []
  for solver in (1, 2):
    mjm = models.chain(3)
    mjm.opt.solver = solver
    mjm.opt.iterations = 200
    mjm.opt.tolerance = 1e-14
    mjm, m, d, hm, hd = helpers.load(mjm)
    hd.arrays["qpos"][0] = [0.2, -0.3, 0.1]
    hd.arrays["qvel"][0] = [2.0, 0.5, -1.0]
    hd.arrays["qpos"][0][0] = 1.62  # beyond the limit
    oracle.run("forward", hm, hd)
    out.append(hd["qacc"][0].copy())
  np.testing.assert_allclose(out[0], out[1], rtol=1e-6, atol=1e-6)
