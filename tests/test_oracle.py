"""Pins the oracle (the CPU restatement of the reference step) against the
reference's own known answers and against what the mathematics fixes.

Reference-held answers: math_test.py:25-88 (closest segment points),
broad_phase_test.py:45-242 (SAP and nxn pair sets), forward_test.py:130-145
(Euler without implicit damping).  Mathematical checks: the mass matrix
against a Jacobian-sum formula, the Cholesky factor and solve against
numpy, bias forces against gravity and finite-difference Coriolis terms,
contacts against closed forms, the converged solver against the unique
minimum of the convex primal cost, free fall against the discrete closed
form, the SAP pair set against brute force, joint motors, springs and
dampers against their physical definitions, implicit damping against the
backward-Euler closed form, and pyramidal contact rows against contact-point
velocities.
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


# ---------------------------------------------------------------------------
# actuation, passive forces and integrators against physical definitions
def test_joint_motor_transmission_and_force():
  """A motor on a hinge with gear g: length g q, moment g on the joint's dof,
  velocity = d(length)/dt, torque g ctrl (ctrl clamped to ctrlrange)."""
  mjm, q = _humanoid_state(4)
  rng = np.random.default_rng(4)
  v = rng.normal(0, 1.0, mjm.nv)
  u = rng.uniform(-1.5, 1.5, mjm.nu)
  mjm, m, d, hm, hd = helpers.load(mjm)
  hd.arrays["qpos"][0], hd.arrays["qvel"][0], hd.arrays["ctrl"][0] = q, v, u
  for s in ("kinematics", "com_pos", "crb", "transmission", "fwd_velocity", "fwd_actuation"):
    oracle.run(s, hm, hd)
  length = hd["actuator_length"][0]
  moment = hd["actuator_moment"][0].reshape(mjm.nu, mjm.nv)
  expect_f = np.zeros(mjm.nv)
  for a in range(mjm.nu):
    j = mjm.actuator_trnid[a, 0]
    g = mjm.actuator_gear[a, 0]
    dof, qa = mjm.jnt_dofadr[j], mjm.jnt_qposadr[j]
    np.testing.assert_allclose(length[a], g * q[qa], rtol=1e-12)
    row = np.zeros(mjm.nv)
    row[dof] = g
    np.testing.assert_array_equal(moment[a], row)
    expect_f[dof] += g * np.clip(u[a], *mjm.actuator_ctrlrange[a])
  # actuator velocity is the time derivative of the length
  eps = 1e-7
  mjm2, _, _, hm2, hd2 = helpers.load(mjm)
  hd2.arrays["qpos"][0] = q
  hd2.arrays["qpos"][0][7:] += eps * v[6:]
  for s in ("kinematics", "com_pos", "transmission"):
    oracle.run(s, hm2, hd2)
  np.testing.assert_allclose(hd["actuator_velocity"][0], (hd2["actuator_length"][0] - length) / eps, rtol=1e-5, atol=1e-5)
  np.testing.assert_allclose(hd["qfrc_actuator"][0], expect_f, rtol=1e-12, atol=1e-12)


PASSIVE_XML = """<mujoco><compiler angle="radian"/><option gravity="0 0 0"/><worldbody>
  <body><joint name="h" type="hinge" axis="0 1 0" stiffness="3" springref="0.2" damping="0.5"/>
    <geom type="sphere" size="0.1" mass="1"/>
    <body pos="0 0 -0.3"><joint name="b" type="ball" stiffness="2" damping="0.7"/>
      <geom type="sphere" size="0.1" mass="1"/></body></body></worldbody></mujoco>"""


def test_spring_and_damper_forces():
  """Hinge: -k (q - springref) and -b qdot.  Ball: the spring torque is
  -k theta axis for a rotation of theta about axis (rest = identity)."""
  mjm, m, d, hm, hd = helpers.load(PASSIVE_XML)
  axis = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])
  theta = 0.7
  qball = np.concatenate([[np.cos(theta / 2)], np.sin(theta / 2) * axis])
  hd.arrays["qpos"][0] = np.concatenate([[0.9], qball])
  v = np.array([1.3, 0.2, -0.4, 0.5])
  hd.arrays["qvel"][0] = v
  oracle.run("passive", hm, hd)
  np.testing.assert_allclose(hd["qfrc_spring"][0], np.concatenate([[-3 * (0.9 - 0.2)], -2 * theta * axis]), atol=1e-12)
  np.testing.assert_allclose(hd["qfrc_damper"][0], -np.array([0.5, 0.7, 0.7, 0.7]) * v, atol=1e-12)
  np.testing.assert_allclose(hd["qfrc_passive"][0], hd["qfrc_spring"][0] + hd["qfrc_damper"][0], atol=1e-12)


@pytest.mark.parametrize("integrator", ["Euler", "implicitfast"])
def test_damped_slide_is_backward_euler_in_the_damping(integrator):
  """A slide joint of mass m and damping b, no other force: both Euler (with
  implicit damping) and implicitfast step qdot' = qdot m / (m + dt b)."""
  xml = f"""<mujoco><option gravity="0 0 0" timestep="0.01" integrator="{integrator}"/><worldbody>
    <body><joint type="slide" axis="1 0 0" damping="3"/><geom type="sphere" size="0.1" mass="2"/></body>
    </worldbody></mujoco>"""
  mjm, m, d, hm, hd = helpers.load(xml)
  hd.arrays["qpos"][0] = [0.25]
  hd.arrays["qvel"][0] = [1.5]
  oracle.run("step", hm, hd)
  dt, mass, b = 0.01, 2.0, 3.0
  v1 = 1.5 * mass / (mass + dt * b)
  np.testing.assert_allclose(hd["qvel"][0], [v1], rtol=1e-12)
  np.testing.assert_allclose(hd["qpos"][0], [0.25 + dt * v1], rtol=1e-12)


@pytest.mark.parametrize("integrator", ["Euler", "implicitfast"])
def test_velocity_actuator_is_implicit_in_implicitfast_only(integrator):
  """A slide joint of mass m and damping b driven by a velocity actuator,
  force kv (u - qdot): affine in qdot, so implicitfast moves kv into the
  integration matrix (forward.py:245-405) and steps qdot' = qdot + dt (kv
  (u - qdot) - b qdot) / (m + dt (kv + b)); Euler's implicit term holds the
  damping only (denominator m + dt b)."""
  xml = f"""<mujoco><option gravity="0 0 0" timestep="0.01" integrator="{integrator}"/><worldbody>
    <body><joint name="j" type="slide" axis="1 0 0" damping="0.5"/><geom type="sphere" size="0.1" mass="2"/>
    </body></worldbody><actuator><velocity joint="j" kv="4"/></actuator></mujoco>"""
  mjm, m, d, hm, hd = helpers.load(xml)
  assert int(m.actuator_affine_bias_gain) == 1
  hd.arrays["qpos"][0] = [0.25]
  hd.arrays["qvel"][0] = [0.7]
  hd.arrays["ctrl"][0] = [1.5]
  oracle.run("step", hm, hd)
  dt, mass, b, kv, u, v0 = 0.01, 2.0, 0.5, 4.0, 1.5, 0.7
  denom = mass + dt * (kv + b if integrator == "implicitfast" else b)
  v1 = v0 + dt * (kv * (u - v0) - b * v0) / denom
  np.testing.assert_allclose(hd["qvel"][0], [v1], rtol=1e-12)
  np.testing.assert_allclose(hd["qpos"][0], [0.25 + dt * v1], rtol=1e-12)


def test_resting_sphere_penetration_closed_form():
  """A free sphere (mass m) settled on a plane with pyramidal condim-3
  contact: at rest the four rows carry m g / 4 each, row force D aref with
  aref = -k imp pos (constraint.py:22-73), D = imp / ((1 - imp) invweight),
  invweight = 2 mu^2 (1 + mu^2) body_invweight0 / impratio and
  body_invweight0 = 1 / m, k dmax^2 = 1 / timeconst^2 (dampratio 1), and imp
  = dmax beyond the impedance width.  So the depth is g (1 - dmax)
  timeconst^2 -- independent of m, which the test varies."""
  depths = []
  for mass in (0.5, 2.0):
    xml = f"""<mujoco><option timestep="0.002" iterations="100" tolerance="1e-12"/><worldbody>
      <geom type="plane" size="5 5 .1" solimp="0.9 0.95 0.00001 0.5 2"/>
      <body pos="0 0 0.1"><freejoint/><geom type="sphere" size="0.1" mass="{mass}"
        solimp="0.9 0.95 0.00001 0.5 2"/></body></worldbody></mujoco>"""
    mjm, m, d, hm, hd = helpers.load(xml, nconmax=8, njmax=16)
    np.testing.assert_allclose(np.asarray(hm.arrays["body_invweight0"])[1, 0], 1.0 / mass, rtol=1e-9)
    for _ in range(600):
      oracle.run("step", hm, hd)
    # settled vertically (the pyramid's friction rows leave a slow ~1e-8
    # rad/s spin, which does not load the normal direction)
    assert abs(hd["qvel"][0][2]) < 1e-9 and np.max(np.abs(hd["qvel"][0])) < 1e-6
    depths.append(0.1 - hd["qpos"][0][2])
  expected = 9.81 * (1.0 - 0.95) * 0.02 ** 2
  np.testing.assert_allclose(depths, [expected, expected], rtol=1e-6)


def test_resting_sphere_penetration_on_the_impedance_curve():
  """As above with the impedance width (0.01) above the resting depth: imp
  follows MuJoCo's solimp sigmoid, imp(x) = dmin + y(x) (dmax - dmin),
  x = depth / width, y = x^p / mid^(p-1) below mid, else 1 - (1-x)^p /
  (1-mid)^(p-1); the steady depth is the root of depth = g (1 - imp) /
  (k imp^2) with k = 1 / (dmax^2 timeconst^2)."""
  dmin, dmax, width, mid, power, tc, g = 0.9, 0.95, 0.01, 0.3, 3.0, 0.02, 9.81
  solimp = f"{dmin} {dmax} {width} {mid} {power}"
  xml = f"""<mujoco><option timestep="0.002" iterations="100" tolerance="1e-12"/><worldbody>
    <geom type="plane" size="5 5 .1" solimp="{solimp}"/>
    <body pos="0 0 0.1"><freejoint/><geom type="sphere" size="0.1" mass="1" solimp="{solimp}"/></body>
    </worldbody></mujoco>"""
  mjm, m, d, hm, hd = helpers.load(xml, nconmax=8, njmax=16)
  for _ in range(800):
    oracle.run("step", hm, hd)
  assert abs(hd["qvel"][0][2]) < 1e-9
  depth = 0.1 - hd["qpos"][0][2]

  def imp(x):
    y = x ** power / mid ** (power - 1) if x < mid else 1.0 - (1.0 - x) ** power / (1.0 - mid) ** (power - 1)
    return dmin + y * (dmax - dmin)

  k = 1.0 / (dmax * dmax * tc * tc)
  expected = scipy.optimize.brentq(lambda z: z - g * (1.0 - imp(z / width)) / (k * imp(z / width) ** 2),
                                   1e-9, width)
  assert expected < width and imp(expected / width) < dmax - 1e-3  # on the curve, not at dmax
  np.testing.assert_allclose(depth, expected, rtol=1e-6)


def test_set_const_closed_forms():
  """mj_setConst quantities of a free sphere (mass m, inertia I = 2/5 m r^2
  about every axis) and of a hinge pendulum: meaninertia is the mean of
  diag(M) = (3m + 3I) / 6; dof_invweight0 is 1/m on the translational and
  1/I on the rotational dofs; body_invweight0 = (1/m, 1/I); the pendulum's
  hinge has dof_invweight0 = 1 / (I_axis + m l^2)."""
  mass, r = 2.5, 0.2
  inertia = 0.4 * mass * r * r
  mjm = mjcf.from_xml_string(f"""<mujoco><worldbody><body><freejoint/>
    <geom type="sphere" size="{r}" mass="{mass}"/></body></worldbody></mujoco>""")
  np.testing.assert_allclose(mjm.stat.meaninertia, (3 * mass + 3 * inertia) / 6, rtol=1e-9)
  np.testing.assert_allclose(mjm.dof_invweight0, [1 / mass] * 3 + [1 / inertia] * 3, rtol=1e-9)
  np.testing.assert_allclose(mjm.body_invweight0[1], [1 / mass, 1 / inertia], rtol=1e-9)
  # pendulum: a point-like sphere at distance l from a hinge about y
  l, ms, rs = 0.7, 1.5, 0.05
  mjm = mjcf.from_xml_string(f"""<mujoco><worldbody><body><joint type="hinge" axis="0 1 0"/>
    <geom type="sphere" pos="{l} 0 0" size="{rs}" mass="{ms}"/></body></worldbody></mujoco>""")
  iaxis = 0.4 * ms * rs * rs + ms * l * l
  np.testing.assert_allclose(mjm.dof_invweight0, [1 / iaxis], rtol=1e-9)
  np.testing.assert_allclose(mjm.stat.meaninertia, iaxis, rtol=1e-9)


# ---------------------------------------------------------------------------
# primitive contacts: distance, position midway between the surfaces, normal
def _contacts(xml, qpos=None):
  mjm, m, d, hm, hd = helpers.load(xml)
  if qpos is not None:
    hd.arrays["qpos"][0] = qpos
  oracle.run("kinematics", hm, hd)
  oracle.run("collision", hm, hd)
  n = int(hd["ncon"][0])
  return (hd["contact_dist"][:n], hd["contact_pos"][:n], hd["contact_frame"][:n].reshape(n, 3, 3),
          hd["contact_geom"][:n])


def test_sphere_sphere_contact_closed_form():
  c1, c2, r1, r2 = np.array([0.1, 0.0, 1.0]), np.array([0.25, 0.1, 1.05]), 0.1, 0.15
  xml = f"""<mujoco><option gravity="0 0 0"/><worldbody>
    <body pos="{c1[0]} {c1[1]} {c1[2]}"><freejoint/><geom type="sphere" size="{r1}"/></body>
    <body pos="{c2[0]} {c2[1]} {c2[2]}"><freejoint/><geom type="sphere" size="{r2}"/></body></worldbody></mujoco>"""
  dist, pos, frame, geom = _contacts(xml)
  assert len(dist) == 1 and list(geom[0]) == [0, 1]
  n = (c2 - c1) / np.linalg.norm(c2 - c1)
  dd = np.linalg.norm(c2 - c1) - r1 - r2
  np.testing.assert_allclose(dist[0], dd, atol=1e-12)
  np.testing.assert_allclose(frame[0][0], n, atol=1e-12)
  np.testing.assert_allclose(pos[0], c1 + n * (r1 + 0.5 * dd), atol=1e-12)
  np.testing.assert_allclose(frame[0] @ frame[0].T, np.eye(3), atol=1e-12)


def test_plane_capsule_contacts_at_both_ends():
  h, half, r = 0.08, 0.2, 0.1
  xml = f"""<mujoco><worldbody><geom type="plane" size="5 5 .1"/>
    <body pos="0.1 0.2 {h}"><freejoint/><geom type="capsule" fromto="{-half} 0 0 {half} 0 0" size="{r}"/></body>
    </worldbody></mujoco>"""
  dist, pos, frame, geom = _contacts(xml)
  assert len(dist) == 2
  np.testing.assert_allclose(dist, [h - r, h - r], atol=1e-12)
  ends = sorted(pos[:, 0])
  np.testing.assert_allclose(ends, [0.1 - half, 0.1 + half], atol=1e-12)
  np.testing.assert_allclose(pos[:, 2], h - r - 0.5 * (h - r), atol=1e-12)
  for f in frame:
    np.testing.assert_allclose(f[0], [0, 0, 1], atol=1e-12)
    np.testing.assert_allclose(abs(f[1] @ [1, 0, 0]), 1.0, atol=1e-12)  # tangent along the capsule axis


def test_plane_box_corner_contacts():
  h, sx, sy, sz = 0.18, 0.3, 0.2, 0.2
  xml = f"""<mujoco><worldbody><geom type="plane" size="5 5 .1"/>
    <body pos="0 0 {h}"><freejoint/><geom type="box" size="{sx} {sy} {sz}"/></body></worldbody></mujoco>"""
  dist, pos, frame, geom = _contacts(xml)
  assert len(dist) == 4
  np.testing.assert_allclose(dist, h - sz, atol=1e-12)
  corners = sorted((round(p[0], 9), round(p[1], 9)) for p in pos)
  assert corners == sorted((x, y) for x in (-sx, sx) for y in (-sy, sy))
  np.testing.assert_allclose(pos[:, 2], (h - sz) - 0.5 * (h - sz), atol=1e-12)


def test_contact_rows_are_contact_point_velocities():
  """Pyramidal rows of a sliding, spinning sphere on a plane: the two rows of
  each friction direction average to the normal row, J_n qdot is the
  approach speed d(dist)/dt, and the half-difference over mu is the contact
  point's velocity along that tangent of the frame."""
  r, z = 0.1, 0.095
  xml = f"""<mujoco><option gravity="0 0 0"/><worldbody><geom type="plane" size="5 5 .1"/>
    <body pos="0 0 {z}"><freejoint/><geom type="sphere" size="{r}" friction="0.7 0.01 0.001"/></body>
    </worldbody></mujoco>"""
  mjm, m, d, hm, hd = helpers.load(xml)
  v = np.array([0.4, -0.3, 0.2, 1.1, -0.6, 0.9])
  hd.arrays["qvel"][0] = v
  oracle.run("fwd_position", hm, hd)
  assert int(hd["ncon"][0]) == 1 and int(hd["nefc"][0]) == 4
  J = hd["efc_J"][:4, :mjm.nv]
  frame = hd["contact_frame"][0].reshape(3, 3)
  # the element-wise max of the two geoms' friction: the plane's default 1
  np.testing.assert_allclose(hd["contact_friction"][0][:2], [1.0, 1.0])
  mu = 1.0
  p = hd["contact_pos"][0]
  c = hd["xpos"][0][1]
  R = hd["xmat"][0][1].reshape(3, 3)
  # free joint: qvel = (world linear velocity of the body origin, local angular velocity)
  w_world = R @ v[3:]
  vp = v[:3] + np.cross(w_world, p - c)
  Jn = 0.5 * (J[0] + J[1])
  np.testing.assert_allclose(0.5 * (J[2] + J[3]), Jn, atol=1e-12)
  np.testing.assert_allclose(Jn @ v, frame[0] @ vp, atol=1e-12)
  np.testing.assert_allclose(Jn @ v, v[2], atol=1e-12)  # d(dist)/dt for a sphere on a plane
  for k, rows in enumerate(((0, 1), (2, 3))):
    jt = (J[rows[0]] - J[rows[1]]) / (2 * mu)
    np.testing.assert_allclose(jt @ v, frame[1 + k] @ vp, atol=1e-12)


# ---------------------------------------------------------------------------
# mj_setConst of an articulated body: a planar two-link arm
def test_set_const_planar_two_link_arm():
  """Two hinges about y; link i a sphere of mass m_i, radius r_i at c_i
  along x from its hinge; the second hinge l1 along x from the first.  At
  qpos0 (straight arm) the textbook double-pendulum inertia is
    M11 = I1 + m1 c1^2 + I2 + m2 (l1 + c2)^2,
    M12 = I2 + m2 c2 (l1 + c2),  M22 = I2 + m2 c2^2,   I_i = 2/5 m_i r_i^2;
  dof_invweight0 = diag(M^-1), meaninertia = tr(M) / 2, and body
  invweights are the traces of J M^-1 J' over 3 at each body's COM: the
  COM of link 2 moves along z by (l1 + c2, c2) per unit hinge rate and
  spins about y by (1, 1); link 1 by (c1, 0) and (1, 0)."""
  m1, r1, c1, l1 = 1.3, 0.07, 0.25, 0.6
  m2, r2, c2 = 0.8, 0.05, 0.3
  mjm = mjcf.from_xml_string(f"""<mujoco><worldbody>
    <body pos="0 0 2"><joint type="hinge" axis="0 1 0"/>
      <geom type="sphere" pos="{c1} 0 0" size="{r1}" mass="{m1}"/>
      <body pos="{l1} 0 0"><joint type="hinge" axis="0 1 0"/>
        <geom type="sphere" pos="{c2} 0 0" size="{r2}" mass="{m2}"/></body></body>
    </worldbody></mujoco>""")
  I1, I2 = 0.4 * m1 * r1 * r1, 0.4 * m2 * r2 * r2
  M = np.array([[I1 + m1 * c1 ** 2 + I2 + m2 * (l1 + c2) ** 2, I2 + m2 * c2 * (l1 + c2)],
                [I2 + m2 * c2 * (l1 + c2), I2 + m2 * c2 ** 2]])
  Minv = np.linalg.inv(M)
  np.testing.assert_allclose(mjm.stat.meaninertia, np.trace(M) / 2, rtol=1e-9)
  np.testing.assert_allclose(mjm.dof_invweight0, np.diag(Minv), rtol=1e-9)
  jt2, jr2 = np.array([l1 + c2, c2]), np.array([1.0, 1.0])
  np.testing.assert_allclose(mjm.body_invweight0[2], [jt2 @ Minv @ jt2 / 3, jr2 @ Minv @ jr2 / 3], rtol=1e-9)
  np.testing.assert_allclose(mjm.body_invweight0[1], [c1 * c1 * Minv[0, 0] / 3, Minv[0, 0] / 3], rtol=1e-9)
  # and the oracle's mass matrix at qpos0 is the same matrix
  mjm, m, d, hm, hd = helpers.load(mjm)
  for s in ("kinematics", "com_pos", "crb"):
    oracle.run(s, hm, hd)
  np.testing.assert_allclose(hd["qM"][0], M, rtol=1e-9, atol=1e-12)


# ---------------------------------------------------------------------------
# forward kinematics in closed form
def test_hinge_chain_kinematics_closed_form():
  """3-link chain hung from (0, 0, 1.5), hinges about y, 0.5 m links along
  -z: with cumulative angles T_k = q_1 + ... + q_k, body k turns by R_y(T_k)
  (quaternion (cos T_k/2, 0, sin T_k/2, 0)), the next hinge sits 0.5 m along
  R_y(T_k)(0, 0, -1) = (-sin T_k, 0, -cos T_k), and each link's COM (and
  capsule centre) halfway down it."""
  mjm = models.chain(3)
  rng = np.random.default_rng(11)
  nworld = 8
  mjm, m, d, hm, hd = helpers.load(mjm, nworld=nworld)
  q = rng.uniform(-1.5, 1.5, (nworld, 3))
  hd.arrays["qpos"][:] = q
  oracle.run("kinematics", hm, hd)
  for w in range(nworld):
    T = np.cumsum(q[w])
    p = np.array([0.0, 0.0, 1.5])
    for k in range(3):
      b = k + 1
      down = np.array([-np.sin(T[k]), 0.0, -np.cos(T[k])])
      R = np.array([[np.cos(T[k]), 0, np.sin(T[k])], [0, 1, 0], [-np.sin(T[k]), 0, np.cos(T[k])]])
      np.testing.assert_allclose(hd["xpos"][w, b], p, atol=1e-12)
      quat = np.array([np.cos(T[k] / 2), 0.0, np.sin(T[k] / 2), 0.0])
      np.testing.assert_allclose(hd["xquat"][w, b] * np.sign(hd["xquat"][w, b, 0]), quat * np.sign(quat[0]),
                                 atol=1e-12)
      np.testing.assert_allclose(hd["xmat"][w, b].reshape(3, 3), R, atol=1e-12)
      np.testing.assert_allclose(hd["xanchor"][w, k], p, atol=1e-12)
      np.testing.assert_allclose(hd["xaxis"][w, k], [0, 1, 0], atol=1e-12)
      np.testing.assert_allclose(hd["xipos"][w, b], p + 0.25 * down, atol=1e-12)
      np.testing.assert_allclose(hd["geom_xpos"][w, b], p + 0.25 * down, atol=1e-12)
      p = p + 0.5 * down


def test_ball_and_free_joint_kinematics_closed_form():
  """A free body carrying a ball-jointed pendulum: the free body's pose is
  qpos[0:7]; the child (at offset o in the parent, its COM at c below the
  ball) has orientation q_free * q_ball and COM x + R(q_free) (o + R(q_ball) c),
  rotations from scipy (an independent quaternion convention check)."""
  from scipy.spatial.transform import Rotation

  o, c = np.array([0.1, -0.2, 0.05]), np.array([0.0, 0.0, -0.4])
  mjm = mjcf.from_xml_string(f"""<mujoco><worldbody>
    <body pos="0 0 1"><freejoint/><geom type="sphere" size="0.1"/>
      <body pos="{o[0]} {o[1]} {o[2]}"><joint type="ball"/>
        <geom type="sphere" pos="{c[0]} {c[1]} {c[2]}" size="0.05"/></body></body>
    </worldbody></mujoco>""")
  rng = np.random.default_rng(12)
  nworld = 6
  mjm, m, d, hm, hd = helpers.load(mjm, nworld=nworld)
  qf = rng.normal(size=(nworld, 4))
  qf /= np.linalg.norm(qf, axis=1, keepdims=True)
  qb = rng.normal(size=(nworld, 4))
  qb /= np.linalg.norm(qb, axis=1, keepdims=True)
  x = rng.normal(size=(nworld, 3))
  hd.arrays["qpos"][:] = np.concatenate([x, qf, qb], axis=1)
  oracle.run("kinematics", hm, hd)
  for w in range(nworld):
    Rf = Rotation.from_quat(np.roll(qf[w], -1))  # scipy takes (x, y, z, w)
    Rb = Rotation.from_quat(np.roll(qb[w], -1))
    np.testing.assert_allclose(hd["xpos"][w, 1], x[w], atol=1e-12)
    np.testing.assert_allclose(hd["xmat"][w, 1].reshape(3, 3), Rf.as_matrix(), atol=1e-12)
    child = Rf * Rb
    np.testing.assert_allclose(hd["xmat"][w, 2].reshape(3, 3), child.as_matrix(), atol=1e-12)
    np.testing.assert_allclose(hd["xpos"][w, 2], x[w] + Rf.apply(o), atol=1e-12)
    np.testing.assert_allclose(hd["xipos"][w, 2], x[w] + Rf.apply(o + Rb.apply(c)), atol=1e-12)
    np.testing.assert_allclose(hd["xanchor"][w, 1], x[w] + Rf.apply(o), atol=1e-12)


def test_actuator_forces_cover_every_tree():
  """Actuator forces reach every kinematic tree.  n_humanoids.xml is a
  sparse model (nv 81), for which the reference sums moment' force and
  moment * qvel over all actuators and dofs (forward.py:427-438, 571-578);
  its dense tile loop ends the last tile group at the number of groups
  (forward.py:629-633), so the 3 trees of one tile size would get only the
  first tree's forces.  Both quantities must equal the full products."""
  mjm = mjcf.from_xml_string((models.ASSETS / "n_humanoids.xml").read_text(), base_dir=str(models.ASSETS))
  _, m, d, hm, hd = helpers.load(mjm, nworld=2)
  assert m.opt.is_sparse
  rng = np.random.default_rng(3)
  hd["qvel"][:] = rng.normal(0, 0.5, hd["qvel"].shape)
  hd["ctrl"][:] = rng.uniform(-1, 1, hd["ctrl"].shape)
  oracle.run("forward", hm, hd)
  for w in range(2):
    mom = hd["actuator_moment"][w].reshape(mjm.nu, mjm.nv)
    np.testing.assert_allclose(hd["qfrc_actuator"][w], mom.T @ hd["actuator_force"][w], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(hd["actuator_velocity"][w], mom @ hd["qvel"][w], rtol=1e-12, atol=1e-12)
    # every humanoid's 21 motors move its own tree
    assert all(np.abs(hd["qfrc_actuator"][w][27 * k + 6:27 * (k + 1)]).max() > 0 for k in range(3))
