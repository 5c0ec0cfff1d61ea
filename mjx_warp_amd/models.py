"""Model builders for the benchmark and test workloads.

* `chain(n)`      — n-link capsule hinge chain above a plane (the CPU-runnable
                    plain-step workload: 3 links of 0.5 m x 0.05 m, 1 kg each,
                    hinges about y, limits +-1.57 rad, first hinge at z = 1.5 m
                    so the hanging chain's tip grazes the plane, Euler
                    dt = 0.002, CG 100 iterations, tol 1e-8).
* `humanoid()`    — the 27-dof humanoid (assets/humanoid.xml) with the
                    headline-benchmark overrides: all 21 motor gears 100,
                    implicitfast, Newton, iterations / ls_iterations 1 / 4
                    (the reference benchmark's defaults, testspeed.py:42-59).
* `humanoid_rollout_state(nworld)` — the differentiable-rollout workload's
                    initial states: the squat keyframe with the torso
                    position and the 21 hinge angles perturbed by N(0, 0.05^2),
                    the torso quaternion rotated by exp of an N(0, 0.05^2)
                    rotation vector (and normalised), qvel = 0 (numpy
                    default_rng(3)).
* `box_world(links, radii)` / `heterogeneous_groups(...)` — the
                    heterogeneous workload's models: a hinge chain of
                    `links` 0.1 m capsule links hung from (0, 0, 0.95) plus
                    free spheres of the given radii (density 1000) in a
                    1 m box (floor and four walls at x, y = +-0.5); one model
                    per (links, spheres) pair, worlds split over the pairs.
                    The reduced workload draws 2-8 links and 1-4 spheres
                    (nv <= 32); `heterogeneous_groups(max_links=32,
                    max_spheres=64)` is the workload at its stated size
                    (nv up to 416), with chains longer than 9 links folded
                    (`chain_fold`) so that they start inside the box.  The
                    chain-workload links (0.5 m) would not fit a 1 m box past
                    two links, so these chains take 0.1 m links.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from . import mjcf

ASSETS = Path(__file__).resolve().parent / "assets"


def chain_xml(n: int = 3, anchor_z: float = 1.5) -> str:
  link = ('<joint name="hinge{i}" type="hinge" axis="0 1 0" range="-1.57 1.57" limited="true"/>'
          '<geom name="link{i}" type="capsule" fromto="0 0 0 0 0 -0.5" size="0.05" mass="1"/>')
  body = ""
  for i in reversed(range(n)):
    pos = f"0 0 {anchor_z}" if i == 0 else "0 0 -0.5"
    body = f'<body name="body{i}" pos="{pos}">' + link.format(i=i) + body + "</body>"
  return f"""
<mujoco model="chain{n}">
  <compiler angle="radian"/>
  <option timestep="0.002" integrator="Euler" solver="CG" iterations="100" tolerance="1e-8"
          ls_iterations="50"/>
  <worldbody>
    <geom name="floor" type="plane" size="0 0 0.05"/>
    {body}
  </worldbody>
</mujoco>
"""


def chain(n: int = 3) -> mjcf.MjModel:
  return mjcf.from_xml_string(chain_xml(n))


def humanoid(gear: float = 100.0, iterations: int = 1, ls_iterations: int = 4,
             integrator: int = 3, solver: int = 2) -> mjcf.MjModel:
  m = mjcf.from_xml_path(str(ASSETS / "humanoid.xml"))
  m.actuator_gear[:, 0] = gear
  m.opt.integrator = integrator
  m.opt.solver = solver
  m.opt.iterations = iterations
  m.opt.ls_iterations = ls_iterations
  return m


def chain_initial_state(nworld: int, seed: int = 0, nq: int = 3):
  """qpos ~ U(-0.5, 0.5), qvel ~ N(0, 0.1) from numpy default_rng(seed)."""
  rng = np.random.default_rng(seed)
  qpos = rng.uniform(-0.5, 0.5, (nworld, nq))
  qvel = rng.normal(0.0, 0.1, (nworld, nq))
  return qpos, qvel


def humanoid_rollout_state(mjm, nworld: int, seed: int = 3):
  """(qpos0, qvel0) [nworld, nq] / [nworld, nv] of the differentiable
  rollout workload (see the module docstring)."""
  rng = np.random.default_rng(seed)
  q = np.tile(np.asarray(mjm.key_qpos[0], np.float64), (nworld, 1))
  q[:, 0:3] += rng.normal(0.0, 0.05, (nworld, 3))
  q[:, 7:] += rng.normal(0.0, 0.05, (nworld, mjm.nq - 7))
  v = rng.normal(0.0, 0.05, (nworld, 3))
  ang = np.linalg.norm(v, axis=1, keepdims=True)
  dq = np.concatenate([np.cos(0.5 * ang), v * (np.sin(0.5 * ang) / np.maximum(ang, 1e-300))], 1)
  a, b = q[:, 3:7], dq
  q[:, 3:7] = np.stack([a[:, 0] * b[:, 0] - a[:, 1] * b[:, 1] - a[:, 2] * b[:, 2] - a[:, 3] * b[:, 3],
                        a[:, 0] * b[:, 1] + a[:, 1] * b[:, 0] + a[:, 2] * b[:, 3] - a[:, 3] * b[:, 2],
                        a[:, 0] * b[:, 2] - a[:, 1] * b[:, 3] + a[:, 2] * b[:, 0] + a[:, 3] * b[:, 1],
                        a[:, 0] * b[:, 3] + a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1] + a[:, 3] * b[:, 0]], 1)
  q[:, 3:7] /= np.linalg.norm(q[:, 3:7], axis=1, keepdims=True)
  return q, np.zeros((nworld, mjm.nv))


def box_world_xml(links: int, radii) -> str:
  """A chain of `links` hinge links (0.1 m capsules of radius 0.02, 0.1 kg,
  hinges about y limited to +-1.57) hung from (0, 0, 0.95), and one free
  sphere per radius (density 1000), inside a 1 m box: a floor and four walls
  (planes at x, y = +-0.5 facing inward).  Newton, implicitfast,
  dt 0.002, dense mass matrix at every size (implicitfast takes no sparse
  one)."""
  body = ""
  for i in reversed(range(links)):
    pos = "0 0 0.95" if i == 0 else "0 0 -0.1"
    body = (f'<body name="link{i}" pos="{pos}"><joint type="hinge" axis="0 1 0" range="-1.57 1.57" '
            f'limited="true"/><geom type="capsule" fromto="0 0 0 0 0 -0.1" size="0.02" mass="0.1"/>'
            + body + "</body>")
  spheres = "".join(f'<body name="ball{k}" pos="0 0 0.5"><freejoint/><geom type="sphere" size="{r:.6g}" '
                    f'density="1000"/></body>' for k, r in enumerate(radii))
  return f"""
<mujoco model="box{links}x{len(radii)}">
  <compiler angle="radian"/>
  <option timestep="0.002" integrator="implicitfast" solver="Newton" iterations="20" ls_iterations="10"
          jacobian="dense"/>
  <worldbody>
    <geom name="floor" type="plane" size="0 0 0.05"/>
    <geom name="wall_px" type="plane" pos="0.5 0 0.5" zaxis="-1 0 0" size="0.5 0.5 0.05"/>
    <geom name="wall_nx" type="plane" pos="-0.5 0 0.5" zaxis="1 0 0" size="0.5 0.5 0.05"/>
    <geom name="wall_py" type="plane" pos="0 0.5 0.5" zaxis="0 -1 0" size="0.5 0.5 0.05"/>
    <geom name="wall_ny" type="plane" pos="0 -0.5 0.5" zaxis="0 1 0" size="0.5 0.5 0.05"/>
    {body}
    {spheres}
  </worldbody>
</mujoco>"""


def box_world(links: int, radii) -> mjcf.MjModel:
  return mjcf.from_xml_string(box_world_xml(links, radii))


def chain_fold(links: int) -> np.ndarray:
  """Hinge angles that keep a chain of `links` 0.1 m links inside the box:
  hanging straight up to 9 links; longer chains fold into a zig-zag in the
  x-z plane (runs of four links 1.45 rad from the vertical, alternating
  direction, joined by a vertical link; every hinge within its +-1.57
  limit), which stays within 0.4 m of x = 0 and drops about 0.3 m per ten
  links."""
  if links <= 9:
    return np.zeros(links)
  phi, sign = [], 1.0
  while len(phi) < links:
    phi += [sign * 1.45] * 4 + [0.0]
    sign = -sign
  phi = np.array(phi[:links])
  return np.diff(np.concatenate([[0.0], phi]))


def box_world_states(mjm, links: int, radii, nworld: int, rng):
  """(qpos, qvel) [nworld, nq] / [nworld, nv]: the chain in its chain_fold
  pose, each sphere at rest at a uniform position inside the box (clear of
  the walls, the floor, the other spheres and the chain -- the column
  |x|, |y| < 0.1 + r of a straight chain, the slab |y| < 0.05 + r,
  |x| < 0.45 + r of a folded one; rejection sampling), identity
  orientation."""
  q = np.zeros((nworld, mjm.nq))
  fold = chain_fold(links)
  q[:, :links] = fold
  folded = np.any(fold != 0.0)
  for w in range(nworld):
    placed = []
    for k, r in enumerate(radii):
      for _ in range(1000):
        p = rng.uniform([-0.5 + r, -0.5 + r, r], [0.5 - r, 0.5 - r, 0.9 - r])
        if not folded and abs(p[0]) < 0.1 + r and abs(p[1]) < 0.1 + r:
          continue
        if folded and abs(p[1]) < 0.05 + r and abs(p[0]) < 0.45 + r:
          continue
        if all(np.linalg.norm(p - c) > r + rc for c, rc in placed):
          break
      placed.append((p, r))
      a = links + 7 * k
      q[w, a:a + 3] = p
      q[w, a + 3] = 1.0
  return q, np.zeros((nworld, mjm.nv))


def _box_world_entry(args):
  return box_world(*args)


def box_worlds(specs, processes: int = 1) -> list:
  """box_world(links, radii) for every (links, radii, ...) of specs, in a
  pool of `processes` worker processes (spawned: they compile MJCF only)
  when there are many."""
  todo = [(links, radii) for links, radii, *_ in specs]
  if processes <= 1 or len(todo) < 2 * processes:
    return [box_world(*a) for a in todo]
  import concurrent.futures
  import multiprocessing

  ctx = multiprocessing.get_context("spawn")
  with concurrent.futures.ProcessPoolExecutor(processes, mp_context=ctx) as ex:
    return list(ex.map(_box_world_entry, todo, chunksize=max(1, len(todo) // (8 * processes))))


def heterogeneous_groups(nworld: int = 16384, seed: int = 2, max_links: int = 8, max_spheres: int = 4):
  """Worlds of the heterogeneous workload split by (links, spheres): links
  ~ U{2..max_links}, spheres ~ U{1..max_spheres} per world (numpy
  default_rng(seed)); each pair's sphere radii ~ U(0.03, 0.1), drawn once
  per pair.  Returns [(links, radii, nworld of the pair)] sorted by pair.
  With the defaults every model has nv = links + 6 spheres <= 32."""
  rng = np.random.default_rng(seed)
  L = rng.integers(2, max_links + 1, nworld)
  S = rng.integers(1, max_spheres + 1, nworld)
  out = []
  for links in range(2, max_links + 1):
    for s in range(1, max_spheres + 1):
      n = int(np.sum((L == links) & (S == s)))
      if n:
        out.append((links, tuple(float(r) for r in rng.uniform(0.03, 0.1, s)), n))
  return out


def contact_heavy_xml(nbody: int = 256, seed: int = 4) -> str:
  """The contact-heavy workload's model: `nbody` free bodies, alternately
  spheres (radius U(0.05, 0.1)) and capsules (radius 0.05, half-length
  U(0.05, 0.15)), density 1000, in a 2 m box of 5 planes (floor and four
  walls at x, y = +-1 facing inward); sizes drawn once (numpy
  default_rng(seed)).  Newton (50 iterations, tol 1e-8), implicitfast,
  dt 0.002, dense mass matrix.  261 geoms: sweep-and-prune broadphase."""
  rng = np.random.default_rng(seed)
  bodies = []
  for i in range(nbody):
    if i % 2 == 0:
      geom = f'<geom type="sphere" size="{rng.uniform(0.05, 0.1):.6g}" density="1000"/>'
    else:
      geom = f'<geom type="capsule" size="0.05 {rng.uniform(0.05, 0.15):.6g}" density="1000"/>'
    bodies.append(f'<body name="b{i}" pos="0 0 {1 + i}"><freejoint/>{geom}</body>')
  return f"""
<mujoco model="contact_heavy{nbody}">
  <option timestep="0.002" integrator="implicitfast" solver="Newton" iterations="50" tolerance="1e-8"
          jacobian="dense"/>
  <worldbody>
    <geom name="floor" type="plane" size="0 0 0.05"/>
    <geom name="wall_px" type="plane" pos="1 0 1" zaxis="-1 0 0" size="1 1 0.05"/>
    <geom name="wall_nx" type="plane" pos="-1 0 1" zaxis="1 0 0" size="1 1 0.05"/>
    <geom name="wall_py" type="plane" pos="0 1 1" zaxis="0 -1 0" size="1 1 0.05"/>
    <geom name="wall_ny" type="plane" pos="0 -1 1" zaxis="0 1 0" size="1 1 0.05"/>
    {"".join(bodies)}
  </worldbody>
</mujoco>"""


def contact_heavy(nbody: int = 256, seed: int = 4) -> mjcf.MjModel:
  return mjcf.from_xml_string(contact_heavy_xml(nbody, seed))


def contact_heavy_states(mjm, nworld: int, rng, height: float = 4.0):
  """qpos [nworld, nq]: every body at a uniform position inside the box
  (its bounding sphere clear of the walls, the floor and the bodies placed
  before it; rejection sampling up to `height`) with a uniform random
  orientation -- the bodies are dropped from rest."""
  nb = mjm.nbody - 1
  q = np.zeros((nworld, mjm.nq))
  rb = np.asarray(mjm.geom_rbound)[5:5 + nb]
  for w in range(nworld):
    placed = []
    for k in range(nb):
      r = rb[k]
      for _ in range(2000):
        p = rng.uniform([-1 + r, -1 + r, r], [1 - r, 1 - r, height - r])
        if all(np.linalg.norm(p - c) > r + rc for c, rc in placed):
          break
      placed.append((p, r))
      quat = rng.normal(size=4)
      q[w, 7 * k:7 * k + 3] = p
      q[w, 7 * k + 3:7 * k + 7] = quat / np.linalg.norm(quat)
  return q

