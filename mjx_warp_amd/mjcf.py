"""MJCF-subset compiler producing a host-side MuJoCo-style model.

`put_model` in the reference consumes a compiled `mujoco.MjModel`
(mujoco_warp/_src/io.py:24).  MuJoCo is not a dependency here, so this
module compiles the subset of MJCF the physics step supports into an
`MjModel` whose array names, shapes and meanings are MuJoCo's, including the
quantities MuJoCo derives at compile time:

* body inertial frames from geoms (mass, COM, principal axes),
* tree tables (body_rootid, body_weldid, dof_parentid, dof_Madr, tree ids),
* geom bounding radii,
* the `mj_setConst` statistics the constraint rows and the solver use:
  `stat.meaninertia` (mean diagonal of M at qpos0), `body_invweight0`
  (mean diagonal of the translational / rotational blocks of
  J_com M^-1 J_com' at qpos0) and `dof_invweight0` (diagonal of M^-1 at
  qpos0, averaged over the translational and rotational dofs of free joints
  and over ball joints).

Supported: <compiler angle autolimits eulerseq>, <option> with <flag>,
<default> classes (geom, joint, motor, position, general, site),
<worldbody> bodies with pos / quat / euler / axisangle / zaxis / xyaxes,
<inertial>, <joint> (free, ball, slide, hinge), <freejoint>, <geom>
(plane, sphere, capsule, cylinder, box, ellipsoid) with fromto / zaxis,
<site>, <actuator> (motor, position, general with joint / jointinparent),
<keyframe>.  Elements MuJoCo uses only for rendering are ignored.
"""

from __future__ import annotations

import copy
import math
import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

from . import types

# MuJoCo defaults (MuJoCo 3.2.7 documentation, "XML Reference")
_GEOM_DEFAULT = dict(
  type="sphere", contype=1, conaffinity=1, condim=3, group=0, priority=0,
  size=[0.0, 0.0, 0.0], friction=[1.0, 0.005, 0.0001], solmix=1.0,
  solref=[0.02, 1.0], solimp=[0.9, 0.95, 0.001, 0.5, 2.0], margin=0.0, gap=0.0,
  density=1000.0, mass=None, pos=[0.0, 0.0, 0.0],
)
_JOINT_DEFAULT = dict(
  type="hinge", pos=[0.0, 0.0, 0.0], axis=[0.0, 0.0, 1.0], range=[0.0, 0.0],
  limited="auto", stiffness=0.0, damping=0.0, armature=0.0, ref=0.0, springref=0.0,
  solreflimit=[0.02, 1.0], solimplimit=[0.9, 0.95, 0.001, 0.5, 2.0], margin=0.0,
  actuatorfrcrange=[0.0, 0.0], actuatorfrclimited="auto",
)
_ACT_DEFAULT = dict(
  gear=[1.0, 0.0, 0.0, 0.0, 0.0, 0.0], ctrlrange=[0.0, 0.0], ctrllimited="auto",
  forcerange=[0.0, 0.0], forcelimited="auto", actrange=[0.0, 0.0], actlimited="auto",
  kp=1.0, kv=0.0, gainprm=None, biasprm=None, dynprm=None, gaintype="fixed",
  biastype="none", dyntype="none",
)
_SITE_DEFAULT = dict(pos=[0.0, 0.0, 0.0], size=[0.005, 0.005, 0.005])

_GEOM_TYPES = {"plane": 0, "hfield": 1, "sphere": 2, "capsule": 3, "ellipsoid": 4,
               "cylinder": 5, "box": 6, "mesh": 7}
_JOINT_TYPES = {"free": 0, "ball": 1, "slide": 2, "hinge": 3}
_DISABLE_FLAGS = {
  "constraint": 1 << 0, "equality": 1 << 1, "frictionloss": 1 << 2, "limit": 1 << 3,
  "contact": 1 << 4, "passive": 1 << 5, "gravity": 1 << 6, "clampctrl": 1 << 7,
  "warmstart": 1 << 8, "filterparent": 1 << 9, "actuation": 1 << 10, "refsafe": 1 << 11,
  "sensor": 1 << 12, "midphase": 1 << 13, "eulerdamp": 1 << 14,
}
_INTEGRATORS = {"euler": 0, "rk4": 1, "implicit": 2, "implicitfast": 3}
_SOLVERS = {"pgs": 0, "cg": 1, "newton": 2}
_CONES = {"pyramidal": 0, "elliptic": 1}
_JACOBIANS = {"dense": 0, "sparse": 1, "auto": 2}

_MINVAL = 1e-15


# ---------------------------------------------------------------------------
# small float64 math (host side, compile time only)

def _qmul(a, b):
  return np.array([
    a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
    a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
    a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1],
    a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0],
  ])


def _qrot(q, v):
  u = np.asarray(q[1:4])
  s = q[0]
  return 2.0 * np.dot(u, v) * u + (s * s - np.dot(u, u)) * np.asarray(v) + 2.0 * s * np.cross(u, v)


def _q2m(q):
  w, x, y, z = q
  return np.array([
    [w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)],
    [2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
    [2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z],
  ])


def _m2q(mat):
  """Unit quaternion (w first, w >= 0) of a proper rotation matrix."""
  m = np.asarray(mat, dtype=np.float64)
  tr = m[0, 0] + m[1, 1] + m[2, 2]
  if tr > 0:
    s = math.sqrt(tr + 1.0) * 2
    q = [0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s]
  elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
    s = math.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
    q = [(m[2, 1] - m[1, 2]) / s, 0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s]
  elif m[1, 1] > m[2, 2]:
    s = math.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
    q = [(m[0, 2] - m[2, 0]) / s, (m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s]
  else:
    s = math.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
    q = [(m[1, 0] - m[0, 1]) / s, (m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s]
  q = np.array(q)
  q /= np.linalg.norm(q)
  if q[0] < 0:
    q = -q
  return q


def _axisangle(axis, angle):
  axis = np.asarray(axis, dtype=np.float64)
  n = np.linalg.norm(axis)
  if n < 1e-15:
    return np.array([1.0, 0.0, 0.0, 0.0])
  axis = axis / n
  return np.concatenate([[math.cos(angle / 2)], axis * math.sin(angle / 2)])


def _z2quat(vec):
  """Quaternion rotating (0, 0, 1) onto vec (minimal rotation)."""
  v = np.asarray(vec, dtype=np.float64)
  v = v / np.linalg.norm(v)
  axis = np.cross([0.0, 0.0, 1.0], v)
  s = np.linalg.norm(axis)
  if s < 1e-10:
    return np.array([1.0, 0.0, 0.0, 0.0]) if v[2] > 0 else np.array([0.0, 1.0, 0.0, 0.0])
  angle = math.atan2(s, v[2])
  return _axisangle(axis / s, angle)


# ---------------------------------------------------------------------------
# host model / data containers (MuJoCo field names)

@dataclass
class MjOption:
  timestep: float = 0.002
  impratio: float = 1.0
  tolerance: float = 1e-8
  ls_tolerance: float = 0.01
  gravity: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, -9.81]))
  integrator: int = 0
  cone: int = 0
  solver: int = 2
  iterations: int = 100
  ls_iterations: int = 50
  disableflags: int = 0
  jacobian: int = 2


@dataclass
class MjStatistic:
  meaninertia: float = 1.0


class MjModel:
  """Compiled model with MuJoCo's field names (numpy arrays, float64)."""

  def __init__(self):
    self.opt = MjOption()
    self.stat = MjStatistic()
    self.names = {}

  def name2id(self, kind: str, name: str) -> int:
    return self.names[kind].index(name)

  def copy(self) -> "MjModel":
    return copy.deepcopy(self)


class MjData:
  """Host state of one world with MuJoCo's field names.

  Only the fields `put_data` reads are allocated; derived fields are zero
  until a forward pass fills them.
  """

  def __init__(self, m: MjModel):
    self.time = 0.0
    self.qpos = m.qpos0.copy()
    self.qvel = np.zeros(m.nv)
    self.act = np.zeros(m.na)
    self.ctrl = np.zeros(m.nu)
    self.qacc = np.zeros(m.nv)
    self.qacc_warmstart = np.zeros(m.nv)
    self.qfrc_applied = np.zeros(m.nv)
    self.xfrc_applied = np.zeros((m.nbody, 6))
    self.mocap_pos = np.array([m.body_pos[b] for b in range(m.nbody) if m.body_mocapid[b] >= 0]).reshape(-1, 3)
    self.mocap_quat = np.array([m.body_quat[b] for b in range(m.nbody) if m.body_mocapid[b] >= 0]).reshape(-1, 4)
    self.ncon = 0
    self.nefc = 0
    self.nl = 0


def mj_resetDataKeyframe(m: MjModel, d: MjData, key: int):
  d.time = float(m.key_time[key])
  d.qpos[:] = m.key_qpos[key]
  d.qvel[:] = m.key_qvel[key]
  d.act[:] = m.key_act[key]
  d.ctrl[:] = m.key_ctrl[key]
  d.qacc_warmstart[:] = 0.0


# ---------------------------------------------------------------------------
# XML parsing

def _floats(s):
  return [float(x) for x in s.split()]


def _parse_value(key, val):
  if key in ("type", "name", "class", "childclass", "joint", "jointinparent", "limited",
             "ctrllimited", "forcelimited", "actlimited", "actuatorfrclimited", "gaintype",
             "biastype", "dyntype", "mocap", "material", "mesh", "align"):
    return val
  try:
    v = _floats(val)
  except ValueError:
    return val
  return v


def _attrs(el):
  return {k: _parse_value(k, v) for k, v in el.attrib.items()}


def _merge_array(default, given):
  """MJCF real arrays: given values override the leading entries."""
  out = list(default) if default is not None else []
  given = list(given)
  if len(given) >= len(out):
    return given
  out[: len(given)] = given
  return out


class _Defaults:
  def __init__(self):
    self.classes = {}

  def build(self, el, parent=None):
    name = el.get("class", "main") if parent is not None else "main"
    base = copy.deepcopy(self.classes[parent]) if parent is not None else {
      "geom": dict(_GEOM_DEFAULT), "joint": dict(_JOINT_DEFAULT), "actuator": dict(_ACT_DEFAULT),
      "site": dict(_SITE_DEFAULT),
    }
    for child in el:
      tag = child.tag
      if tag == "default":
        continue
      kind = {"geom": "geom", "joint": "joint", "site": "site", "motor": "actuator",
              "position": "actuator", "general": "actuator", "velocity": "actuator"}.get(tag)
      if kind is None:
        continue
      for k, v in _attrs(child).items():
        if isinstance(v, list) and isinstance(base[kind].get(k), list):
          base[kind][k] = _merge_array(base[kind][k], v)
        else:
          base[kind][k] = v
    self.classes[name] = base
    for child in el:
      if child.tag == "default":
        self.build(child, name)

  def get(self, kind, cls):
    return copy.deepcopy(self.classes[cls or "main"][kind])


def _resolve(defaults, kind, el, childclass):
  cls = el.get("class", childclass)
  a = defaults.get(kind, cls)
  for k, v in _attrs(el).items():
    if isinstance(v, list) and isinstance(a.get(k), list):
      a[k] = _merge_array(a[k], v) if k not in ("fromto", "pos", "quat", "axis") else v
    else:
      a[k] = v
  return a


class _Compiler:
  def __init__(self, root: ET.Element, extra_classes=None):
    self.root = root
    self._settings(root)
    self.defaults = _Defaults()
    dflt = root.find("default")
    if dflt is not None:
      self.defaults.build(dflt)
    else:
      self.defaults.build(ET.Element("default"))
    # default classes of attached models, under their prefixed names
    self.defaults.classes.update(extra_classes or {})

  def _settings(self, root):
    comp = root.find("compiler")
    self.angle_deg = True
    self.autolimits = True
    self.eulerseq = "xyz"
    if comp is not None:
      self.angle_deg = comp.get("angle", "degree") == "degree"
      self.autolimits = comp.get("autolimits", "true") == "true"
      self.eulerseq = comp.get("eulerseq", "xyz")

  # ---- orientation ----
  def _angle(self, a):
    return a * math.pi / 180.0 if self.angle_deg else a

  def orientation(self, a):
    if "quat" in a:
      q = np.array(a["quat"], dtype=np.float64)
      return q / np.linalg.norm(q)
    if "axisangle" in a:
      v = a["axisangle"]
      return _axisangle(v[:3], self._angle(v[3]))
    if "euler" in a:
      q = np.array([1.0, 0.0, 0.0, 0.0])
      for c, ang in zip(self.eulerseq, a["euler"]):
        ax = {"x": [1, 0, 0], "y": [0, 1, 0], "z": [0, 0, 1]}[c.lower()]
        r = _axisangle(ax, self._angle(ang))
        q = _qmul(q, r) if c.islower() else _qmul(r, q)
      return q / np.linalg.norm(q)
    if "zaxis" in a:
      return _z2quat(a["zaxis"])
    if "xyaxes" in a:
      v = np.array(a["xyaxes"], dtype=np.float64)
      x = v[:3] / np.linalg.norm(v[:3])
      y = v[3:] - x * np.dot(x, v[3:])
      y /= np.linalg.norm(y)
      z = np.cross(x, y)
      return _m2q(np.stack([x, y, z], axis=1))
    return np.array([1.0, 0.0, 0.0, 0.0])

  # ---- compile ----
  def compile(self) -> MjModel:
    m = MjModel()
    self._option(m)
    self.bodies = []  # dicts
    self.joints = []
    self.geoms = []
    self.sites = []
    world = self.root.find("worldbody")
    self.bodies.append(dict(name="world", parent=-1, pos=np.zeros(3), quat=np.array([1.0, 0, 0, 0]),
                            mocap=False, inertial=None, el=world))
    self._body_children(world, 0, None)
    self._finish(m)
    return m

  def _option(self, m):
    opt = self.root.find("option")
    if opt is None:
      return
    a = opt.attrib
    o = m.opt
    if "timestep" in a: o.timestep = float(a["timestep"])
    if "impratio" in a: o.impratio = float(a["impratio"])
    if "tolerance" in a: o.tolerance = float(a["tolerance"])
    if "ls_tolerance" in a: o.ls_tolerance = float(a["ls_tolerance"])
    if "gravity" in a: o.gravity = np.array(_floats(a["gravity"]))
    if "integrator" in a: o.integrator = _INTEGRATORS[a["integrator"].lower()]
    if "cone" in a: o.cone = _CONES[a["cone"].lower()]
    if "solver" in a: o.solver = _SOLVERS[a["solver"].lower()]
    if "iterations" in a: o.iterations = int(a["iterations"])
    if "ls_iterations" in a: o.ls_iterations = int(a["ls_iterations"])
    if "jacobian" in a: o.jacobian = _JACOBIANS[a["jacobian"].lower()]
    flag = opt.find("flag")
    if flag is not None:
      for k, v in flag.attrib.items():
        if k in _DISABLE_FLAGS and v == "disable":
          o.disableflags |= _DISABLE_FLAGS[k]

  def _body_children(self, el, bid, childclass):
    for child in el:
      if child.tag in _UNSUPPORTED_ELEMENTS:
        raise NotImplementedError(f"MJCF element <{child.tag}> is not supported by this compiler")
      if child.tag == "geom":
        self._geom(child, bid, childclass)
      elif child.tag == "site":
        self._site(child, bid, childclass)
      elif child.tag in ("joint", "freejoint"):
        self._joint(child, bid, childclass)
      elif child.tag == "inertial":
        self.bodies[bid]["inertial"] = _attrs(child)
    for child in el:
      if child.tag == "body":
        a = _attrs(child)
        cc = a.get("childclass", childclass)
        nb = dict(name=a.get("name", ""), parent=bid, pos=np.array(a.get("pos", [0.0, 0.0, 0.0])),
                  quat=self.orientation(a), mocap=a.get("mocap", "false") == "true", inertial=None, el=child)
        self.bodies.append(nb)
        self._body_children(child, len(self.bodies) - 1, cc)

  def _joint(self, el, bid, childclass):
    if el.tag == "freejoint":
      a = dict(_JOINT_DEFAULT)
      a.update(_attrs(el))
      a["type"] = "free"
    else:
      a = _resolve(self.defaults, "joint", el, childclass)
    a["body"] = bid
    self.joints.append(a)

  def _geom(self, el, bid, childclass):
    a = _resolve(self.defaults, "geom", el, childclass)
    a["body"] = bid
    self.geoms.append(a)

  def _site(self, el, bid, childclass):
    a = _resolve(self.defaults, "site", el, childclass)
    a["body"] = bid
    self.sites.append(a)

  # ---- geom shape, mass and inertia ----
  def _geom_frame(self, g):
    t = _GEOM_TYPES[g["type"]]
    size = np.array(_merge_array([0.0, 0.0, 0.0], g.get("size", [0.0, 0.0, 0.0]))[:3], dtype=np.float64)
    if "fromto" in g:
      ft = np.array(g["fromto"], dtype=np.float64)
      p0, p1 = ft[:3], ft[3:]
      pos = 0.5 * (p0 + p1)
      quat = _z2quat(p1 - p0)
      half = 0.5 * np.linalg.norm(p1 - p0)
      if t in (3, 5):
        size[1] = half
      elif t in (4, 6):
        # the two sizes orthogonal to the axis; one given value serves both
        if size[1] == 0.0:
          size[1] = size[0]
        size[2] = half
    else:
      pos = np.array(g.get("pos", [0.0, 0.0, 0.0]), dtype=np.float64)
      quat = self.orientation(g)
    if t == 2:
      size[1] = size[2] = 0.0
    return t, size, pos, quat

  @staticmethod
  def _shape_mass_inertia(t, size, density):
    r = size[0]
    if t == 2:  # sphere
      vol = 4.0 / 3.0 * math.pi * r ** 3
      mass = density * vol
      return mass, np.full(3, 0.4 * mass * r * r)
    if t == 3:  # capsule: cylinder plus two hemispheres
      h = 2.0 * size[1]
      ms = density * 4.0 / 3.0 * math.pi * r ** 3
      mc = density * math.pi * r * r * h
      ixx = mc * (3 * r * r + h * h) / 12.0 + ms * (0.4 * r * r + 0.25 * h * h + 3.0 * h * r / 8.0)
      izz = mc * r * r / 2.0 + ms * 0.4 * r * r
      return ms + mc, np.array([ixx, ixx, izz])
    if t == 5:  # cylinder
      h = 2.0 * size[1]
      mass = density * math.pi * r * r * h
      ixx = mass * (3 * r * r + h * h) / 12.0
      return mass, np.array([ixx, ixx, mass * r * r / 2.0])
    if t == 6:  # box
      a, b, c = 2 * size
      mass = density * a * b * c
      return mass, mass / 12.0 * np.array([b * b + c * c, a * a + c * c, a * a + b * b])
    if t == 4:  # ellipsoid
      a, b, c = size
      mass = density * 4.0 / 3.0 * math.pi * a * b * c
      return mass, mass / 5.0 * np.array([b * b + c * c, a * a + c * c, a * a + b * b])
    return 0.0, np.zeros(3)

  @staticmethod
  def _rbound(t, size):
    if t == 0:
      return 0.0
    if t == 2:
      return size[0]
    if t == 3:
      return size[0] + size[1]
    if t == 5:
      return math.sqrt(size[0] ** 2 + size[1] ** 2)
    if t == 4:
      return float(np.max(size))
    if t == 6:
      return float(np.linalg.norm(size))
    return 0.0

  def _finish(self, m: MjModel):
    nbody = len(self.bodies)
    njnt = len(self.joints)
    ngeom = len(self.geoms)
    nsite = len(self.sites)
    m.nbody, m.njnt, m.ngeom, m.nsite = nbody, njnt, ngeom, nsite
    m.names = {"body": [b["name"] for b in self.bodies],
               "joint": [j.get("name", "") for j in self.joints],
               "geom": [g.get("name", "") for g in self.geoms],
               "site": [s.get("name", "") for s in self.sites]}

    # --- bodies ---
    m.body_parentid = np.array([b["parent"] for b in self.bodies], dtype=np.int32)
    m.body_parentid[0] = 0
    m.body_pos = np.array([b["pos"] for b in self.bodies], dtype=np.float64)
    m.body_quat = np.array([b["quat"] for b in self.bodies], dtype=np.float64)
    mocap = [b["mocap"] for b in self.bodies]
    m.body_mocapid = np.full(nbody, -1, dtype=np.int32)
    k = 0
    for i in range(nbody):
      if mocap[i]:
        m.body_mocapid[i] = k
        k += 1
    m.nmocap = k

    # --- joints / dofs ---
    m.jnt_type = np.array([_JOINT_TYPES[j["type"]] for j in self.joints], dtype=np.int32)
    m.jnt_bodyid = np.array([j["body"] for j in self.joints], dtype=np.int32)
    # joints are listed in body order already (bodies are visited depth first)
    qposadr, dofadr = [], []
    nq = nv = 0
    for jt in m.jnt_type:
      qposadr.append(nq)
      dofadr.append(nv)
      nq += {0: 7, 1: 4, 2: 1, 3: 1}[int(jt)]
      nv += {0: 6, 1: 3, 2: 1, 3: 1}[int(jt)]
    m.nq, m.nv = nq, nv
    m.jnt_qposadr = np.array(qposadr, dtype=np.int32)
    m.jnt_dofadr = np.array(dofadr, dtype=np.int32)
    m.jnt_pos = np.array([j.get("pos", [0, 0, 0]) for j in self.joints], dtype=np.float64).reshape(njnt, 3)
    axes = []
    for j in self.joints:
      ax = np.array(j.get("axis", [0, 0, 1]), dtype=np.float64)
      if j["type"] in ("free", "ball"):
        ax = np.array([0.0, 0.0, 1.0])
      axes.append(ax / np.linalg.norm(ax))
    m.jnt_axis = np.array(axes, dtype=np.float64).reshape(njnt, 3)
    m.jnt_stiffness = np.array([float(np.atleast_1d(j.get("stiffness", 0.0))[0]) for j in self.joints])
    rng, lim = [], []
    for j in self.joints:
      r = np.array(j.get("range", [0.0, 0.0]), dtype=np.float64)
      if j["type"] in ("hinge", "ball"):
        r = np.array([self._angle(x) for x in r])
      rng.append(r)
      lmt = j.get("limited", "auto")
      if lmt == "auto":
        limited = self.autolimits and "range" in j and (r[0] != 0 or r[1] != 0) and j["type"] != "free"
        limited = limited and not (j.get("range") == _JOINT_DEFAULT["range"])
      else:
        limited = lmt == "true"
      lim.append(1 if limited else 0)
    m.jnt_range = np.array(rng, dtype=np.float64).reshape(njnt, 2)
    m.jnt_limited = np.array(lim, dtype=np.int32)
    m.jnt_solref = np.array([j.get("solreflimit") for j in self.joints], dtype=np.float64).reshape(njnt, 2)
    m.jnt_solimp = np.array([j.get("solimplimit") for j in self.joints], dtype=np.float64).reshape(njnt, 5)
    m.jnt_margin = np.array([float(np.atleast_1d(j.get("margin", 0.0))[0]) for j in self.joints])
    afl, afr = [], []
    for j in self.joints:
      rr = np.array(j.get("actuatorfrcrange", [0.0, 0.0]), dtype=np.float64)
      flag = j.get("actuatorfrclimited", "auto")
      afl.append(bool(flag == "true" or (flag == "auto" and self.autolimits and (rr[0] != 0 or rr[1] != 0))))
      afr.append(rr)
    m.jnt_actfrclimited = np.array(afl, dtype=bool)
    m.jnt_actfrcrange = np.array(afr, dtype=np.float64).reshape(njnt, 2)

    dof_bodyid, dof_jntid, dof_armature, dof_damping = [], [], [], []
    for j, jd in enumerate(self.joints):
      n = {0: 6, 1: 3, 2: 1, 3: 1}[int(m.jnt_type[j])]
      for _ in range(n):
        dof_bodyid.append(jd["body"])
        dof_jntid.append(j)
        dof_armature.append(float(np.atleast_1d(jd.get("armature", 0.0))[0]))
        dof_damping.append(float(np.atleast_1d(jd.get("damping", 0.0))[0]))
    m.dof_bodyid = np.array(dof_bodyid, dtype=np.int32)
    m.dof_jntid = np.array(dof_jntid, dtype=np.int32)
    m.dof_armature = np.array(dof_armature, dtype=np.float64)
    m.dof_damping = np.array(dof_damping, dtype=np.float64)

    # body joint / dof tables
    m.body_jntnum = np.zeros(nbody, dtype=np.int32)
    m.body_jntadr = np.full(nbody, -1, dtype=np.int32)
    m.body_dofnum = np.zeros(nbody, dtype=np.int32)
    m.body_dofadr = np.full(nbody, -1, dtype=np.int32)
    for j in range(njnt):
      b = m.jnt_bodyid[j]
      if m.body_jntnum[b] == 0:
        m.body_jntadr[b] = j
      m.body_jntnum[b] += 1
    for i in range(nv):
      b = m.dof_bodyid[i]
      if m.body_dofnum[b] == 0:
        m.body_dofadr[b] = i
      m.body_dofnum[b] += 1

    # dof_parentid: previous dof of the same body, else last dof of the
    # nearest ancestor with dofs
    m.dof_parentid = np.full(nv, -1, dtype=np.int32)
    last_dof = np.full(nbody, -1, dtype=np.int64)
    for b in range(1, nbody):
      p = m.body_parentid[b]
      prev = last_dof[p]
      for i in range(m.body_dofadr[b], m.body_dofadr[b] + m.body_dofnum[b]) if m.body_dofnum[b] else []:
        m.dof_parentid[i] = prev
        prev = i
      last_dof[b] = prev
    depth = np.zeros(nv, dtype=np.int64)
    for i in range(nv):
      depth[i] = 0 if m.dof_parentid[i] < 0 else depth[m.dof_parentid[i]] + 1
    m.dof_Madr = np.zeros(nv, dtype=np.int32)
    adr = 0
    for i in range(nv):
      m.dof_Madr[i] = adr
      adr += depth[i] + 1
    m.nM = int(adr)

    # weld / root ids and tree ids
    m.body_weldid = np.zeros(nbody, dtype=np.int32)
    m.body_rootid = np.zeros(nbody, dtype=np.int32)
    for b in range(1, nbody):
      p = m.body_parentid[b]
      m.body_weldid[b] = b if m.body_jntnum[b] > 0 else m.body_weldid[p]
      m.body_rootid[b] = b if p == 0 else m.body_rootid[p]
    m.body_treeid = np.full(nbody, -1, dtype=np.int32)
    tree_of_root = {}
    for b in range(1, nbody):
      r = m.body_rootid[b]
      # a tree is a root subtree containing dofs
      pass
    has_dof = np.zeros(nbody, dtype=bool)
    for b in range(nbody - 1, 0, -1):
      has_dof[b] |= m.body_dofnum[b] > 0
      has_dof[m.body_parentid[b]] |= has_dof[b]
    ntree = 0
    for b in range(1, nbody):
      if m.body_parentid[b] == 0 and has_dof[b]:
        tree_of_root[b] = ntree
        ntree += 1
    for b in range(1, nbody):
      r = m.body_rootid[b]
      if r in tree_of_root:
        m.body_treeid[b] = tree_of_root[r]
    m.ntree = ntree
    m.dof_treeid = np.array([m.body_treeid[b] for b in m.dof_bodyid], dtype=np.int32)

    # --- qpos0 / springs ---
    m.qpos0 = np.zeros(nq)
    m.qpos_spring = np.zeros(nq)
    for j, jd in enumerate(self.joints):
      qa = m.jnt_qposadr[j]
      jt = int(m.jnt_type[j])
      b = jd["body"]
      if jt == 0:
        m.qpos0[qa:qa + 3] = m.body_pos[b]
        m.qpos0[qa + 3:qa + 7] = m.body_quat[b]
        m.qpos_spring[qa:qa + 7] = m.qpos0[qa:qa + 7]
      elif jt == 1:
        m.qpos0[qa:qa + 4] = [1.0, 0.0, 0.0, 0.0]
        m.qpos_spring[qa:qa + 4] = m.qpos0[qa:qa + 4]
      else:
        ref = float(np.atleast_1d(jd.get("ref", 0.0))[0])
        sref = float(np.atleast_1d(jd.get("springref", 0.0))[0])
        if jt == 3:
          ref, sref = self._angle(ref), self._angle(sref)
        m.qpos0[qa] = ref
        m.qpos_spring[qa] = sref

    # --- geoms ---
    gt, gsize, gpos, gquat, gmass, ginert = [], [], [], [], [], []
    for g in self.geoms:
      t, size, pos, quat = self._geom_frame(g)
      gt.append(t)
      gsize.append(size)
      gpos.append(pos)
      gquat.append(quat)
      mass, inert = self._shape_mass_inertia(t, size, float(np.atleast_1d(g.get("density", 1000.0))[0]))
      if g.get("mass") is not None:
        gm = float(np.atleast_1d(g["mass"])[0])
        inert = inert * (gm / mass) if mass > 0 else inert
        mass = gm
      gmass.append(mass)
      ginert.append(inert)
    m.geom_type = np.array(gt, dtype=np.int32)
    m.geom_size = np.array(gsize, dtype=np.float64).reshape(ngeom, 3)
    m.geom_pos = np.array(gpos, dtype=np.float64).reshape(ngeom, 3)
    m.geom_quat = np.array(gquat, dtype=np.float64).reshape(ngeom, 4)
    m.geom_bodyid = np.array([g["body"] for g in self.geoms], dtype=np.int32)
    m.geom_contype = np.array([int(np.atleast_1d(g["contype"])[0]) for g in self.geoms], dtype=np.int32)
    m.geom_conaffinity = np.array([int(np.atleast_1d(g["conaffinity"])[0]) for g in self.geoms], dtype=np.int32)
    m.geom_condim = np.array([int(np.atleast_1d(g["condim"])[0]) for g in self.geoms], dtype=np.int32)
    m.geom_priority = np.array([int(np.atleast_1d(g["priority"])[0]) for g in self.geoms], dtype=np.int32)
    m.geom_solmix = np.array([float(np.atleast_1d(g["solmix"])[0]) for g in self.geoms])
    m.geom_solref = np.array([g["solref"] for g in self.geoms], dtype=np.float64).reshape(ngeom, 2)
    m.geom_solimp = np.array([g["solimp"] for g in self.geoms], dtype=np.float64).reshape(ngeom, 5)
    m.geom_friction = np.array([g["friction"] for g in self.geoms], dtype=np.float64).reshape(ngeom, 3)
    m.geom_margin = np.array([float(np.atleast_1d(g["margin"])[0]) for g in self.geoms])
    m.geom_gap = np.array([float(np.atleast_1d(g["gap"])[0]) for g in self.geoms])
    m.geom_dataid = np.full(ngeom, -1, dtype=np.int32)
    m.geom_rbound = np.array([self._rbound(t, s) for t, s in zip(gt, gsize)], dtype=np.float64)
    aabb = np.zeros((ngeom, 2, 3))
    for i, (t, s) in enumerate(zip(gt, gsize)):
      if t == 2:
        aabb[i, 1] = s[0]
      elif t == 3:
        aabb[i, 1] = [s[0], s[0], s[0] + s[1]]
      elif t == 5:
        aabb[i, 1] = [s[0], s[0], s[1]]
      elif t in (4, 6):
        aabb[i, 1] = s
      elif t == 0:
        aabb[i, 1] = [s[0] if s[0] > 0 else 1e10, s[1] if s[1] > 0 else 1e10, 0.0]
    m.geom_aabb = aabb
    m.body_geomnum = np.zeros(nbody, dtype=np.int32)
    m.body_geomadr = np.full(nbody, -1, dtype=np.int32)
    # geoms must be grouped by body in body order (MuJoCo's layout)
    order = np.argsort(m.geom_bodyid, kind="stable")
    if not np.array_equal(order, np.arange(ngeom)):
      raise ValueError("internal: geoms not in body order")
    for i in range(ngeom):
      b = m.geom_bodyid[i]
      if m.body_geomnum[b] == 0:
        m.body_geomadr[b] = i
      m.body_geomnum[b] += 1
    m.body_contype = np.zeros(nbody, dtype=np.int32)
    m.body_conaffinity = np.zeros(nbody, dtype=np.int32)
    for i in range(ngeom):
      m.body_contype[m.geom_bodyid[i]] |= m.geom_contype[i]
      m.body_conaffinity[m.geom_bodyid[i]] |= m.geom_conaffinity[i]

    # --- body inertial frames ---
    m.body_mass = np.zeros(nbody)
    m.body_ipos = np.zeros((nbody, 3))
    m.body_iquat = np.tile([1.0, 0.0, 0.0, 0.0], (nbody, 1))
    m.body_inertia = np.zeros((nbody, 3))
    for b in range(1, nbody):
      ine = self.bodies[b]["inertial"]
      if ine is not None:
        m.body_mass[b] = float(ine["mass"][0])
        m.body_ipos[b] = ine.get("pos", [0.0, 0.0, 0.0])
        if "fullinertia" in ine:
          f = ine["fullinertia"]
          full = np.array([[f[0], f[3], f[4]], [f[3], f[1], f[5]], [f[4], f[5], f[2]]])
          ev, vecs = np.linalg.eigh(full)
          idx = np.argsort(-ev)
          ev, vecs = ev[idx], vecs[:, idx]
          if np.linalg.det(vecs) < 0:
            vecs[:, 2] = -vecs[:, 2]
          m.body_inertia[b] = ev
          m.body_iquat[b] = _m2q(vecs)
        else:
          m.body_inertia[b] = ine["diaginertia"]
          m.body_iquat[b] = self.orientation(ine)
        continue
      gs = [i for i in range(ngeom) if m.geom_bodyid[i] == b and gt[i] != 0]
      mass = sum(gmass[i] for i in gs)
      if mass <= 0:
        continue
      com = sum(gmass[i] * m.geom_pos[i] for i in gs) / mass
      full = np.zeros((3, 3))
      for i in gs:
        rot = _q2m(m.geom_quat[i])
        full += rot @ np.diag(ginert[i]) @ rot.T
        dd = m.geom_pos[i] - com
        full += gmass[i] * (np.dot(dd, dd) * np.eye(3) - np.outer(dd, dd))
      ev, vecs = np.linalg.eigh(full)
      idx = np.argsort(-ev, kind="stable")
      ev, vecs = ev[idx], vecs[:, idx]
      if np.linalg.det(vecs) < 0:
        vecs[:, 2] = -vecs[:, 2]
      m.body_mass[b] = mass
      m.body_ipos[b] = com
      m.body_inertia[b] = ev
      m.body_iquat[b] = _m2q(vecs)

    # --- sites ---
    m.site_bodyid = np.array([s["body"] for s in self.sites], dtype=np.int32)
    m.site_pos = np.array([s.get("pos", [0, 0, 0]) for s in self.sites], dtype=np.float64).reshape(nsite, 3)
    m.site_quat = np.array([self.orientation(s) for s in self.sites], dtype=np.float64).reshape(nsite, 4)

    # --- actuators ---
    self._actuators(m)

    # --- keyframes ---
    self._keyframes(m)

    # --- misc sizes ---
    m.na = 0
    m.nexclude = 0
    m.exclude_signature = np.zeros(0, dtype=np.int32)
    m.nmesh = 0
    m.mesh_vertadr = np.zeros(0, dtype=np.int32)
    m.mesh_vertnum = np.zeros(0, dtype=np.int32)
    m.mesh_vert = np.zeros((0, 3))
    m.neq = m.nsensor = m.ntendon = 0

    set_const(m)

  def _actuators(self, m):
    acts = []
    el = self.root.find("actuator")
    if el is not None:
      for a in el:
        if a.tag not in ("motor", "position", "general", "velocity"):
          raise NotImplementedError(f"actuator <{a.tag}> is unsupported")
        d = _resolve(self.defaults, "actuator", a, None)
        d["tag"] = a.tag
        acts.append(d)
    nu = len(acts)
    m.nu = nu
    m.actuator_trntype = np.zeros(nu, dtype=np.int32)
    m.actuator_dyntype = np.zeros(nu, dtype=np.int32)
    m.actuator_gaintype = np.zeros(nu, dtype=np.int32)
    m.actuator_biastype = np.zeros(nu, dtype=np.int32)
    m.actuator_trnid = np.full((nu, 2), -1, dtype=np.int32)
    m.actuator_actadr = np.full(nu, -1, dtype=np.int32)
    m.actuator_ctrllimited = np.zeros(nu, dtype=bool)
    m.actuator_forcelimited = np.zeros(nu, dtype=bool)
    m.actuator_actlimited = np.zeros(nu, dtype=bool)
    m.actuator_dynprm = np.zeros((nu, 10))
    m.actuator_gainprm = np.zeros((nu, 10))
    m.actuator_biasprm = np.zeros((nu, 10))
    m.actuator_ctrlrange = np.zeros((nu, 2))
    m.actuator_forcerange = np.zeros((nu, 2))
    m.actuator_actrange = np.zeros((nu, 2))
    m.actuator_gear = np.zeros((nu, 6))
    jnames = m.names["joint"]
    for i, a in enumerate(acts):
      if "joint" in a:
        m.actuator_trntype[i] = 0
        m.actuator_trnid[i, 0] = jnames.index(a["joint"])
      elif "jointinparent" in a:
        m.actuator_trntype[i] = 1
        m.actuator_trnid[i, 0] = jnames.index(a["jointinparent"])
      else:
        raise NotImplementedError("only joint transmissions are supported")
      m.actuator_gear[i] = _merge_array([1.0, 0, 0, 0, 0, 0], a["gear"])[:6]
      cr = np.array(a["ctrlrange"], dtype=np.float64)
      fr = np.array(a["forcerange"], dtype=np.float64)
      m.actuator_ctrlrange[i] = cr
      m.actuator_forcerange[i] = fr
      def _lim(flag, r):
        return flag == "true" or (flag == "auto" and self.autolimits and (r[0] != 0 or r[1] != 0))
      m.actuator_ctrllimited[i] = _lim(a["ctrllimited"], cr)
      m.actuator_forcelimited[i] = _lim(a["forcelimited"], fr)
      if a["tag"] == "motor":
        m.actuator_gainprm[i, 0] = 1.0
      elif a["tag"] == "position":
        kp = float(np.atleast_1d(a.get("kp", 1.0))[0])
        kv = float(np.atleast_1d(a.get("kv", 0.0))[0])
        m.actuator_gainprm[i, 0] = kp
        m.actuator_biastype[i] = 1
        m.actuator_biasprm[i, :3] = [0.0, -kp, -kv]
      elif a["tag"] == "velocity":
        kv = float(np.atleast_1d(a.get("kv", 1.0))[0])
        m.actuator_gainprm[i, 0] = kv
        m.actuator_biastype[i] = 1
        m.actuator_biasprm[i, :3] = [0.0, 0.0, -kv]
      else:  # general
        m.actuator_gaintype[i] = {"fixed": 0, "affine": 1}[a.get("gaintype", "fixed")]
        m.actuator_biastype[i] = {"none": 0, "affine": 1}[a.get("biastype", "none")]
        if a.get("gainprm") is not None:
          m.actuator_gainprm[i, : len(a["gainprm"])] = a["gainprm"]
        else:
          m.actuator_gainprm[i, 0] = 1.0
        if a.get("biasprm") is not None:
          m.actuator_biasprm[i, : len(a["biasprm"])] = a["biasprm"]
        if a.get("dyntype", "none") != "none":
          raise NotImplementedError("actuator dynamics are unsupported")

  def _keyframes(self, m):
    keys = []
    el = self.root.find("keyframe")
    if el is not None:
      keys = [k for k in el if k.tag == "key"]
    nkey = len(keys)
    m.nkey = nkey
    m.key_time = np.zeros(nkey)
    m.key_qpos = np.tile(m.qpos0, (nkey, 1))
    m.key_qvel = np.zeros((nkey, m.nv))
    m.key_act = np.zeros((nkey, 0))
    m.key_ctrl = np.zeros((nkey, m.nu))
    m.names["key"] = [k.get("name", "") for k in keys]
    for i, k in enumerate(keys):
      if "time" in k.attrib:
        m.key_time[i] = float(k.get("time"))
      if "qpos" in k.attrib:
        m.key_qpos[i] = _floats(k.get("qpos"))
      if "qvel" in k.attrib:
        m.key_qvel[i] = _floats(k.get("qvel"))
      if "ctrl" in k.attrib:
        m.key_ctrl[i] = _floats(k.get("ctrl"))


# ---------------------------------------------------------------------------
# compile-time constants (MuJoCo's mj_setConst quantities, in float64)

def _fk_com(m: MjModel, qpos):
  """Body frames, subtree COM and cdof at qpos (float64)."""
  nbody = m.nbody
  xpos = np.zeros((nbody, 3))
  xquat = np.tile([1.0, 0.0, 0.0, 0.0], (nbody, 1))
  xanchor = np.zeros((m.njnt, 3))
  xaxis = np.zeros((m.njnt, 3))
  for b in range(1, nbody):
    p = m.body_parentid[b]
    ja, jn = m.body_jntadr[b], m.body_jntnum[b]
    if jn == 1 and m.jnt_type[ja] == 0:
      qa = m.jnt_qposadr[ja]
      pos = qpos[qa:qa + 3].copy()
      quat = qpos[qa + 3:qa + 7].copy()
      xanchor[ja] = pos
      xaxis[ja] = m.jnt_axis[ja]
    else:
      pos = _q2m(xquat[p]) @ m.body_pos[b] + xpos[p]
      quat = _qmul(xquat[p], m.body_quat[b])
      for j in range(ja, ja + jn):
        qa = m.jnt_qposadr[j]
        anchor = _qrot(quat, m.jnt_pos[j]) + pos
        axis = _qrot(quat, m.jnt_axis[j])
        if m.jnt_type[j] == 1:
          quat = _qmul(quat, qpos[qa:qa + 4])
          pos = anchor - _qrot(quat, m.jnt_pos[j])
        elif m.jnt_type[j] == 2:
          pos = pos + axis * (qpos[qa] - m.qpos0[qa])
        elif m.jnt_type[j] == 3:
          quat = _qmul(quat, _axisangle(m.jnt_axis[j], qpos[qa] - m.qpos0[qa]))
          pos = anchor - _qrot(quat, m.jnt_pos[j])
        xanchor[j] = anchor
        xaxis[j] = axis
    xpos[b] = pos
    xquat[b] = quat / np.linalg.norm(quat)
  xipos = np.array([xpos[b] + _qrot(xquat[b], m.body_ipos[b]) for b in range(nbody)])
  mass = m.body_mass
  tot = np.zeros((nbody, 3))
  msum = mass.copy()
  tot[:] = xipos * mass[:, None]
  for b in range(nbody - 1, 0, -1):
    tot[m.body_parentid[b]] += tot[b]
    msum[m.body_parentid[b]] += msum[b]
  subtree_com = np.where(msum[:, None] > _MINVAL, tot / np.maximum(msum, _MINVAL)[:, None], xipos)
  cdof = np.zeros((m.nv, 6))
  for j in range(m.njnt):
    b = m.jnt_bodyid[j]
    da = m.jnt_dofadr[j]
    off = subtree_com[m.body_rootid[b]] - xanchor[j]
    xm = _q2m(xquat[b])
    if m.jnt_type[j] == 0:
      for i in range(3):
        cdof[da + i, 3 + i] = 1.0
      for i in range(3):
        cdof[da + 3 + i, :3] = xm[:, i]
        cdof[da + 3 + i, 3:] = np.cross(xm[:, i], off)
    elif m.jnt_type[j] == 1:
      for i in range(3):
        cdof[da + i, :3] = xm[:, i]
        cdof[da + i, 3:] = np.cross(xm[:, i], off)
    elif m.jnt_type[j] == 2:
      cdof[da, 3:] = xaxis[j]
    else:
      cdof[da, :3] = xaxis[j]
      cdof[da, 3:] = np.cross(xaxis[j], off)
  return xpos, xquat, xipos, subtree_com, cdof


def _chain_dofs(m: MjModel, b) -> np.ndarray:
  """Dofs of body b and its ancestors, ascending: the columns where a point
  on b has a nonzero Jacobian."""
  out = []
  c = b
  while c != 0:
    out.extend(range(m.body_dofadr[c], m.body_dofadr[c] + m.body_dofnum[c]))
    c = m.body_parentid[c]
  return np.array(sorted(out), dtype=np.int64)


def _body_jacobian(m: MjModel, point, b, subtree_com, cdof):
  """Translational and rotational Jacobians (3 x nv) of a point on body b."""
  jp = np.zeros((3, m.nv))
  jr = np.zeros((3, m.nv))
  if b == 0:
    return jp, jr
  off = point - subtree_com[m.body_rootid[b]]
  for i in _chain_dofs(m, b):
    jr[:, i] = cdof[i, :3]
    jp[:, i] = cdof[i, 3:] + np.cross(cdof[i, :3], off)
  return jp, jr


def mass_matrix(m: MjModel, qpos) -> np.ndarray:
  """Dense joint-space inertia (armature included) at qpos, float64; each
  body adds to the rows / columns of its own and its ancestors' dofs."""
  _, xquat, xipos, subtree_com, cdof = _fk_com(m, qpos)
  M = np.diag(m.dof_armature.astype(np.float64)).copy()
  for b in range(1, m.nbody):
    if m.body_mass[b] <= 0 and not np.any(m.body_inertia[b]):
      continue
    jp, jr = _body_jacobian(m, xipos[b], b, subtree_com, cdof)
    idx = _chain_dofs(m, b)
    jp, jr = jp[:, idx], jr[:, idx]
    R = _q2m(_qmul(xquat[b], m.body_iquat[b]))
    Iw = R @ np.diag(m.body_inertia[b]) @ R.T
    M[np.ix_(idx, idx)] += m.body_mass[b] * jp.T @ jp + jr.T @ Iw @ jr
  return M


def _tree_blocks(m: MjModel):
  """[begin, end) dof ranges of the kinematic trees (dofs of a tree are
  contiguous, roots have dof_parentid -1)."""
  roots = [i for i in range(m.nv) if m.dof_parentid[i] == -1]
  return [(r, roots[k + 1] if k + 1 < len(roots) else m.nv) for k, r in enumerate(roots)]


def set_const(m: MjModel):
  """stat.meaninertia, dof_invweight0 and body_invweight0 at qpos0."""
  if m.nv == 0:
    m.stat.meaninertia = 1.0
    m.dof_invweight0 = np.zeros(0)
    m.body_invweight0 = np.zeros((m.nbody, 2))
    return
  M = mass_matrix(m, m.qpos0)
  m.stat.meaninertia = float(np.trace(M) / m.nv)
  # M is block diagonal over the kinematic trees: invert block by block
  Minv = np.zeros_like(M)
  for a, e in _tree_blocks(m):
    Minv[a:e, a:e] = np.linalg.inv(M[a:e, a:e])
  dinv = np.diag(Minv).copy()
  m.dof_invweight0 = np.zeros(m.nv)
  for j in range(m.njnt):
    da = m.jnt_dofadr[j]
    if m.jnt_type[j] == 0:
      m.dof_invweight0[da:da + 3] = dinv[da:da + 3].mean()
      m.dof_invweight0[da + 3:da + 6] = dinv[da + 3:da + 6].mean()
    elif m.jnt_type[j] == 1:
      m.dof_invweight0[da:da + 3] = dinv[da:da + 3].mean()
    else:
      m.dof_invweight0[da] = dinv[da]
  _, _, xipos, subtree_com, cdof = _fk_com(m, m.qpos0)
  m.body_invweight0 = np.zeros((m.nbody, 2))
  for b in range(1, m.nbody):
    if m.body_weldid[b] == 0:
      continue
    jp, jr = _body_jacobian(m, xipos[b], b, subtree_com, cdof)
    idx = _chain_dofs(m, b)
    J = np.vstack([jp, jr])[:, idx]
    A = J @ Minv[np.ix_(idx, idx)] @ J.T
    m.body_invweight0[b, 0] = max(_MINVAL, np.trace(A[:3, :3]) / 3.0)
    m.body_invweight0[b, 1] = max(_MINVAL, np.trace(A[3:, 3:]) / 3.0)


# ---------------------------------------------------------------------------
# <replicate>, <frame> and <attach> (with <model file=...> assets): expanded
# into a plain worldbody before compiling.
#
# * frame: its position / orientation compose into its children's (bodies,
#   geoms, sites; a geom's fromto endpoints move with it);
# * replicate count=N: N copies of its children, copy k under the frame of
#   k times `offset` and k compositions of its rotation (euler / quat /
#   axisangle), every name in copy k suffixed with sep + k;
# * attach model=M body=B prefix=P: a copy of body B of the asset model M
#   (an MJCF file beside the including one) with every name prefixed by P,
#   the orientations of M's subtree converted with M's own compiler
#   settings, M's default classes merged under prefixed names (M's "main"
#   becomes P + "main", the attached body's default class), and M's
#   actuators of the subtree's joints appended (names prefixed, and suffixed
#   inside a replicate).
# Lights and cameras, which the step does not use, are dropped in expanded
# subtrees.

_NAMED = ("body", "joint", "freejoint", "geom", "site")


class _Expander:
  def __init__(self, root, base_dir):
    self.root = root
    self.base_dir = base_dir
    self.comp = _Compiler.__new__(_Compiler)  # orientation() of the including model
    _Compiler._settings(self.comp, root)
    self.models = {}
    self.classes = {}
    self.actuators = []

  def _model(self, name):
    if name in self.models:
      return self.models[name]
    asset = self.root.find("asset")
    el = None if asset is None else next((a for a in asset.findall("model") if a.get("name") == name), None)
    if el is None or "file" not in el.attrib:
      raise ValueError(f"attach: no asset <model name={name!r} file=...>")
    if self.base_dir is None:
      raise NotImplementedError("attach of a model file needs the including file's path (from_xml_path)")
    import os

    path = os.path.join(self.base_dir, el.get("file"))
    with open(path) as f:
      sub = ET.fromstring(f.read())
    sub_base = os.path.dirname(os.path.abspath(path))
    sub, sub_classes, sub_acts = _Expander(sub, sub_base).run()
    comp = _Compiler.__new__(_Compiler)
    _Compiler._settings(comp, sub)
    if comp.angle_deg != self.comp.angle_deg:
      raise NotImplementedError("attach: the attached model's angle unit differs from the including model's")
    dflt = _Defaults()
    dflt.build(sub.find("default") if sub.find("default") is not None else ET.Element("default"))
    dflt.classes.update(sub_classes)
    self.models[name] = (sub, comp, dflt, sub_acts)
    return self.models[name]

  def run(self):
    world = self.root.find("worldbody")
    if world is not None:
      self._expand(world)
    act = self.root.find("actuator")
    if self.actuators:
      if act is None:
        act = ET.SubElement(self.root, "actuator")
      for a in self.actuators:
        act.append(a)
    return self.root, self.classes, self.actuators

  # in place: frames, replicates and attaches among el's children (and
  # deeper, inside bodies) become plain elements
  def _expand(self, el):
    out = []
    for child in list(el):
      out.extend(self._flatten(child, np.zeros(3), np.array([1.0, 0, 0, 0]), ""))
    for child in list(el):
      el.remove(child)
    for c in out:
      el.append(c)

  def _flatten(self, el, pos, quat, suffix):
    """el as plain elements under the frame (pos, quat), names suffixed."""
    tag = el.tag
    if tag == "frame":
      fp, fq = self._frame_of(el, self.comp)
      p2, q2 = pos + _qrot(quat, fp), _qmul(quat, fq)
      return [c for child in el for c in self._flatten(child, p2, q2, suffix)]
    if tag == "replicate":
      count = int(el.get("count", "1"))
      sep = el.get("sep", "")
      off = np.array(_floats(el.get("offset", "0 0 0")))
      _, rq = self._frame_of(el, self.comp)
      width = len(str(max(count - 1, 0)))
      out, qk = [], np.array([1.0, 0, 0, 0])
      for k in range(count):
        p2, q2 = pos + _qrot(quat, k * off), _qmul(quat, qk)
        for child in el:
          out.extend(self._flatten(child, p2, q2, suffix + sep + str(k).zfill(width)))
        qk = _qmul(qk, rq)
      return out
    if tag == "attach":
      return [self._attach(el, pos, quat, suffix)]
    if tag in ("light", "camera") and (suffix or np.any(pos) or quat[0] != 1.0):
      return []
    c = copy.deepcopy(el)
    if tag in ("body", "geom", "site"):
      self._place(c, pos, quat, self.comp)
    _rename(c, "", suffix)
    if tag == "body":
      self._expand(c)
    return [c]

  def _frame_of(self, el, comp):
    a = _attrs(el)
    return np.array(a.get("pos", [0.0, 0.0, 0.0]), dtype=np.float64), comp.orientation(a)

  @staticmethod
  def _place(c, pos, quat, comp):
    """c's own frame moved by (pos, quat); orientation written as quat."""
    a = _attrs(c)
    if "fromto" in a:
      f = np.array(a["fromto"], dtype=np.float64)
      f0, f1 = pos + _qrot(quat, f[:3]), pos + _qrot(quat, f[3:])
      c.set("fromto", " ".join(repr(float(x)) for x in np.concatenate([f0, f1])))
      return
    p = np.array(a.get("pos", [0.0, 0.0, 0.0]), dtype=np.float64)
    q = comp.orientation(a)
    for k in ("quat", "euler", "axisangle", "zaxis", "xyaxes"):
      c.attrib.pop(k, None)
    p2, q2 = pos + _qrot(quat, p), _qmul(quat, q)
    c.set("pos", " ".join(repr(float(x)) for x in p2))
    c.set("quat", " ".join(repr(float(x)) for x in q2))

  def _attach(self, el, pos, quat, suffix):
    sub, comp, dflt, sub_acts = self._model(el.get("model"))
    prefix = el.get("prefix", "")
    body = next((b for b in sub.iter("body") if b.get("name") == el.get("body")), None)
    if body is None:
      raise ValueError(f"attach: model {el.get('model')!r} has no body {el.get('body')!r}")
    c = copy.deepcopy(body)
    # the subtree's orientations in the attached model's own convention
    for e in c.iter():
      if e is not c and e.tag in ("body", "geom", "site"):
        self._place(e, np.zeros(3), np.array([1.0, 0, 0, 0]), comp)
    self._place(c, np.zeros(3), np.array([1.0, 0, 0, 0]), comp)
    self._place(c, pos, quat, self.comp)
    for e in c.iter():
      for k in ("class", "childclass"):
        if k in e.attrib:
          e.set(k, prefix + e.get(k))
    if "childclass" not in c.attrib:
      c.set("childclass", prefix + "main")
    for name, cls in dflt.classes.items():
      self.classes[prefix + name] = copy.deepcopy(cls)
    joints = {e.get("name") for e in c.iter() if e.tag in ("joint", "freejoint") and e.get("name")}
    _rename(c, prefix, suffix)
    act = sub.find("actuator")
    for a in [] if act is None else list(act):
      if a.get("joint") in joints:
        a2 = copy.deepcopy(a)
        a2.set("joint", prefix + a.get("joint") + suffix)
        if a2.get("name"):
          a2.set("name", prefix + a2.get("name") + suffix)
        a2.set("class", prefix + a.get("class", "main"))
        self.actuators.append(a2)
    return c


def _rename(el, prefix, suffix):
  if not prefix and not suffix:
    return
  for e in el.iter():
    if e.tag in _NAMED and e.get("name"):
      e.set("name", prefix + e.get("name") + suffix)


# ---------------------------------------------------------------------------
# entry points

_FIX_ATTR = re.compile(r'(=\s*"[^"]*")(?=[A-Za-z_])')


# Elements that change the model's structure and that this compiler does
# not expand: compiling without them would silently drop bodies.
_UNSUPPORTED_ELEMENTS = ("composite", "flexcomp", "include", "plugin")


def from_xml_string(xml: str, base_dir=None) -> MjModel:
  """Compiles an MJCF string (MjModel.from_xml_string counterpart); model
  files named by <attach> are read relative to base_dir."""
  try:
    root = ET.fromstring(xml)
  except ET.ParseError:
    # MJCF parsers accept attributes without separating whitespace
    # (e.g. type="ball"solreflimit=...); insert it and retry.
    root = ET.fromstring(_FIX_ATTR.sub(r"\1 ", xml))
  for el in root.iter():
    if el.tag in _UNSUPPORTED_ELEMENTS:
      raise NotImplementedError(f"MJCF element <{el.tag}> is not supported by this compiler")
  root, classes, _ = _Expander(root, base_dir).run()
  if root.find("equality") is not None and len(root.find("equality")):
    m = _Compiler(root, classes).compile()
    m.neq = len(root.find("equality"))
    return m
  m = _Compiler(root, classes).compile()
  if root.find("sensor") is not None:
    m.nsensor = len(root.find("sensor"))
  if root.find("tendon") is not None:
    m.ntendon = len(root.find("tendon"))
  return m


def from_xml_path(path: str) -> MjModel:
  import os

  with open(path) as f:
    return from_xml_string(f.read(), base_dir=os.path.dirname(os.path.abspath(path)))
