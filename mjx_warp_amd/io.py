"""Host <-> device transfer: put_model, make_data, put_data, get_data_into.

Counterparts of mujoco_warp/_src/io.py:24-831.  The derived index tables the
kernels use (body tree levels, mass-matrix addressing, per-tree tiles,
limited-joint list, line-search candidates, subtree masses) are built here,
with the same definitions as the reference's, plus `body_dof_tree`, a
precomputed (body, dof) ancestry table.
"""

from __future__ import annotations

import numpy as np

from . import _native
from . import types

_SUPPORTED_INTEGRATORS = (0, 3)
_SUPPORTED_CONES = (0,)
_SUPPORTED_SOLVERS = (1, 2)


def _torch():
  import torch

  return torch


def default_device():
  torch = _torch()
  return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


def is_sparse(mjm) -> bool:
  """support.py:29-32: sparse when jacobian is sparse, or auto and nv >= 60."""
  if mjm.opt.jacobian == 2:
    return mjm.nv >= 60
  return mjm.opt.jacobian == 1


def _tensor(x, kind, device):
  torch = _torch()
  dt = {"R": torch.float32, "I": torch.int32, "B": torch.bool}[kind]
  a = np.asarray(x)
  if kind == "R":
    a = a.astype(np.float32)
  elif kind == "I":
    a = a.astype(np.int32)
  else:
    a = a.astype(bool)
  return torch.from_numpy(np.ascontiguousarray(a).copy()).to(device=device, dtype=dt)


def model_tables(mjm) -> dict:
  """Derived index tables of io.py:76-254 (as numpy arrays)."""
  nv, nbody = mjm.nv, mjm.nbody
  t = {}
  t["dof_tri_row"], t["dof_tri_col"] = np.tril_indices(nv)

  # sparse-M addressing (io.py:79-103)
  fi, fj = [], []
  for i in range(nv):
    j = i
    while j > -1:
      fi.append(i)
      fj.append(j)
      j = mjm.dof_parentid[j]
  t["qM_fullm_i"], t["qM_fullm_j"] = np.array(fi, np.int32), np.array(fj, np.int32)
  mi, mj, madr = [], [], []
  for i in range(nv):
    adr, j = mjm.dof_Madr[i], i
    while True:
      adr, j = adr + 1, mjm.dof_parentid[j]
      if j == -1:
        break
      mi.append(i)
      mj.append(j)
      madr.append(adr)
  t["qM_mulm_i"], t["qM_mulm_j"], t["qM_madr_ij"] = (np.array(x, np.int32) for x in (mi, mj, madr))

  limited = [j for j in range(mjm.njnt) if mjm.jnt_limited[j] and mjm.jnt_type[j] in (2, 3)]
  t["jnt_limited_slide_hinge_adr"] = np.array(limited, np.int32)

  # breadth-first body levels (io.py:113-121)
  depth = np.zeros(nbody, dtype=np.int64)
  levels = {}
  for b in range(nbody):
    depth[b] = 0 if b == 0 else depth[mjm.body_parentid[b]] + 1
    levels.setdefault(int(depth[b]), []).append(b)
  t["body_tree"] = np.concatenate([levels[k] for k in range(len(levels))]).astype(np.int32)
  t["body_treeadr"] = np.cumsum([0] + [len(levels[k]) for k in range(len(levels))])[:-1].astype(np.int32)
  t["max_level_width"] = max(len(v) for v in levels.values())

  sparse = is_sparse(mjm)
  t["qLD_update_tree"] = np.zeros((0, 3), np.int32)
  t["qLD_update_treeadr"] = np.zeros(0, np.int32)
  t["qLD_tile"] = np.zeros(0, np.int32)
  t["qLD_tileadr"] = np.zeros(0, np.int32)
  t["qLD_tilesize"] = np.zeros(0, np.int32)
  roots = [i for i in range(nv) if mjm.dof_parentid[i] == -1]
  if sparse:
    upd, ddepth = {}, np.zeros(nv, dtype=np.int64)
    for k in range(nv):
      ddepth[k] = 0 if mjm.dof_parentid[k] < 0 else ddepth[mjm.dof_parentid[k]] + 1
      i = mjm.dof_parentid[k]
      madr_ki = mjm.dof_Madr[k] + 1
      while i > -1:
        upd.setdefault(int(ddepth[i]), []).append((i, k, madr_ki))
        i = mjm.dof_parentid[i]
        madr_ki += 1
    if upd:
      t["qLD_update_tree"] = np.array([u for k in range(len(upd)) for u in upd[k]], np.int32).reshape(-1, 3)
      t["qLD_update_treeadr"] = np.cumsum([0] + [len(upd[k]) for k in range(len(upd))])[:-1].astype(np.int32)
  # one dense tile per kinematic tree, grouped by tile size (io.py:150-162).
  # The reference builds these for dense models only; here the kernels work
  # on the dense matrix for sparse models too, so every model has them.
  tiles = {}
  for n, beg in enumerate(roots):
    end = nv if n == len(roots) - 1 else roots[n + 1]
    tiles.setdefault(end - beg, []).append(beg)
  keys = sorted(tiles)
  if keys:
    t["qLD_tile"] = np.concatenate([tiles[k] for k in keys]).astype(np.int32)
    t["qLD_tileadr"] = np.cumsum([0] + [len(tiles[k]) for k in keys])[:-1].astype(np.int32)
    t["qLD_tilesize"] = np.array(keys, np.int32)

  # actuator_moment tiles: (tree dofs, tree actuators) (io.py:164-204)
  for k in ("offset_nv", "offset_nu", "tileadr", "tilesize_nv", "tilesize_nu"):
    t["actuator_moment_" + k] = np.zeros(0, np.int32)
  if roots:  # (as the dense tiles above: every model)
    tree_id = mjm.dof_treeid[roots]
    num_trees = int(np.max(tree_id))
    act_tree = mjm.body_treeid[mjm.jnt_bodyid[mjm.actuator_trnid[:, 0]]] if mjm.nu else np.zeros(0)
    counts = np.array([int(np.sum(act_tree == k)) for k in range(num_trees + 1)])
    groups = {}
    act_beg = 0
    for n, beg in enumerate(roots):
      end = nv if n == len(roots) - 1 else roots[n + 1]
      na = int(counts[int(tree_id[n])])
      groups.setdefault((end - beg, na), []).append((beg, act_beg))
      act_beg += na
    keys = sorted(groups)
    t["actuator_moment_offset_nv"] = np.array([g[0] for k in keys for g in groups[k]], np.int32)
    t["actuator_moment_offset_nu"] = np.array([g[1] for k in keys for g in groups[k]], np.int32)
    t["actuator_moment_tileadr"] = np.cumsum([0] + [len(groups[k]) for k in keys])[:-1].astype(np.int32)
    t["actuator_moment_tilesize_nv"] = np.array([k[0] for k in keys], np.int32)
    t["actuator_moment_tilesize_nu"] = np.array([k[1] for k in keys], np.int32)

  t["alpha_candidate"] = np.linspace(0.0, 1.0, mjm.opt.ls_iterations).astype(np.float32)

  # subtree mass; the world body is added to itself once more (io.py:249-252)
  sm = np.array(mjm.body_mass, dtype=np.float64).copy()
  for i in range(nbody - 1, -1, -1):
    sm[mjm.body_parentid[i]] += sm[i]
  t["subtree_mass"] = sm

  # body_dof_tree[b, i]: dof i's body is an ancestor-or-self of body b != 0
  bdt = np.zeros((nbody, nv), dtype=bool)
  for b in range(1, nbody):
    anc = set()
    c = b
    while c != 0:
      anc.add(c)
      c = mjm.body_parentid[c]
    for i in range(nv):
      bdt[b, i] = mjm.dof_bodyid[i] in anc
  t["body_dof_tree"] = bdt
  t["nxn_pair"] = nxn_pairs(mjm)
  return t


NXN_TABLE_MAX = 100


def nxn_pairs(mjm) -> np.ndarray:
  """The nxn broadphase's candidate pairs (g1 < g2, lexicographic as
  collision_driver.py:465-478) that pass the model-only part of the geom
  filter (collision_driver.py:27-49: contype/conaffinity mask, distinct
  weld bodies), packed g1 | g2 << 15 | parent_child << 30; parent_child
  marks the pairs the parent filter drops unless it is disabled.  Empty for
  models over NXN_TABLE_MAX geoms (the kernels then enumerate all pairs)."""
  ng = int(mjm.ngeom)
  if ng > NXN_TABLE_MAX:
    return np.zeros(0, np.int32)
  bodyid = np.asarray(mjm.geom_bodyid)
  contype = np.asarray(mjm.geom_contype)
  conaff = np.asarray(mjm.geom_conaffinity)
  weld = np.asarray(mjm.body_weldid)
  parent = np.asarray(mjm.body_parentid)
  out = []
  for g1 in range(ng):
    for g2 in range(g1 + 1, ng):
      w1, w2 = int(weld[bodyid[g1]]), int(weld[bodyid[g2]])
      wp1, wp2 = int(weld[parent[w1]]), int(weld[parent[w2]])
      mask = (int(contype[g1]) & int(conaff[g2])) or (int(contype[g2]) & int(conaff[g1]))
      if not mask or w1 == w2:
        continue
      pc = w1 != 0 and w2 != 0 and (w1 == wp2 or w2 == wp1)
      out.append(g1 | (g2 << 15) | (int(pc) << 30))
  return np.asarray(out, np.int32)


def put_model(mjm, device=None) -> types.Model:
  """Builds the device Model from a compiled MjModel (io.py:24-334)."""
  if getattr(mjm, "neq", 0) > 0:
    raise NotImplementedError("Equality constraints are unsupported.")
  if getattr(mjm, "nsensor", 0) > 0:
    raise NotImplementedError("Sensors are unsupported.")
  if getattr(mjm, "ntendon", 0) > 0:
    raise NotImplementedError("Tendons are unsupported.")
  if mjm.opt.integrator not in _SUPPORTED_INTEGRATORS:
    raise NotImplementedError(f"Integrator: {mjm.opt.integrator} is unsupported.")
  if mjm.opt.cone not in _SUPPORTED_CONES:
    raise NotImplementedError(f"Cone: {mjm.opt.cone} is unsupported.")
  if mjm.opt.solver not in _SUPPORTED_SOLVERS:
    raise NotImplementedError(f"Solver: {mjm.opt.solver} is unsupported.")

  device = device or default_device()
  tables = model_tables(mjm)
  m = types.Model()
  for k in ("nq", "nv", "nu", "na", "nbody", "njnt", "ngeom", "nsite", "nmocap", "nM"):
    setattr(m, k, int(getattr(mjm, k)))
  m.nexclude = int(getattr(mjm, "nexclude", 0))
  m.nlsp = int(mjm.opt.ls_iterations)
  m.nmesh = int(getattr(mjm, "nmesh", 0))
  m.nmeshvert = int(np.asarray(getattr(mjm, "mesh_vert", np.zeros((0, 3)))).reshape(-1, 3).shape[0])
  m.nlevel = len(tables["body_treeadr"])
  m.nlimited = len(tables["jnt_limited_slide_hinge_adr"])
  m.ntril = len(tables["dof_tri_row"])
  m.nfullm = len(tables["qM_fullm_i"])
  m.nmulm = len(tables["qM_mulm_i"])
  m.nqLD_update = len(tables["qLD_update_tree"])
  m.nqLD_update_level = len(tables["qLD_update_treeadr"])
  m.nqLD_tile = len(tables["qLD_tile"])
  m.nqLD_group = len(tables["qLD_tileadr"])
  m.nact_tile = len(tables["actuator_moment_offset_nv"])
  m.nact_group = len(tables["actuator_moment_tileadr"])
  m.max_level_width = int(tables["max_level_width"])
  m.nnxn = len(tables["nxn_pair"])
  # dofs of the largest kinematic tree: the LDS a tree block's factor takes
  m.tree_dof_max = int(np.max(tables["qLD_tilesize"])) if len(tables["qLD_tilesize"]) else 0
  m.actuator_affine_bias_gain = int(
    bool(np.any(np.asarray(mjm.actuator_biastype) == 1) or np.any(np.asarray(mjm.actuator_gaintype) == 1)))

  o = m.opt
  o.timestep = float(mjm.opt.timestep)
  o.impratio = float(mjm.opt.impratio)
  o.tolerance = float(mjm.opt.tolerance)
  o.ls_tolerance = float(mjm.opt.ls_tolerance)
  o.gravity = np.array(mjm.opt.gravity, dtype=np.float32)
  o.integrator = int(mjm.opt.integrator)
  o.cone = int(mjm.opt.cone)
  o.solver = int(mjm.opt.solver)
  o.iterations = int(mjm.opt.iterations)
  o.ls_iterations = int(mjm.opt.ls_iterations)
  o.disableflags = int(mjm.opt.disableflags)
  o.is_sparse = is_sparse(mjm)
  o.ls_parallel = False
  m.stat.meaninertia = float(mjm.stat.meaninertia)

  _, arrays = _native.layout()["model"]
  host = model_arrays(mjm, _native.model_sizes(m), tables)
  for kind, name, _ in arrays:
    setattr(m, name, _tensor(host[name], kind, device))
  return m


def model_arrays(mjm, sizes: dict, tables: dict | None = None) -> dict:
  """Every Model array of the C ABI as a numpy array in the MjModel's own
  precision, shaped by the field list (the input of put_model)."""
  tables = tables if tables is not None else model_tables(mjm)
  _, arrays = _native.layout()["model"]
  out = {}
  for _, name, dims in arrays:
    src = tables[name] if name in tables else getattr(mjm, name)
    out[name] = np.asarray(src).reshape(_native.shape_of(dims, sizes))
  return out


def _alloc_data(mjm, nworld, nconmax, njmax, device) -> types.Data:
  torch = _torch()
  d = types.Data()
  d.nworld = int(nworld)
  d.nconmax = int(nconmax)
  d.njmax = int(njmax)
  sparse = is_sparse(mjm)
  d.qM_rows, d.qM_cols = (1, mjm.nM) if sparse else (mjm.nv, mjm.nv)
  # dense working copies of qM / qLD: separate arrays for sparse models only
  d.qMd_n = mjm.nv if sparse else 0
  # device buffer for the fused step's model / data description
  d.ndesc = _native.step_desc_words()
  m_like = types.Model()
  tables = model_tables(mjm)
  for k in ("nq", "nv", "nu", "na", "nbody", "njnt", "ngeom", "nsite", "nmocap", "nM"):
    setattr(m_like, k, int(getattr(mjm, k)))
  sizes = dict(nworld=d.nworld, nconmax=d.nconmax, njmax=d.njmax, qM_rows=d.qM_rows, qM_cols=d.qM_cols,
               qMd_n=d.qMd_n, ndesc=d.ndesc, nworld2=2 * d.nworld, nworld1=d.nworld + 1, nworld_ngeom=d.nworld * mjm.ngeom,
               nq=mjm.nq, nv=mjm.nv, nu=mjm.nu, na=mjm.na, nbody=mjm.nbody, njnt=mjm.njnt, ngeom=mjm.ngeom,
               nsite=mjm.nsite, nmocap=mjm.nmocap, nM=mjm.nM, nlsp=int(mjm.opt.ls_iterations))
  _, arrays = _native.layout()["data"]
  dt = {"R": torch.float32, "I": torch.int32, "B": torch.bool}
  for kind, name, dims in arrays:
    node, attr = _native.data_node(d, name)
    setattr(node, attr, torch.zeros(_native.shape_of(dims, sizes), dtype=dt[kind], device=device))
  d.sap_segment_index = torch.arange(0, (d.nworld + 1) * mjm.ngeom, max(mjm.ngeom, 1), dtype=torch.int32,
                                     device=device)[: d.nworld + 1].contiguous()
  del tables
  return d


def make_data(mjm, nworld: int = 1, nconmax: int = -1, njmax: int = -1, device=None) -> types.Data:
  """Zero-initialised Data with qpos = qpos0 (io.py:395-506)."""
  device = device or default_device()
  if nconmax == -1:
    nconmax = 512
  if njmax == -1:
    njmax = 512
  d = _alloc_data(mjm, nworld, nconmax, njmax, device)
  torch = _torch()
  d.qpos = torch.from_numpy(np.tile(np.asarray(mjm.qpos0, np.float32), (nworld, 1))).to(device).contiguous()
  if mjm.nmocap:
    mb = [b for b in range(mjm.nbody) if mjm.body_mocapid[b] >= 0]
    d.mocap_pos = _tensor(np.tile(np.asarray(mjm.body_pos)[mb], (nworld, 1, 1)), "R", device)
    d.mocap_quat = _tensor(np.tile(np.asarray(mjm.body_quat)[mb], (nworld, 1, 1)), "R", device)
  return d


def _dense_qM(mjm, qM_sparse):
  """MuJoCo's sparse qM (dof_Madr layout) -> dense symmetric matrix."""
  nv = mjm.nv
  M = np.zeros((nv, nv))
  for i in range(nv):
    adr = mjm.dof_Madr[i]
    j = i
    while j >= 0:
      M[i, j] = M[j, i] = qM_sparse[adr]
      adr += 1
      j = mjm.dof_parentid[j]
  return M


def put_data(mjm, mjd, nworld: int = 1, nconmax: int = -1, njmax: int = -1, device=None) -> types.Data:
  """Tiles one host MjData across nworld worlds (io.py:509-725)."""
  device = device or default_device()
  ncon = int(getattr(mjd, "ncon", 0))
  nefc = int(getattr(mjd, "nefc", 0))
  if nconmax == -1:
    nconmax = max(512, ncon * nworld)
  if njmax == -1:
    njmax = max(512, nefc * nworld)
  if nworld * nefc > njmax:
    raise ValueError("nworld * nefc > njmax")
  if nworld * ncon > nconmax:
    raise ValueError("nworld * ncon > nconmax")
  d = make_data(mjm, nworld, nconmax, njmax, device)
  d.time = float(getattr(mjd, "time", 0.0))
  d.nl = int(getattr(mjd, "nl", 0))

  def tile(name, kind="R", shape=None):
    v = getattr(mjd, name, None)
    if v is None:
      return None
    v = np.asarray(v)
    if shape is not None:
      v = v.reshape(shape)
    return _tensor(np.tile(v, (nworld,) + (1,) * v.ndim), kind, device)

  nb, nj, ng, ns = mjm.nbody, mjm.njnt, mjm.ngeom, mjm.nsite
  simple = {
    "qpos": None, "qvel": None, "qacc_warmstart": None, "qfrc_applied": None, "qacc": None, "ctrl": None,
    "act": None, "act_dot": None, "xfrc_applied": (nb, 6), "mocap_pos": (mjm.nmocap, 3),
    "mocap_quat": (mjm.nmocap, 4), "xanchor": (nj, 3), "xaxis": (nj, 3), "xmat": (nb, 3, 3), "xpos": (nb, 3),
    "xquat": (nb, 4), "xipos": (nb, 3), "ximat": (nb, 3, 3), "subtree_com": (nb, 3), "geom_xpos": (ng, 3),
    "geom_xmat": (ng, 3, 3), "site_xpos": (ns, 3), "site_xmat": (ns, 3, 3), "cinert": (nb, 10),
    "cdof": (mjm.nv, 6), "crb": (nb, 10), "qLDiagInv": None, "actuator_velocity": None, "actuator_force": None,
    "actuator_length": None, "cvel": (nb, 6), "cdof_dot": (mjm.nv, 6), "qfrc_bias": None, "qfrc_passive": None,
    "qfrc_spring": None, "qfrc_damper": None, "qfrc_actuator": None, "qfrc_smooth": None,
    "qfrc_constraint": None, "qacc_smooth": None,
  }
  for name, shape in simple.items():
    t = tile(name, "R", shape)
    if t is not None:
      setattr(d, name, t.reshape(getattr(d, name).shape).contiguous())
  qM = getattr(mjd, "qM", None)
  if qM is not None and np.asarray(qM).size == mjm.nM and mjm.nv:
    # the host qM is MuJoCo's sparse layout; dense models keep it dense,
    # sparse models keep both the sparse layout and the dense working copy
    M = _dense_qM(mjm, np.asarray(qM))
    dense = _tensor(np.tile(M, (nworld, 1, 1)), "R", device)
    try:
      dense_L = _tensor(np.tile(np.linalg.cholesky(M), (nworld, 1, 1)), "R", device)
    except np.linalg.LinAlgError:
      dense_L = None
    if is_sparse(mjm):
      d.qM = _tensor(np.tile(np.asarray(qM, np.float64).reshape(1, mjm.nM), (nworld, 1, 1)), "R", device)
      d.qMd = dense
      if dense_L is not None:
        d.qLDd = dense_L
      qLD = getattr(mjd, "qLD", None)
      if qLD is not None and np.asarray(qLD).size == mjm.nM:
        d.qLD = _tensor(np.tile(np.asarray(qLD, np.float64).reshape(1, mjm.nM), (nworld, 1, 1)), "R", device)
    else:
      d.qM = dense
      if dense_L is not None:
        d.qLD = dense_L
  mom = getattr(mjd, "actuator_moment", None)
  if mom is not None and np.asarray(mom).size == mjm.nu * mjm.nv:
    d.actuator_moment = _tensor(np.tile(np.asarray(mom).reshape(mjm.nu, mjm.nv), (nworld, 1, 1)), "R", device)

  # contact and constraint pools: world i holds entries [i*n, (i+1)*n)
  torch = _torch()
  if ncon:
    c = mjd.contact
    for name, width in (("dist", None), ("pos", 3), ("frame", 9), ("includemargin", None), ("friction", 5),
                        ("solref", 2), ("solreffriction", 2), ("solimp", 5)):
      v = np.asarray(getattr(c, name), np.float32).reshape(ncon, -1)
      full = getattr(d.contact, name).reshape(nconmax, -1).cpu().numpy()
      full[: nworld * ncon] = np.tile(v, (nworld, 1))
      setattr(d.contact, name, _tensor(full.reshape(getattr(d.contact, name).shape), "R", device))
    for name in ("dim", "geom", "efc_address"):
      v = np.asarray(getattr(c, name), np.int32).reshape(ncon, -1)
      full = getattr(d.contact, name).reshape(nconmax, -1).cpu().numpy()
      rep = np.tile(v, (nworld, 1))
      if name == "efc_address":
        rep = rep + (np.repeat(np.arange(nworld), ncon) * ncon)[:, None]
      full[: nworld * ncon] = rep
      setattr(d.contact, name, _tensor(full.reshape(getattr(d.contact, name).shape), "I", device))
    wid = np.zeros(nconmax, np.int32)
    wid[: nworld * ncon] = np.repeat(np.arange(nworld), ncon)
    d.contact.worldid = _tensor(wid, "I", device)
    d.ncon = torch.tensor([ncon * nworld], dtype=torch.int32, device=device)
  wc = np.stack([np.arange(nworld) * ncon, np.full(nworld, ncon)], axis=1).astype(np.int32)
  d.world_con = _tensor(wc, "I", device)
  if nefc:
    J = np.asarray(mjd.efc_J, np.float32).reshape(nefc, mjm.nv)
    full = np.zeros((njmax, mjm.nv), np.float32)
    full[: nworld * nefc] = np.tile(J, (nworld, 1))
    d.efc.J = _tensor(full, "R", device)
    for name in ("D", "pos", "aref", "force", "margin"):
      v = np.zeros(njmax, np.float32)
      v[: nworld * nefc] = np.tile(np.asarray(getattr(mjd, "efc_" + name), np.float32), nworld)
      setattr(d.efc, name, _tensor(v, "R", device))
    wid = np.zeros(njmax, np.int32)
    wid[: nworld * nefc] = np.repeat(np.arange(nworld), nefc)
    d.efc.worldid = _tensor(wid, "I", device)
    d.nefc = torch.tensor([nefc * nworld], dtype=torch.int32, device=device)
  we = np.stack([np.arange(nworld) * nefc, np.full(nworld, nefc)], axis=1).astype(np.int32)
  d.world_efc = _tensor(we, "I", device)
  return d


def get_data_into(result, mjm, d: types.Data):
  """Copies world 0 of d into a host MjData (io.py:728-831)."""
  if d.nworld > 1:
    raise NotImplementedError("only nworld == 1 supported for now")
  ncon = int(d.ncon.cpu()[0])
  nefc = int(d.nefc.cpu()[0])
  result.time = d.time
  for name in ("qpos", "qvel", "qacc_warmstart", "qfrc_applied", "mocap_pos", "mocap_quat", "qacc", "xanchor",
               "xaxis", "xpos", "xquat", "xipos", "subtree_com", "geom_xpos", "site_xpos", "cinert", "cdof",
               "crb", "qLDiagInv", "ctrl", "actuator_velocity", "actuator_force", "actuator_length", "cvel",
               "cdof_dot", "qfrc_bias", "qfrc_passive", "qfrc_spring", "qfrc_damper", "qfrc_actuator",
               "qfrc_smooth", "qfrc_constraint", "qacc_smooth", "act", "act_dot"):
    setattr(result, name, getattr(d, name).cpu().numpy()[0].astype(np.float64))
  for name in ("xmat", "ximat", "geom_xmat", "site_xmat"):
    setattr(result, name, getattr(d, name).cpu().numpy()[0].reshape(-1, 9).astype(np.float64))
  result.actuator_moment = d.actuator_moment.cpu().numpy()[0].astype(np.float64)
  if is_sparse(mjm):
    # io.py:803-810: the sparse layout is MuJoCo's own
    result.qM = d.qM.cpu().numpy()[0, 0].astype(np.float64)
    result.qLD = d.qLD.cpu().numpy()[0, 0].astype(np.float64)
  else:
    qM = d.qM.cpu().numpy()[0]
    qLD = d.qLD.cpu().numpy()[0]
    sq, sl = [], []
    for i in range(mjm.nv):
      j = i
      while j >= 0:
        sq.append(qM[i, j])
        sl.append(qLD[i, j])
        j = mjm.dof_parentid[j]
    result.qM = np.array(sq)
    result.qLD = np.array(sl)
  result.ncon = ncon
  result.nefc = nefc

  class _C:
    pass

  c = _C()
  for name in ("dist", "pos", "frame", "includemargin", "friction", "solref", "solreffriction", "solimp", "dim",
               "geom", "efc_address"):
    v = getattr(d.contact, name).cpu().numpy()[:ncon]
    setattr(c, name, v.reshape(ncon, int(np.prod(v.shape[1:]))) if v.ndim > 1 else v)
  result.contact = c
  for name in ("D", "pos", "aref", "force", "margin"):
    setattr(result, "efc_" + name, getattr(d.efc, name).cpu().numpy()[:nefc].astype(np.float64))
  result.efc_J = d.efc.J.cpu().numpy()[:nefc].astype(np.float64)
