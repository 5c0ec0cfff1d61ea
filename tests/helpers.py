"""Shared helpers of the test suite: model / data construction and the
conversion of device Model / Data into the oracle's host arrays."""

from __future__ import annotations

import numpy as np
import torch

from mjx_warp_amd import _native
from mjx_warp_amd import io
from oracle import oracle


def host_model(m, real=np.float64) -> oracle.HostModel:
  msizes, marrays = _native.layout()["model"]
  sizes = {s: int(getattr(m, s)) for s in msizes}
  o = m.opt
  opt = dict(timestep=o.timestep, impratio=o.impratio, tolerance=o.tolerance, ls_tolerance=o.ls_tolerance,
             gravity=np.asarray(o.gravity, dtype=np.float64), integrator=o.integrator, cone=o.cone,
             solver=o.solver, iterations=o.iterations, ls_iterations=o.ls_iterations,
             disableflags=o.disableflags, is_sparse=int(bool(o.is_sparse)), ls_parallel=int(bool(o.ls_parallel)))
  arrays = {name: getattr(m, name).detach().cpu().numpy() for _, name, _ in marrays}
  return oracle.HostModel(sizes, opt, float(m.stat.meaninertia), arrays, real)


def host_data(m, d, real=np.float64) -> oracle.HostData:
  _, darrays = _native.layout()["data"]
  sizes = dict(nworld=d.nworld, nconmax=d.nconmax, njmax=d.njmax, qM_rows=d.qM_rows, qM_cols=d.qM_cols,
               qMd_n=d.qMd_n)
  arrays = {}
  for _, name, _ in darrays:
    node, attr = _native.data_node(d, name)
    arrays[name] = getattr(node, attr).detach().cpu().numpy()
  return oracle.HostData(sizes, arrays, real)


def data_numpy(d, name):
  node, attr = _native.data_node(d, name)
  return getattr(node, attr).detach().cpu().numpy()


def set_field(d, name, value):
  node, attr = _native.data_node(d, name)
  t = getattr(node, attr)
  setattr(node, attr, torch.as_tensor(np.asarray(value), dtype=t.dtype, device=t.device).reshape(t.shape).contiguous())


def world_blocks(pool_adr_num, nworld):
  """(first, count) of each world's block from a world_* array."""
  a = np.asarray(pool_adr_num).reshape(nworld, 2)
  return [(int(a[w, 0]), int(a[w, 1])) for w in range(nworld)]


def gather_world_rows(arr, blocks):
  """Concatenates each world's block of a pool array in world order."""
  return np.concatenate([arr[b:b + n] for b, n in blocks], axis=0) if blocks else arr[:0]


def load(xml_or_mjm, nworld=1, keyframe=None, device="cpu", nconmax=None, njmax=None, real=np.float64):
  """(mjm, m, d, host model, host data) for an MJCF string or compiled model."""
  from mjx_warp_amd import mjcf

  mjm = mjcf.from_xml_string(xml_or_mjm) if isinstance(xml_or_mjm, str) else xml_or_mjm
  m = io.put_model(mjm, device=device)
  d = io.make_data(mjm, nworld=nworld, nconmax=nconmax or max(64, 16 * nworld),
                   njmax=njmax or max(256, 64 * nworld), device=device)
  hm = host_model_exact(mjm, m, real)
  hd = host_data(m, d, real)
  hd.arrays["qpos"][:] = mjm.qpos0[None]
  if keyframe is not None:
    set_field(d, "qpos", np.tile(mjm.key_qpos[keyframe], (nworld, 1)))
    set_field(d, "qvel", np.tile(mjm.key_qvel[keyframe], (nworld, 1)))
    hd.arrays["qpos"][:] = mjm.key_qpos[keyframe][None]
    hd.arrays["qvel"][:] = mjm.key_qvel[keyframe][None]
  return mjm, m, d, hm, hd


def host_model_exact(mjm, m, real=np.float64) -> oracle.HostModel:
  """Oracle model whose arrays come straight from the compiled MjModel (no
  float32 round trip), so float64 checks see the model's exact values."""
  base = host_model(m, real)
  arrays = io.model_arrays(mjm, base.sizes)
  o = mjm.opt
  opt = dict(base.opt, timestep=o.timestep, impratio=o.impratio, tolerance=o.tolerance,
             ls_tolerance=o.ls_tolerance, gravity=np.asarray(o.gravity, np.float64))
  return oracle.HostModel(base.sizes, opt, float(mjm.stat.meaninertia), arrays, real)


def pairs_of(hd, world=0):
  """Sorted (g1, g2) pairs of one world from the broadphase pool."""
  adr, n = hd["world_pair"].reshape(-1, 2)[world]
  p = hd["collision_pair"].reshape(-1, 2)[adr:adr + n]
  return sorted(tuple(sorted((int(a), int(b)))) for a, b in p)


def oracle_stepper(mjm, m_cpu, real=np.float64, nconmax_w=64, njmax_w=277, nthreads=1):
  """step(qpos, qvel, act, ctrl, warm) -> (qpos', qvel', act', warm') over
  [n, .] arrays, one oracle step of n worlds (host data cached per n) on
  nthreads host threads."""
  hm = host_model_exact(mjm, m_cpu, real) if real == np.float64 else host_model(m_cpu, real)
  cache = {}

  def step(q, v, a, u, w):
    n = q.shape[0]
    if n not in cache:
      d = io.make_data(mjm, nworld=n, nconmax=nconmax_w * n, njmax=njmax_w * n, device="cpu")
      cache[n] = host_data(m_cpu, d, real)
    hd = cache[n]
    hd["qpos"][:] = q
    hd["qvel"][:] = v
    if mjm.na:
      hd["act"][:] = a
    if mjm.nu:
      hd["ctrl"][:] = u
    hd["qacc_warmstart"][:] = w
    oracle.run("step", hm, hd, nthreads=nthreads)
    act = hd["act"].copy() if mjm.na else np.zeros((n, 0))
    return (hd["qpos"].astype(np.float64), hd["qvel"].astype(np.float64), act.astype(np.float64),
            hd["qacc_warmstart"].astype(np.float64))

  return step


def mjdata_from_oracle(mjm, hd, w=0):
  """A host MjData (MuJoCo's field names and host layouts) holding world w
  of an oracle HostData after a forward pass: the state, the derived
  fields, qM in MuJoCo's sparse layout, the world's contacts and
  constraint rows -- what mj_forward leaves in an mjData, the input of the
  reference's put_data (test_util.py:31-44)."""
  from mjx_warp_amd import mjcf

  mjd = mjcf.MjData(mjm)
  nb, nv = mjm.nbody, mjm.nv
  for name in ("qpos", "qvel", "qacc", "qacc_warmstart", "qfrc_applied", "ctrl", "act", "act_dot", "xpos",
               "xquat", "xipos", "xanchor", "xaxis", "subtree_com", "geom_xpos", "site_xpos", "cinert", "cdof",
               "crb", "cvel", "cdof_dot", "actuator_length", "actuator_velocity", "actuator_force", "qfrc_bias",
               "qfrc_spring", "qfrc_damper", "qfrc_passive", "qfrc_actuator", "qfrc_smooth", "qfrc_constraint",
               "qacc_smooth", "xfrc_applied", "mocap_pos", "mocap_quat", "qLDiagInv"):
    setattr(mjd, name, np.asarray(hd[name][w], np.float64).copy())
  for name in ("xmat", "ximat", "geom_xmat", "site_xmat"):
    setattr(mjd, name, np.asarray(hd[name][w], np.float64).reshape(-1, 9).copy())
  mjd.actuator_moment = np.asarray(hd["actuator_moment"][w], np.float64).copy()
  # dense M -> MuJoCo's sparse qM (row i: i, then its ancestors' columns)
  M = np.asarray(hd["qMd"][w] if hd.sizes["qMd_n"] else hd["qM"][w], np.float64)
  sq = []
  for i in range(nv):
    j = i
    while j >= 0:
      sq.append(M[i, j])
      j = mjm.dof_parentid[j]
  mjd.qM = np.array(sq)
  c0, nc = (int(x) for x in hd["world_con"].reshape(-1, 2)[w])

  class _Contact:
    pass

  c = _Contact()
  for name, kind in (("dist", 1), ("pos", 3), ("frame", 9), ("includemargin", 1), ("friction", 5), ("solref", 2),
                     ("solreffriction", 2), ("solimp", 5), ("dim", 1), ("geom", 2), ("efc_address", 1)):
    v = np.asarray(hd["contact_" + name])
    v = v.reshape(v.shape[0], -1)[c0:c0 + nc]
    setattr(c, name, v if kind > 1 else v[:, 0])
  mjd.contact = c
  mjd.ncon = nc
  r0, nr = (int(x) for x in hd["world_efc"].reshape(-1, 2)[w])
  mjd.nefc = nr
  mjd.efc_J = np.asarray(hd["efc_J"][r0:r0 + nr], np.float64).copy()
  for name in ("D", "pos", "aref", "force", "margin"):
    setattr(mjd, "efc_" + name, np.asarray(hd["efc_" + name][r0:r0 + nr], np.float64).copy())
  return mjd
