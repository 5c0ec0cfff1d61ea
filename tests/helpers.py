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
  sizes = dict(nworld=d.nworld, nconmax=d.nconmax, njmax=d.njmax, qM_rows=d.qM_rows, qM_cols=d.qM_cols)
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
