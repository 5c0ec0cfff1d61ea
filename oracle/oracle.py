"""ctypes front end of liboracle.so — test infrastructure only.

The oracle runs on host numpy arrays with the same field names and layout as
the device Model / Data (the X-macro lists of include/mjx_warp_amd.h), in
float32 (parity with the kernels) or float64 (accuracy anchor).  Arrays are
updated in place.
"""

from __future__ import annotations

import ctypes
import functools
import re
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
HEADER = HERE.parent / "include" / "mjx_warp_amd.h"
LIBRARY = HERE / "liboracle.so"

STAGES = ("kinematics", "com_pos", "crb", "factor_m", "rne", "transmission", "com_vel", "passive",
          "nxn_broadphase", "sap_broadphase", "primitive_narrowphase", "collision", "make_constraint",
          "fwd_position", "fwd_velocity", "fwd_actuation", "fwd_acceleration", "solve", "euler", "implicit",
          "forward", "step")


def _macro(text, name):
  m = re.search(r"#define\s+" + name + r"\(X\)((?:.*\\\n)*.*)\n", text)
  return m.group(1)


@functools.lru_cache(maxsize=None)
def layout():
  text = HEADER.read_text()
  out = {}
  for s in ("MODEL", "DATA"):
    sizes = re.findall(r"X\((\w+)\)", _macro(text, f"MJH_{s}_SIZES"))
    arrays = [(k, n, tuple(x.strip() for x in dims.split(",")))
              for k, n, dims in re.findall(r"X\(([RIB]),\s*(\w+),\s*\(([^)]*)\)\)", _macro(text, f"MJH_{s}_ARRAYS"))]
    out[s.lower()] = (sizes, arrays)
  return out


@functools.lru_cache(maxsize=None)
def structs(real):
  R = ctypes.c_float if real == np.float32 else ctypes.c_double
  (msizes, marrays), (dsizes, darrays) = layout()["model"], layout()["data"]

  class Opt(ctypes.Structure):
    _fields_ = [("timestep", R), ("impratio", R), ("tolerance", R), ("ls_tolerance", R), ("gravity", R * 3),
                ("integrator", ctypes.c_int), ("cone", ctypes.c_int), ("solver", ctypes.c_int),
                ("iterations", ctypes.c_int), ("ls_iterations", ctypes.c_int), ("disableflags", ctypes.c_int),
                ("is_sparse", ctypes.c_int), ("ls_parallel", ctypes.c_int)]

  class Model(ctypes.Structure):
    _fields_ = ([(s, ctypes.c_int32) for s in msizes] + [("opt", Opt), ("meaninertia", R)]
                + [(n, ctypes.c_void_p) for _, n, _ in marrays])

  class Data(ctypes.Structure):
    _fields_ = [(s, ctypes.c_int32) for s in dsizes] + [(n, ctypes.c_void_p) for _, n, _ in darrays]

  return Model, Data


@functools.lru_cache(maxsize=None)
def lib():
  if not LIBRARY.exists():
    raise RuntimeError(f"{LIBRARY} is missing; run python oracle/build.py")
  L = ctypes.CDLL(str(LIBRARY))
  for s in STAGES:
    for suf in ("f32", "f64"):
      f = getattr(L, f"oracle_{s}_{suf}")
      f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
      f.restype = ctypes.c_int
  return L


_NP = {"I": np.int32, "B": np.uint8}


class HostModel:
  """Numpy copy of a model: sizes, options and arrays by flat field name."""

  def __init__(self, sizes: dict, opt: dict, meaninertia: float, arrays: dict, real=np.float64):
    self.real = real
    self.sizes = dict(sizes)
    self.opt = dict(opt)
    self.meaninertia = meaninertia
    _, marrays = layout()["model"]
    self.arrays = {}
    for kind, name, _ in marrays:
      dt = real if kind == "R" else _NP[kind]
      self.arrays[name] = np.ascontiguousarray(np.asarray(arrays[name]).astype(dt))

  def struct(self):
    Model, _ = structs(self.real)
    s = Model()
    for k, v in self.sizes.items():
      setattr(s, k, int(v))
    for k in ("timestep", "impratio", "tolerance", "ls_tolerance"):
      setattr(s.opt, k, float(self.opt[k]))
    for i in range(3):
      s.opt.gravity[i] = float(self.opt["gravity"][i])
    for k in ("integrator", "cone", "solver", "iterations", "ls_iterations", "disableflags", "is_sparse",
              "ls_parallel"):
      setattr(s.opt, k, int(self.opt[k]))
    s.meaninertia = self.meaninertia
    for name, a in self.arrays.items():
      setattr(s, name, a.ctypes.data if a.size else None)
    return s


class HostData:
  """Numpy copy of a data: sizes and arrays by flat field name."""

  def __init__(self, sizes: dict, arrays: dict, real=np.float64):
    self.real = real
    self.sizes = dict(sizes)
    _, darrays = layout()["data"]
    self.arrays = {}
    for kind, name, _ in darrays:
      dt = real if kind == "R" else _NP[kind]
      self.arrays[name] = np.ascontiguousarray(np.asarray(arrays[name]).astype(dt))

  def __getitem__(self, k):
    return self.arrays[k]

  def struct(self):
    _, Data = structs(self.real)
    s = Data()
    for k in ("nworld", "nconmax", "njmax", "qM_rows", "qM_cols", "qMd_n"):
      setattr(s, k, int(self.sizes[k]))
    for name, a in self.arrays.items():
      setattr(s, name, a.ctypes.data if a.size else None)
    if int(self.sizes["qMd_n"]) == 0:
      # dense model: the dense working matrices are qM and qLD themselves
      s.qMd, s.qLDd = s.qM, s.qLD
    return s


def run(stage: str, m: HostModel, d: HostData, nthreads: int = 1):
  """Runs one reference stage of the oracle on (m, d) in place."""
  if m.real != d.real:
    raise ValueError("model and data precision differ")
  suf = "f32" if m.real == np.float32 else "f64"
  ms, ds = m.struct(), d.struct()
  if stage in ("step", "implicit") and ms.opt.is_sparse and ms.opt.integrator == 3:
    # implicitfast works on the dense matrices only; the reference refuses
    # a sparse model (forward.py:268)
    raise NotImplementedError("implicitfast with a sparse mass matrix is not implemented")
  rc = getattr(lib(), f"oracle_{stage}_{suf}")(ctypes.byref(ms), ctypes.byref(ds), int(nthreads))
  if rc != 0:
    raise RuntimeError(f"oracle {stage} failed: {rc}")


def closest_segments(a0, a1, b0, b1, real=np.float64):
  """math.py:209-242 closest points between two segments."""
  L = lib()
  suf = "f32" if real == np.float32 else "f64"
  f = getattr(L, f"oracle_closest_segments_{suf}")
  args = [np.ascontiguousarray(np.asarray(x, dtype=real)) for x in (a0, a1, b0, b1)]
  ra = np.zeros(3, real)
  rb = np.zeros(3, real)
  f(*[a.ctypes.data_as(ctypes.c_void_p) for a in args], ra.ctypes.data_as(ctypes.c_void_p),
    rb.ctypes.data_as(ctypes.c_void_p))
  return ra, rb
