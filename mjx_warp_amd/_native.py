"""ctypes binding of libmjx_warp_amd.so.

The C structs mjhModel / mjhData are generated here from the X-macro field
lists of include/mjx_warp_amd.h, so the Python layout cannot drift from the
C one.  `call()` marshals a Model / Data pair (torch tensors on the HIP
device) into those structs and invokes an entry point on torch's current
stream.  There is no fallback: if the library is missing or a tensor is not
on the GPU, the call raises.
"""

from __future__ import annotations

import ctypes
import functools
import math
import re
from pathlib import Path

_PKG = Path(__file__).resolve().parent
HEADER = _PKG.parent / "include" / "mjx_warp_amd.h"
LIBRARY = _PKG / "libmjx_warp_amd.so"

MJH_OK = 0
MJH_ERR_INVALID = -1
MJH_ERR_UNSUPPORTED = -2
MJH_ERR_HIP = -3
MJH_ERR_CAPACITY = -4


def _macro(text: str, name: str) -> str:
  m = re.search(r"#define\s+" + name + r"\(X\)((?:.*\\\n)*.*)\n", text)
  if not m:
    raise RuntimeError(f"{name} not found in {HEADER}")
  return m.group(1)


@functools.lru_cache(maxsize=None)
def layout():
  """Parses the field lists: sizes and (kind, name, dims) per struct."""
  text = HEADER.read_text()
  out = {}
  for struct in ("MODEL", "DATA"):
    sizes = re.findall(r"X\((\w+)\)", _macro(text, f"MJH_{struct}_SIZES"))
    arrays = []
    for kind, name, dims in re.findall(r"X\(([RIB]),\s*(\w+),\s*\(([^)]*)\)\)", _macro(text, f"MJH_{struct}_ARRAYS")):
      arrays.append((kind, name, tuple(x.strip() for x in dims.split(","))))
    out[struct.lower()] = (sizes, arrays)
  entry = re.findall(r"^int\s+(mjh_\w+)\(", text, flags=re.M)
  out["entry_points"] = entry + ["mjh_last_error"]
  return out


class OptionStruct(ctypes.Structure):
  _fields_ = [
    ("timestep", ctypes.c_float), ("impratio", ctypes.c_float), ("tolerance", ctypes.c_float),
    ("ls_tolerance", ctypes.c_float), ("gravity", ctypes.c_float * 3), ("integrator", ctypes.c_int),
    ("cone", ctypes.c_int), ("solver", ctypes.c_int), ("iterations", ctypes.c_int),
    ("ls_iterations", ctypes.c_int), ("disableflags", ctypes.c_int), ("is_sparse", ctypes.c_int),
    ("ls_parallel", ctypes.c_int),
  ]


@functools.lru_cache(maxsize=None)
def structs():
  lay = layout()
  msizes, marrays = lay["model"]
  dsizes, darrays = lay["data"]

  class ModelStruct(ctypes.Structure):
    _fields_ = ([(s, ctypes.c_int32) for s in msizes] + [("opt", OptionStruct), ("meaninertia", ctypes.c_float)]
                + [(n, ctypes.c_void_p) for _, n, _ in marrays])

  class DataStruct(ctypes.Structure):
    _fields_ = [(s, ctypes.c_int32) for s in dsizes] + [(n, ctypes.c_void_p) for _, n, _ in darrays]

  return ModelStruct, DataStruct


def data_node(d, name: str):
  """(object, attribute) holding flat field `name` of a Data."""
  if name.startswith("contact_"):
    return d.contact, name[len("contact_"):]
  if name.startswith("efc_"):
    return d.efc, name[len("efc_"):]
  return d, name


@functools.lru_cache(maxsize=None)
def step_desc_words() -> int:
  """MJH_STEP_DESC_WORDS: the size of Data.step_desc."""
  return int(re.search(r"#define\s+MJH_STEP_DESC_WORDS\s+(\d+)", HEADER.read_text()).group(1))


class GroupTuneStruct(ctypes.Structure):
  """mjhGroupTune: lanes per world (0, 32, 64), LDS-solver row capacity
  (<= 0: the library's), Jacobian rows staged in LDS (< 0: the library's)."""
  _fields_ = [("team", ctypes.c_int), ("rmax", ctypes.c_int), ("rj", ctypes.c_int)]


MJH_MULTI_BUCKETS = 1
MJH_MULTI_QUEUE = 2


def model_sizes(m) -> dict:
  msizes, _ = layout()["model"]
  return {s: int(getattr(m, s)) for s in msizes}


def data_sizes(m, d) -> dict:
  s = model_sizes(m)
  s.update(nworld=int(d.nworld), nconmax=int(d.nconmax), njmax=int(d.njmax), qM_rows=int(d.qM_rows),
           qM_cols=int(d.qM_cols), qMd_n=int(d.qMd_n), ndesc=int(getattr(d, "ndesc", step_desc_words())))
  s.update(nworld2=2 * s["nworld"], nworld1=s["nworld"] + 1, nworld_ngeom=s["nworld"] * s["ngeom"])
  return s


def shape_of(dims, sizes) -> tuple:
  return tuple(int(eval(x, {}, sizes)) for x in dims)


_KIND_DTYPE = {"R": "float32", "I": "int32", "B": "bool"}


def _ptr(t, name, kind, shape):
  import torch

  want = {"R": torch.float32, "I": torch.int32, "B": torch.bool}[kind]
  if not isinstance(t, torch.Tensor):
    raise TypeError(f"{name}: expected a torch.Tensor, got {type(t).__name__}")
  if t.dtype != want:
    raise ValueError(f"{name}: dtype {t.dtype}, expected {want}")
  if t.numel() != math.prod(shape):
    raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {shape}")
  if not t.is_contiguous():
    raise ValueError(f"{name}: tensor must be contiguous")
  if t.numel() and t.device.type != "cuda":
    raise ValueError(f"{name}: tensor is on {t.device}, the kernels need the HIP device")
  return t.data_ptr() if t.numel() else None


def model_struct(m):
  if m._abi is not None:
    return m._abi
  ModelStruct, _ = structs()
  _, arrays = layout()["model"]
  sizes = model_sizes(m)
  s = ModelStruct()
  for k, v in sizes.items():
    setattr(s, k, v)
  o = m.opt
  s.opt.timestep = o.timestep
  s.opt.impratio = o.impratio
  s.opt.tolerance = o.tolerance
  s.opt.ls_tolerance = o.ls_tolerance
  for i in range(3):
    s.opt.gravity[i] = float(o.gravity[i])
  s.opt.integrator = int(o.integrator)
  s.opt.cone = int(o.cone)
  s.opt.solver = int(o.solver)
  s.opt.iterations = int(o.iterations)
  s.opt.ls_iterations = int(o.ls_iterations)
  s.opt.disableflags = int(o.disableflags)
  s.opt.is_sparse = int(bool(o.is_sparse))
  s.opt.ls_parallel = int(bool(o.ls_parallel))
  s.meaninertia = float(m.stat.meaninertia)
  for kind, name, dims in arrays:
    setattr(s, name, _ptr(getattr(m, name), name, kind, shape_of(dims, sizes)))
  object.__setattr__(m, "_abi", s)
  return s


def data_struct(m, d):
  if d._abi is not None and d._abi_model is m:
    return d._abi
  _, DataStruct = structs()
  _, arrays = layout()["data"]
  sizes = data_sizes(m, d)
  if sizes["ndesc"] not in (0, step_desc_words()):
    # the step writes a whole description into a non-empty step_desc
    raise ValueError(f"step_desc must hold 0 or {step_desc_words()} words, not {sizes['ndesc']}")
  s = DataStruct()
  for k in ("nworld", "nconmax", "njmax", "qM_rows", "qM_cols", "qMd_n"):
    setattr(s, k, sizes[k])
  for kind, name, dims in arrays:
    node, attr = data_node(d, name)
    setattr(s, name, _ptr(getattr(node, attr), name, kind, shape_of(dims, sizes)))
  if sizes["qMd_n"] == 0:
    # dense model: the kernels' dense working matrices are qM and qLD
    s.qMd, s.qLDd = s.qM, s.qLD
  object.__setattr__(d, "_abi", s)
  object.__setattr__(d, "_abi_model", m)
  return s


@functools.lru_cache(maxsize=None)
def lib():
  """Loads the library; raises if it has not been built."""
  if not LIBRARY.exists():
    raise RuntimeError(f"{LIBRARY} is missing: build it with __graft_entry__.build() "
                       "(python -c 'import __graft_entry__ as g; g.build()')")
  L = ctypes.CDLL(str(LIBRARY))
  P = ctypes.c_void_p
  L.mjh_last_error.restype = ctypes.c_char_p
  L.mjh_abi_version.restype = ctypes.c_int
  for name in layout()["entry_points"]:
    if name in ("mjh_last_error", "mjh_abi_version"):
      continue
    getattr(L, name).restype = ctypes.c_int
  L.mjh_check.argtypes = [P, P]
  L.mjh_solve_m.argtypes = [P, P, P, P, P]
  L.mjh_mul_m.argtypes = [P, P, P, P, P, P]
  L.mjh_xfrc_accumulate.argtypes = [P, P, P, P]
  L.mjh_rollout.argtypes = [P, P, ctypes.c_int, ctypes.c_int, P]
  L.mjh_multi_create.argtypes = [ctypes.c_int, P, P, P]
  L.mjh_multi_create_tuned.argtypes = [ctypes.c_int, P, P, P, ctypes.c_int, P]
  L.mjh_multi_step.argtypes = [P, P]
  L.mjh_multi_step_static.argtypes = [ctypes.c_int, P, P, P]
  L.mjh_multi_destroy.argtypes = [P]
  L.mjh_transition_fd.argtypes = [P, P, P, ctypes.c_float, ctypes.c_int, P, P, P]
  L.mjh_adjoint_step.argtypes = [P, P, P, P, P, P, P, P, P, P, P]
  L.mjh_adjoint_state_grad.argtypes = [P, P, P, P, P, P, P]
  L.mjh_adjoint_vjp.argtypes = [P, P, P, P, P, P, P, P, P]
  L.mjh_adjoint_tape_bytes.argtypes = [P, P, P]
  L.mjh_adjoint_tape_save.argtypes = [P, P, P, P]
  L.mjh_adjoint_vjp_taped.argtypes = [P, P, P, P, P, P, P, P, P, P]
  return L


def _raise(rc: int, what: str):
  msg = lib().mjh_last_error().decode()
  if rc == MJH_ERR_UNSUPPORTED:
    raise NotImplementedError(msg)
  if rc in (MJH_ERR_INVALID, MJH_ERR_CAPACITY):
    raise ValueError(f"{what}: {msg}")
  raise RuntimeError(f"{what} failed ({rc}): {msg}")


def _stream():
  import torch

  return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def check(m, d):
  ms = model_struct(m)
  ds = data_struct(m, d)
  rc = lib().mjh_check(ctypes.byref(ms), ctypes.byref(ds))
  if rc != MJH_OK:
    _raise(rc, "mjh_check")


def call(name: str, m, d, *extra):
  """Invokes entry point `name` on (m, d) with extra pointer/int arguments."""
  ms = model_struct(m)
  ds = data_struct(m, d)
  if getattr(d, "_checked", None) is not m:
    check(m, d)
    object.__setattr__(d, "_checked", m)
  fn = getattr(lib(), name)
  args = [ctypes.byref(ms), ctypes.byref(ds)] + list(extra) + [_stream()]
  rc = fn(*args)
  if rc != MJH_OK:
    _raise(rc, name)
