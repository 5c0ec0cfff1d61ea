"""Enums and the Model / Data containers of the drop-in API.

Names, values and field layout follow mujoco_warp/_src/types.py:27-800.
Arrays are torch tensors (float32 / int32 / bool) on the HIP device, laid
out world-major exactly as the reference's Warp arrays: per-world arrays are
[nworld, ...]; contacts and constraint rows are global pools of capacity
nconmax / njmax whose entries carry a world id.

The field list itself lives in include/mjx_warp_amd.h (one X-macro list that
also defines the C ABI); `_native.layout()` parses it so Python, C and the
kernels agree on every name and shape.
"""

from __future__ import annotations

import enum

MJ_MINVAL = 1e-15
MJ_MINIMP = 0.0001
MJ_MAXIMP = 0.9999
MJ_NREF = 2
MJ_NIMP = 5


class DisableBit(enum.IntFlag):
  """Disable default feature bitflags (types.py:27-53)."""

  CONSTRAINT = 1 << 0
  LIMIT = 1 << 3
  CONTACT = 1 << 4
  PASSIVE = 1 << 5
  GRAVITY = 1 << 6
  CLAMPCTRL = 1 << 7
  ACTUATION = 1 << 10
  REFSAFE = 1 << 11
  EULERDAMP = 1 << 14
  FILTERPARENT = 1 << 9


class TrnType(enum.IntEnum):
  JOINT = 0
  JOINTINPARENT = 1


class DynType(enum.IntEnum):
  NONE = 0
  FILTEREXACT = 3


class GainType(enum.IntEnum):
  FIXED = 0
  AFFINE = 1


class BiasType(enum.IntEnum):
  NONE = 0
  AFFINE = 1


class JointType(enum.IntEnum):
  FREE = 0
  BALL = 1
  SLIDE = 2
  HINGE = 3

  def dof_width(self) -> int:
    return {0: 6, 1: 3, 2: 1, 3: 1}[self.value]

  def qpos_width(self) -> int:
    return {0: 7, 1: 4, 2: 1, 3: 1}[self.value]


class ConeType(enum.IntEnum):
  PYRAMIDAL = 0


class IntegratorType(enum.IntEnum):
  EULER = 0
  IMPLICITFAST = 3


class GeomType(enum.IntEnum):
  PLANE = 0
  SPHERE = 2
  CAPSULE = 3
  BOX = 6


class SolverType(enum.IntEnum):
  CG = 1
  NEWTON = 2


class _Node:
  """Attribute container whose assignments mark the owning Model/Data stale,
  so the cached C description is rebuilt before the next call."""

  def __init__(self, owner=None):
    object.__setattr__(self, "_owner", owner)

  _HOST_ONLY = ("time", "nl")

  def __setattr__(self, name, value):
    object.__setattr__(self, name, value)
    if name.startswith("_") or name in self._HOST_ONLY:
      return
    owner = self._owner if self._owner is not None else self
    object.__setattr__(owner, "_abi", None)
    object.__setattr__(owner, "_checked", None)


class Option(_Node):
  """Physics options (types.py:218-250)."""


class Statistic(_Node):
  """Model statistics in qpos0 (types.py:253-261)."""


class Contact(_Node):
  """Contact pool (types.py:615-645)."""


class Constraint(_Node):
  """Constraint rows and solver state (types.py:264-363)."""


class Model(_Node):
  """Model definition and parameters (types.py:366-612)."""

  def __init__(self):
    super().__init__(None)
    object.__setattr__(self, "_abi", None)
    self.opt = Option(self)
    self.stat = Statistic(self)


class Data(_Node):
  """Dynamic state (types.py:648-800).  `time` and `nl` are host values."""

  def __init__(self):
    super().__init__(None)
    object.__setattr__(self, "_abi", None)
    self.contact = Contact(self)
    self.efc = Constraint(self)
    self.time = 0.0
    self.nl = 0
