"""The C-ABI library loads and exports every entry point declared in
include/mjx_warp_amd.h (no compute: runs without a GPU)."""

import ctypes
import re
from pathlib import Path

import pytest

from mjx_warp_amd import _native

ROOT = Path(__file__).resolve().parent.parent


def test_header_entry_points_exported():
  lib = ctypes.CDLL(str(_native.LIBRARY))
  names = re.findall(r"^(?:int|const char\*)\s+(mjh_\w+)\(", (ROOT / "include" / "mjx_warp_amd.h").read_text(), re.M)
  assert len(names) >= 30
  for n in names:
    assert hasattr(lib, n), n


def test_abi_version_and_error_string():
  lib = _native.lib()
  assert lib.mjh_abi_version() == 1
  assert isinstance(lib.mjh_last_error(), bytes)


def test_struct_layout_matches_header():
  ModelStruct, DataStruct = _native.structs()
  msizes, marrays = _native.layout()["model"]
  dsizes, darrays = _native.layout()["data"]
  assert len(ModelStruct._fields_) == len(msizes) + 2 + len(marrays)
  assert len(DataStruct._fields_) == len(dsizes) + len(darrays)
  # option block: 7 floats + 8 ints
  assert ctypes.sizeof(_native.OptionStruct) == 15 * 4


def test_check_rejects_null():
  lib = _native.lib()
  assert lib.mjh_check(None, None) == _native.MJH_ERR_INVALID
  assert b"null" in lib.mjh_last_error()


def test_device_call_without_gpu_tensors_raises():
  import numpy as np
  import mjx_warp_amd as mjw
  from mjx_warp_amd import models

  mjm = models.chain(3)
  m = mjw.put_model(mjm, device="cpu")
  d = mjw.make_data(mjm, nworld=2, device="cpu")
  with pytest.raises(ValueError, match="HIP device"):
    mjw.kinematics(m, d)
