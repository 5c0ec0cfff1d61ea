"""MJCF compiler: the reference's fixtures compile with the element counts
their XML declares; structural elements the compiler does not expand are
refused instead of dropped; models with equality constraints, tendons or
sensors reach put_model, which raises the reference's NotImplementedError
(io.py:25-32)."""

import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np
import pytest

import mjx_warp_amd as mjw
from mjx_warp_amd import mjcf

GOLDEN = Path(__file__).resolve().parent / "golden"
ASSETS = Path(mjw.__file__).resolve().parent / "assets"


@pytest.mark.parametrize("path", [GOLDEN / "pendula.xml", GOLDEN / "constraints.xml", GOLDEN / "collision.xml",
                                  ASSETS / "humanoid.xml"])
def test_fixture_counts(path):
  # the fixtures have attributes without separating whitespace, as MuJoCo
  # accepts; the compiler's own repair makes them parse here too
  root = ET.fromstring(mjcf._FIX_ATTR.sub(r"\1 ", path.read_text()))
  world = root.find("worldbody")
  m = mjcf.from_xml_path(str(path))
  assert m.nbody == 1 + len(world.findall(".//body"))
  assert m.njnt == len(world.findall(".//joint")) + len(world.findall(".//freejoint"))
  assert m.ngeom == len(world.findall(".//geom"))
  act = root.find("actuator")
  assert m.nu == (0 if act is None else len(list(act)))


@pytest.mark.parametrize("element", [
  '<composite type="grid" count="2 2 1"/>',
  '<body><include file="x.xml"/></body>',
])
def test_structural_elements_are_refused(element):
  with pytest.raises(NotImplementedError, match="not supported"):
    mjcf.from_xml_string(f"<mujoco><worldbody>{element}</worldbody></mujoco>")


def test_replicate_and_frame_expand():
  """A replicate of count 4 turning 90 degrees about z, around a frame at
  (1, 0, 0.5) holding a body with a free sphere: four bodies at the corners
  (cos, sin of k 90 deg) x 1, height 0.5, names suffixed 0..3 after sep."""
  m = mjcf.from_xml_string("""<mujoco><worldbody>
    <replicate count="4" euler="0 0 90" sep="_">
      <frame pos="1 0 0.5"><body name="b"><freejoint/><geom type="sphere" size=".1"/></body></frame>
    </replicate></worldbody></mujoco>""")
  assert m.nbody == 5 and m.nv == 24
  for k in range(4):
    b = m.names["body"].index(f"b_{k}")
    np.testing.assert_allclose(m.body_pos[b], [np.cos(k * np.pi / 2), np.sin(k * np.pi / 2), 0.5], atol=1e-12)
  # a frame's rotation carries a geom's fromto endpoints with it
  m = mjcf.from_xml_string("""<mujoco><worldbody><body name="a"><freejoint/>
    <frame euler="0 0 90"><geom type="capsule" fromto="0 0 0 1 0 0" size=".1"/></frame></body>
    </worldbody></mujoco>""")
  np.testing.assert_allclose(m.geom_pos[0], [0, 0.5, 0], atol=1e-12)


def test_attach_needs_a_model_asset():
  with pytest.raises(ValueError, match="attach"):
    mjcf.from_xml_string('<mujoco><worldbody><attach model="m" body="b" prefix="_"/></worldbody></mujoco>')


@pytest.mark.parametrize("section,message", [
  ('<equality><weld body1="a"/></equality>', "Equality"),
  ('<tendon><fixed name="t"><joint joint="j" coef="1"/></fixed></tendon>', "Tendons"),
  ('<sensor><jointpos joint="j"/></sensor>', "Sensors"),
])
def test_unsupported_features_raise_in_put_model(section, message):
  xml = ('<mujoco><worldbody><body name="a"><joint name="j"/><geom size=".1"/></body></worldbody>'
         f'{section}</mujoco>')
  m = mjcf.from_xml_string(xml)
  with pytest.raises(NotImplementedError, match=message):
    mjw.put_model(m, device="cpu")


def test_replicate_attach_n_humanoids():
  """The reference's n_humanoids.xml (test_data/humanoid): <replicate
  count=3 euler="0 0 16.36"> of a <frame pos="0 -4 0"> holding <attach
  model="humanoid" body="torso" prefix="_">.  Three humanoids: element
  counts three times the humanoid's plus the floor, torso k at
  Rz(16.36 k deg) (0, -4, 1.282), and each copy's block of the mass matrix
  and its mj_setConst quantities equal to the single humanoid's turned (its
  free joint's translational dofs are world-frame and turn with the copy, the
  rest of the matrix is the humanoid's)."""
  from mjx_warp_amd import models

  m = mjcf.from_xml_path(str(models.ASSETS / "n_humanoids.xml"))
  h = mjcf.from_xml_path(str(models.ASSETS / "humanoid.xml"))
  assert (m.nq, m.nv, m.nu, m.njnt) == (3 * h.nq, 3 * h.nv, 3 * h.nu, 3 * h.njnt)
  assert m.nbody == 1 + 3 * (h.nbody - 1) and m.ngeom == 1 + 3 * (h.ngeom - 1)
  assert m.opt.timestep == 0.005
  th = np.deg2rad(16.36)
  for k in range(3):
    b = m.names["body"].index(f"_torso-{k}")
    c, s = np.cos(k * th), np.sin(k * th)
    np.testing.assert_allclose(m.body_pos[b], [4 * s, -4 * c, 1.282], atol=1e-9)
    np.testing.assert_allclose(abs(m.body_quat[b][0]), abs(np.cos(k * th / 2)), atol=1e-9)
  np.testing.assert_allclose(m.actuator_gear[:, 0], np.tile(h.actuator_gear[:, 0], 3))
  M, Mh = mjcf.mass_matrix(m, m.qpos0), mjcf.mass_matrix(h, h.qpos0)
  for k in range(3):
    blk = slice(k * h.nv, (k + 1) * h.nv)
    # copy k is the humanoid turned by Rz(k theta): the free joint's
    # translational dofs are world-frame, so M_k = T M T' with T = diag(Rz, I)
    T = np.eye(h.nv)
    c, s = np.cos(k * th), np.sin(k * th)
    T[:3, :3] = [[c, -s, 0], [s, c, 0], [0, 0, 1]]
    np.testing.assert_allclose(M[blk, blk], T @ Mh @ T.T, atol=1e-9)
    np.testing.assert_allclose(m.dof_invweight0[blk], h.dof_invweight0, rtol=1e-9)
  off = M.copy()
  for k in range(3):
    off[k * h.nv:(k + 1) * h.nv, k * h.nv:(k + 1) * h.nv] = 0
  assert not np.any(off)  # three separate trees
