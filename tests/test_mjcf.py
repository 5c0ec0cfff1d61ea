"""MJCF compiler: the reference's fixtures compile with the element counts
their XML declares; structural elements the compiler does not expand are
refused instead of dropped; models with equality constraints, tendons or
sensors reach put_model, which raises the reference's NotImplementedError
(io.py:25-32)."""

import xml.etree.ElementTree as ET
from pathlib import Path

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
  '<replicate count="2" euler="0 0 90"><body><geom size=".1"/></body></replicate>',
  '<frame pos="0 0 1"><body><geom size=".1"/></body></frame>',
  '<body><attach model="m" body="b" prefix="_"/></body>',
])
def test_structural_elements_are_refused(element):
  with pytest.raises(NotImplementedError, match="not supported"):
    mjcf.from_xml_string(f"<mujoco><worldbody>{element}</worldbody></mujoco>")


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
