"""put_data / get_data_into, the state wire format (io.py:509-831).

The input of put_data is a host MjData after a forward pass; here that is
world 0 of the float64 oracle copied into MuJoCo's host layouts
(helpers.mjdata_from_oracle), on the reference's own pendula scene and on
the humanoid pressed into the floor (contacts and constraint rows).

* make_data and put_data build the same fields with the same shapes and
  dtypes (support_test.py:75-86);
* put_data tiles the state across worlds and lays each world's contacts
  and rows out as one block of the pools (world_con / world_efc);
* get_data_into(put_data(mjd)) returns mjd (to float32 rounding): the
  state, the derived fields, MuJoCo's sparse qM, contacts and rows;
* capacity errors are ValueError (io.py:529-530);
* on the GPU, put_data -> step -> get_data_into equals the float32 oracle's
  step from the same host data.
"""

import numpy as np
import pytest
import torch

import mjx_warp_amd as mjw
from mjx_warp_amd import io, mjcf, models
from oracle import oracle

import helpers


def _forwarded(name):
  """(mjm, mjd): a host MjData after one forward pass of the float64 oracle."""
  rng = np.random.default_rng(5)
  if name == "humanoid":
    mjm = models.humanoid(iterations=10, ls_iterations=20)
    q = mjm.key_qpos[0].copy()
    q[2] -= 0.01
    q[7:] += rng.normal(0, 0.1, mjm.nq - 7)
  else:
    mjm = mjcf.from_xml_path(str(helpers_golden(name)))
    q = mjm.qpos0.copy()
    for j in range(mjm.njnt):
      if mjm.jnt_type[j] in (2, 3):
        q[mjm.jnt_qposadr[j]] += rng.normal(0, 0.3)
  mjm, m, d, hm, hd = helpers.load(mjm, nworld=1, nconmax=64, njmax=277)
  hd.arrays["qpos"][0] = q
  hd.arrays["qvel"][0] = rng.normal(0, 0.3, mjm.nv)
  if mjm.nu:
    hd.arrays["ctrl"][0] = rng.uniform(-1, 1, mjm.nu)
  oracle.run("forward", hm, hd)
  return mjm, helpers.mjdata_from_oracle(mjm, hd, 0)


def helpers_golden(name):
  from pathlib import Path
  return Path(__file__).resolve().parent / "golden" / name


def _fields(d, prefix=""):
  out = {}
  for k, v in vars(d).items():
    if isinstance(v, torch.Tensor):
      out[prefix + k] = v
    elif hasattr(v, "__dict__") and not k.startswith("_") and k not in ("owner",):
      out.update(_fields(v, prefix + k + "."))
  return out


@pytest.mark.parametrize("name", ["pendula.xml", "humanoid"])
def test_make_data_and_put_data_agree_on_fields_and_shapes(name):
  mjm, mjd = _forwarded(name)
  nworld = 3
  nconmax, njmax = max(512, 3 * mjd.ncon), max(512, 3 * mjd.nefc)
  pd = io.put_data(mjm, mjd, nworld=nworld, nconmax=nconmax, njmax=njmax, device="cpu")
  md = io.make_data(mjm, nworld=nworld, nconmax=nconmax, njmax=njmax, device="cpu")
  a, b = _fields(pd), _fields(md)
  assert a.keys() == b.keys()
  for k in a:
    assert a[k].shape == b[k].shape and a[k].dtype == b[k].dtype, (k, a[k].shape, b[k].shape)
  # default capacities follow the reference's (io.py:517-524)
  dd = io.put_data(mjm, mjd, nworld=2, device="cpu")
  assert dd.nconmax == max(512, 2 * mjd.ncon) and dd.njmax == max(512, 2 * mjd.nefc)


def test_put_data_tiles_worlds_and_pools():
  mjm, mjd = _forwarded("humanoid")
  assert mjd.ncon > 0 and mjd.nefc > 0
  nworld = 4
  d = io.put_data(mjm, mjd, nworld=nworld, nconmax=64 * nworld, njmax=277 * nworld, device="cpu")
  for w in range(nworld):
    np.testing.assert_allclose(d.qpos[w].numpy(), mjd.qpos, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(d.qvel[w].numpy(), mjd.qvel, rtol=1e-6, atol=1e-7)
  wc = d.world_con.numpy()
  we = d.world_efc.numpy()
  assert np.array_equal(wc, np.stack([np.arange(nworld) * mjd.ncon, np.full(nworld, mjd.ncon)], 1))
  assert np.array_equal(we, np.stack([np.arange(nworld) * mjd.nefc, np.full(nworld, mjd.nefc)], 1))
  assert int(d.ncon[0]) == nworld * mjd.ncon and int(d.nefc[0]) == nworld * mjd.nefc
  wid = d.efc.worldid.numpy()[: nworld * mjd.nefc]
  assert np.array_equal(wid, np.repeat(np.arange(nworld), mjd.nefc))
  for w in range(nworld):
    blk = slice(w * mjd.nefc, (w + 1) * mjd.nefc)
    np.testing.assert_allclose(d.efc.J.numpy()[blk], mjd.efc_J, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(d.efc.D.numpy()[blk], mjd.efc_D, rtol=1e-6)
    cb = slice(w * mjd.ncon, (w + 1) * mjd.ncon)
    np.testing.assert_allclose(d.contact.dist.numpy()[cb], mjd.contact.dist, rtol=1e-6, atol=1e-7)
    assert np.array_equal(d.contact.geom.numpy()[cb], mjd.contact.geom)
  # the dense mass matrix and its Cholesky factor from MuJoCo's sparse qM
  M = io._dense_qM(mjm, mjd.qM)
  np.testing.assert_allclose(d.qM[2].numpy(), M, rtol=1e-6, atol=1e-6)
  np.testing.assert_allclose(d.qLD[2].numpy(), np.linalg.cholesky(M), rtol=1e-5, atol=1e-5)


def test_put_data_capacity_errors():
  mjm, mjd = _forwarded("humanoid")
  with pytest.raises(ValueError):
    io.put_data(mjm, mjd, nworld=4, nconmax=4 * mjd.ncon, njmax=4 * mjd.nefc - 1, device="cpu")
  with pytest.raises(ValueError):
    io.put_data(mjm, mjd, nworld=4, nconmax=4 * mjd.ncon - 1, njmax=4 * mjd.nefc, device="cpu")


FIELDS_VEC = ("qpos", "qvel", "qacc", "qacc_warmstart", "ctrl", "xpos", "xquat", "xipos", "xanchor", "xaxis",
              "subtree_com", "geom_xpos", "cinert", "cdof", "crb", "cvel", "cdof_dot", "actuator_length",
              "actuator_velocity", "actuator_force", "qfrc_bias", "qfrc_passive", "qfrc_actuator", "qfrc_smooth",
              "qfrc_constraint", "qacc_smooth", "xmat", "ximat", "geom_xmat", "actuator_moment", "qM")


@pytest.mark.parametrize("name", ["pendula.xml", "humanoid"])
def test_get_data_into_inverts_put_data(name):
  mjm, mjd = _forwarded(name)
  d = io.put_data(mjm, mjd, nworld=1, device="cpu")
  out = mjcf.MjData(mjm)
  io.get_data_into(out, mjm, d)
  for f in FIELDS_VEC:
    a, b = np.asarray(getattr(out, f)), np.asarray(getattr(mjd, f))
    scale = 1 + np.abs(b).max() if b.size else 1
    np.testing.assert_allclose(a.reshape(b.shape), b, atol=1e-6 * scale, err_msg=f)
  assert out.ncon == mjd.ncon and out.nefc == mjd.nefc
  n = mjd.ncon
  for f in ("dist", "pos", "frame", "friction", "solref", "solimp", "includemargin", "dim", "geom") if n else ():
    a = np.asarray(getattr(out.contact, f)).reshape(n, -1)
    b = np.asarray(getattr(mjd.contact, f)).reshape(a.shape)
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-6, err_msg=f)
  for f in ("J", "D", "pos", "aref", "force", "margin"):
    a, b = getattr(out, "efc_" + f), getattr(mjd, "efc_" + f)
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-6 * (1 + np.abs(b).max() if b.size else 1), err_msg=f)
  with pytest.raises(NotImplementedError):
    io.get_data_into(out, mjm, io.put_data(mjm, mjd, nworld=2, device="cpu"))


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [False, True])
def test_put_data_step_get_data_into_matches_oracle(fused):
  mjm, mjd = _forwarded("humanoid")
  m = mjw.put_model(mjm)
  d = mjw.put_data(mjm, mjd, nworld=1, nconmax=64, njmax=277)
  (mjw.step_fused if fused else mjw.step)(m, d)
  torch.cuda.synchronize()
  out = mjcf.MjData(mjm)
  mjw.get_data_into(out, mjm, d)
  # the oracle from the same host data
  mc = mjw.put_model(mjm, device="cpu")
  dc = io.put_data(mjm, mjd, nworld=1, nconmax=64, njmax=277, device="cpu")
  hm = helpers.host_model(mc, np.float32)
  hd = helpers.host_data(mc, dc, np.float32)
  oracle.run("step", hm, hd)
  for f, tol in (("qpos", 1e-4), ("qvel", 2e-3), ("qacc", 2e-3), ("xpos", 1e-4), ("qfrc_smooth", 2e-4)):
    ref = hd[f][0].astype(np.float64)
    np.testing.assert_allclose(np.asarray(getattr(out, f)).reshape(ref.shape), ref,
                               atol=tol * (1 + np.abs(ref).max()), err_msg=f)
  assert out.ncon == int(hd["ncon"][0]) and out.nefc == int(hd["nefc"][0])
  np.testing.assert_allclose(out.efc_force, hd["efc_force"][: out.nefc], atol=2e-3 * (1 + np.abs(hd["efc_force"]).max()))
