"""World-steps/s of the batched physics step on one MI355X.

Workload (default): the 27-dof humanoid, B = 8192 worlds, all 21 motor gears
100, ctrl ~ U(-1, 1) resampled every step on the device, Newton solver,
implicitfast, dt = 0.005, FP32, from the "squat" keyframe.  Solver
iterations / line-search iterations default to the model file's 10 / 20;
--iterations 1 --ls_iterations 4 selects the reference benchmark's defaults
(unstable for this model with implicitfast: worlds reach NaN within tens of
steps).

One step = the whole of mujoco_warp.step for all B worlds (forward
kinematics, inertia, collision, constraint rows, solver, integration) plus
the ctrl resample.  `value` is B * steps / wall time of the timed loop with
all state resident in HBM; the fused step (a forward launch and a
solve+integrate launch, plus a tail for worlds over the solver's row
capacity) is the headline, the plain per-stage step is reported beside it,
and so are the batched-persistent form (10 fused steps captured as one HIP
graph, replayed), forward + adjoint world-steps/s of a differentiable
50-step rollout of 1024 humanoids (`forward_adjoint`), and a heterogeneous
batch of 16,384 worlds over 28 different models (`heterogeneous`).
`roofline.traffic` is the fused step's HBM traffic measured with rocprofv3
for the default configuration (profiles/step_humanoid_b8192_newton10.md).

Prints one JSON line on stdout; progress goes to stderr.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
  sys.path.insert(0, ROOT)

METRIC = "world-steps/s, 27-DoF humanoid B=8192; fwd+adjoint world-steps/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
  print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(mjm, ncon_w: float, nefc_w: float) -> float:
  """Bytes of Data fields one world-step must read or write, counted once in
  FP32: reads R = nq + 3nv + nu + na + 6 nbody + 7 nmocap; writes
  W = nq + 23nv + 57 nbody + 6 njnt + 12 ngeom + 12 nsite + 3nu + nu nv + 2na
  + 2 nv^2 (dense qM and qLD); plus 30 words per contact and nv + 6 words per
  constraint row, at the measured per-world means."""
  R = mjm.nq + 3 * mjm.nv + mjm.nu + mjm.na + 6 * mjm.nbody + 7 * mjm.nmocap
  W = (mjm.nq + 23 * mjm.nv + 57 * mjm.nbody + 6 * mjm.njnt + 12 * mjm.ngeom + 12 * mjm.nsite + 3 * mjm.nu
       + mjm.nu * mjm.nv + 2 * mjm.na + 2 * mjm.nv * mjm.nv)
  return 4.0 * (R + W + 30 * ncon_w + (mjm.nv + 6) * nefc_w)


PROFILE = os.path.join(ROOT, "profiles", "step_humanoid_b8192_newton10.md")


def kernel_sources_sha() -> str:
  """sha256 over the stage kernels' sources (csrc/*.hpp and kernels.hip in
  name order; api.hip holds the host code): the profile summary records
  it, so a measured figure is used only for the kernels it was measured
  on."""
  import glob
  import hashlib

  h = hashlib.sha256()
  # adjoint.hpp holds the step adjoint's kernel (built from api.hip only),
  # not a stage of the step
  hpp = [p for p in glob.glob(os.path.join(ROOT, "mjx_warp_amd", "csrc", "*.hpp"))
         if os.path.basename(p) != "adjoint.hpp"]
  for path in sorted(hpp) + [os.path.join(ROOT, "mjx_warp_amd", "csrc", "kernels.hip")]:
    with open(path, "rb") as f:
      h.update(os.path.basename(path).encode() + b"\0" + f.read())
  return h.hexdigest()[:16]


def measured_traffic(args):
  """HBM bytes of one fused step measured with rocprofv3 FETCH_SIZE /
  WRITE_SIZE (scaled by a plain copy's counters) for the default workload,
  as recorded in the profile summary; None for other configurations or
  when the profile was taken on other kernel sources."""
  if (args.batch, args.iterations, args.ls_iterations) != (8192, 10, 20):
    return None
  traffic, sha = None, None
  try:
    with open(PROFILE) as f:
      for line in f:
        if line.startswith("fused step traffic per launch"):
          traffic = float(line.rsplit(":", 1)[1])
        if line.startswith("kernel sources sha256"):
          sha = line.rsplit(":", 1)[1].strip()
  except OSError:
    pass
  return traffic if sha == kernel_sources_sha() else None


def setup(args, device="cuda"):
  import numpy as np
  import torch

  import mjx_warp_amd as mjw
  from mjx_warp_amd import models

  mjm = models.humanoid(gear=100.0, iterations=args.iterations, ls_iterations=args.ls_iterations)
  m = mjw.put_model(mjm, device=device)
  B = args.batch
  d = mjw.make_data(mjm, nworld=B, nconmax=64 * B, njmax=(4 * 64 + 21) * B, device=device)
  d.qpos = torch.tensor(np.tile(mjm.key_qpos[0], (B, 1)), dtype=torch.float32, device=device)
  return mjm, m, d


def time_form(args, fused: bool, preroll: int = 0):
  """Times args.steps steps after args.warmup warm-up steps and `preroll`
  more untimed steps (1000 take the humanoids from the squat keyframe to
  lying on the floor, 100-150 constraint rows each)."""
  import torch

  import mjx_warp_amd as mjw

  mjm, m, d = setup(args)
  gen = torch.Generator(device="cuda")
  gen.manual_seed(1)
  stepfn = mjw.step_fused if fused else mjw.step
  if args.graph:
    # one captured step replayed per timed step (the reference benchmark's
    # form, test_util.py:99-118); ctrl is resampled eagerly between replays
    graph = mjw.StepGraph(m, d, nstep=1, fused=fused)
    stepfn = lambda m_, d_: graph.replay()
  u = torch.empty_like(d.ctrl)

  def one_step():
    torch.rand(u.shape, generator=gen, device="cuda", out=u)
    d.ctrl.copy_(u.mul_(2.0).sub_(1.0))
    stepfn(m, d)

  for _ in range(args.warmup + preroll):
    one_step()
  torch.cuda.synchronize()
  stream = torch.cuda.current_stream()
  ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
  ncon = torch.zeros((), dtype=torch.int64, device="cuda")
  nefc = torch.zeros((), dtype=torch.int64, device="cuda")
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for k in range(args.steps):
    torch.rand(u.shape, generator=gen, device="cuda", out=u)
    d.ctrl.copy_(u.mul_(2.0).sub_(1.0))
    ev[k][0].record(stream)
    stepfn(m, d)
    ev[k][1].record(stream)
    ncon += d.ncon[0]
    nefc += d.nefc[0]
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  step_ms = sum(a.elapsed_time(b) for a, b in ev) / args.steps
  B = args.batch
  res = dict(wall=wall, ms_per_step=1000.0 * wall / args.steps, kernel_ms=step_ms,
             value=B * args.steps / wall, ncon_w=float(ncon.item()) / args.steps / B,
             nefc_w=float(nefc.item()) / args.steps / B,
             finite=bool(torch.isfinite(d.qpos).all().item()))
  return mjm, res


def time_rollout_graph(args, nstep: int = 10):
  """The batched-persistent form: nstep fused steps, each after copying its
  own ctrl row into d.ctrl, captured once as a HIP graph and replayed; the
  nstep ctrl rows are resampled from U(-1, 1) before each replay, so every
  step sees fresh controls as in the eager loop."""
  import torch

  import mjx_warp_amd as mjw
  from mjx_warp_amd import _native

  _, m, d = setup(args)
  gen = torch.Generator(device="cuda")
  gen.manual_seed(1)
  u = torch.empty((nstep,) + tuple(d.ctrl.shape), dtype=torch.float32, device="cuda")
  _native.check(m, d)
  object.__setattr__(d, "_checked", m)
  graph = torch.cuda.CUDAGraph()
  side = torch.cuda.Stream()
  side.wait_stream(torch.cuda.current_stream())
  with torch.cuda.graph(graph, stream=side):
    for k in range(nstep):
      d.ctrl.copy_(u[k])
      _native.call("mjh_step_fused", m, d)
  torch.cuda.current_stream().wait_stream(side)
  reps = max(1, args.steps // nstep)

  def one():
    torch.rand(u.shape, generator=gen, device="cuda", out=u)
    u.mul_(2.0).sub_(1.0)
    graph.replay()

  for _ in range(max(1, args.warmup // nstep)):
    one()
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(reps):
    one()
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  return {"value": args.batch * reps * nstep / wall, "ms_per_step": 1000.0 * wall / (reps * nstep),
          "steps_per_graph": nstep, "replays": reps}


def time_adjoint(args, nworld: int = 1024, nstep: int = 50, eps: float = 4e-3, reps: int = 2, fd: bool = True):
  """Forward + adjoint world-steps/s of the differentiable rollout: the
  humanoid (gear 100, implicitfast, Newton) from perturbed squat states,
  nworld worlds, nstep steps, ctrl ~ U(-1, 1) [nstep, nworld, 21], loss
  sum_t (torso z - 1.3)^2 + 1e-3 |ctrl|^2, gradients with respect to ctrl and
  qvel0.  The value is the analytic adjoint (each backward step recomputes
  its forward quantities with one fused step and takes one vector-Jacobian
  product); the finite-difference adjoint (centered transition Jacobians,
  eps) is timed beside it."""
  import torch

  import mjx_warp_amd as mjw
  from mjx_warp_amd import derivative, models

  mjm = models.humanoid(gear=100.0, iterations=args.iterations, ls_iterations=args.ls_iterations)
  m = mjw.put_model(mjm)
  q0, v0 = models.humanoid_rollout_state(mjm, nworld)
  gen = torch.Generator(device="cuda")
  gen.manual_seed(1)
  ctrl = torch.rand((nstep, nworld, mjm.nu), generator=gen, device="cuda") * 2.0 - 1.0
  tq0 = torch.tensor(q0, dtype=torch.float32, device="cuda")
  tv0 = torch.tensor(v0, dtype=torch.float32, device="cuda")

  def measure(r):
    def fwd_adj():
      qv = tv0.clone().requires_grad_(True)
      u = ctrl.clone().requires_grad_(True)
      qpos, _ = r(tq0, qv, u)
      loss = ((qpos[..., 2] - 1.3) ** 2).sum() + 1e-3 * (u**2).sum()
      loss.backward()
      return u.grad, qv.grad

    def fwd():
      # the rollout alone: no tape writes
      with torch.no_grad():
        r.forward(tq0, tv0, ctrl, record=False)

    fwd_adj()
    fwd()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
      fwd()
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
      gu, gv = fwd_adj()
    torch.cuda.synchronize()
    t_all = (time.perf_counter() - t0) / reps
    finite = bool(torch.isfinite(gu).all().item() and torch.isfinite(gv).all().item())
    return {"value": nworld * nstep / t_all, "forward_value": nworld * nstep / t_fwd, "cost_ratio": t_all / t_fwd,
            "ms_forward": 1000.0 * t_fwd, "ms_forward_adjoint": 1000.0 * t_all, "gradients_finite": finite}, gu

  ra = derivative.Rollout(mjm, m, nworld, nstep, nconmax=64 * nworld, njmax=277 * nworld)
  out, gu_a = measure(ra)
  taped = ra.tape is not None
  tape_mb = ra.tape.numel() / 2**20 if taped else 0.0
  del ra
  out.update({"unit": "world-steps/s", "nworld": nworld, "nstep": nstep, "adjoint": "analytic",
              "taped": taped, "tape_mib": tape_mb,
              "workload": "humanoid 27-dof rollout, gradients of sum_t (z_t - 1.3)^2 + 1e-3 |u_t|^2 w.r.t. "
                          "ctrl [50, 1024, 21] and qvel0 [1024, 27]; analytic step adjoint (one "
                          "vector-Jacobian product per step, each step's fields taped in the forward pass)"})
  rr = derivative.Rollout(mjm, m, nworld, nstep, nconmax=64 * nworld, njmax=277 * nworld, tape=False)
  r2, _ = measure(rr)
  del rr
  out["recompute"] = {k: r2[k] for k in ("value", "cost_ratio", "ms_forward_adjoint")}
  if fd:
    rf = derivative.Rollout(mjm, m, nworld, nstep, nconmax=64 * nworld, njmax=277 * nworld, eps=eps, adjoint="fd")
    f, gu_f = measure(rf)
    f.update({"eps": eps, "fd_worlds": nworld * rf.tfd.per_world,
              "ctrl_grad_rel_diff_to_analytic": float(torch.linalg.norm(gu_f - gu_a) / torch.linalg.norm(gu_a))})
    out["finite_difference"] = f
    del rf
  return out


def time_heterogeneous(args, nworld: int = 16384, steps: int = 50, graph_steps: int = 10):
  """World-steps/s of the heterogeneous workload: 16,384 worlds, each a
  hinge chain of 2-8 links plus 1-4 free spheres in a walled box, split
  into one model per (links, spheres) pair (28 models, nv 8-32), all
  advanced together: grouped (one launch per stage over every model),
  per model (one fused step per model), the grouped step captured as a
  HIP graph of graph_steps steps and replayed, and the form the autotuner
  picks (per model: lanes per world and LDS budget; buckets by LDS class or
  one launch per stage; grouped or per model), which is `value`."""
  import numpy as np
  import torch

  import mjx_warp_amd as mjw
  from mjx_warp_amd import models

  rng = np.random.default_rng(2)
  groups = []
  spec = models.heterogeneous_groups(nworld)
  for links, radii, n in spec:
    mjm = models.box_world(links, radii)
    q, v = models.box_world_states(mjm, links, radii, n, rng)
    m = mjw.put_model(mjm)
    ncon = 8 * (links + len(radii))
    d = mjw.make_data(mjm, nworld=n, nconmax=ncon * n, njmax=(4 * ncon + mjm.njnt) * n)
    d.qpos.copy_(torch.tensor(q, dtype=torch.float32))
    groups.append((m, d))
  ms = mjw.MultiStep(groups)
  out = {"nworld": nworld, "models": len(groups), "steps": steps,
         "nv_range": [min(int(m.nv) for m, _ in groups), max(int(m.nv) for m, _ in groups)]}
  # every form is timed from the same state
  state = [{f: getattr(d, f).clone() for f in ("qpos", "qvel", "qacc_warmstart")} for _, d in groups]

  def restore():
    for (_, d), st in zip(groups, state):
      for f, v in st.items():
        getattr(d, f).copy_(v)

  for mode in ("grouped", "per_group"):
    ms.mode = mode
    restore()
    for _ in range(3):
      ms.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
      ms.step()
    torch.cuda.synchronize()
    out[mode] = nworld * steps / (time.perf_counter() - t0)
  ms.mode = "grouped"
  restore()
  g = torch.cuda.CUDAGraph()
  side = torch.cuda.Stream()
  side.wait_stream(torch.cuda.current_stream())
  with torch.cuda.graph(g, stream=side):
    for _ in range(graph_steps):
      ms.step()
  torch.cuda.current_stream().wait_stream(side)
  restore()
  g.replay()
  torch.cuda.synchronize()
  reps = max(1, steps // graph_steps)
  t0 = time.perf_counter()
  for _ in range(reps):
    g.replay()
  torch.cuda.synchronize()
  out["grouped_graph"] = nworld * reps * graph_steps / (time.perf_counter() - t0)
  ms.close()
  # the same grouped step with work-queue launches
  mq = mjw.MultiStep(groups, queue=True)
  restore()
  for _ in range(3):
    mq.step()
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(steps):
    mq.step()
  torch.cuda.synchronize()
  out["grouped_queue"] = nworld * steps / (time.perf_counter() - t0)
  mq.close()
  t0 = time.perf_counter()
  restore()
  tuned = mjw.MultiStep(groups, autotune=True)
  out["autotune_s"] = time.perf_counter() - t0
  restore()
  for _ in range(3):
    tuned.step()
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(steps):
    tuned.step()
  torch.cuda.synchronize()
  out["autotuned"] = nworld * steps / (time.perf_counter() - t0)
  teams = {}
  for t in tuned.tunes or []:
    teams[f"team{t[0]}_rj{t[2]}"] = teams.get(f"team{t[0]}_rj{t[2]}", 0) + 1
  out["autotuned_choice"] = {"mode": tuned.mode, "buckets": tuned.buckets, "queue": tuned.queue, "per_model": teams}
  out["value"] = out["autotuned"]
  out["unit"] = "world-steps/s"
  out["finite"] = bool(all(torch.isfinite(d.qpos).all().item() for _, d in groups))
  tuned.close()
  return out


def stated_groups(nworld: int, processes: int = 16):
  """(Model, Data) groups of the heterogeneous workload at its stated shape
  (chains of U{2..32} links, U{1..64} spheres in a box of 5 planes), one per
  (links, spheres) pair present among `nworld` worlds, with their initial
  poses; returns the groups and the seconds spent compiling the models."""
  import numpy as np
  import torch

  import mjx_warp_amd as mjw
  from mjx_warp_amd import models

  t0 = time.perf_counter()
  spec = models.heterogeneous_groups(nworld, max_links=32, max_spheres=64)
  mjms = models.box_worlds(spec, processes=min(processes, os.cpu_count() or 1))
  t_compile = time.perf_counter() - t0
  rng = np.random.default_rng(2)
  groups = []
  for (links, radii, n), mjm in zip(spec, mjms):
    q, _ = models.box_world_states(mjm, links, radii, n, rng)
    m = mjw.put_model(mjm)
    ncon_w = 3 * (links + len(radii))
    d = mjw.make_data(mjm, nworld=n, nconmax=ncon_w * n, njmax=(4 * ncon_w + mjm.njnt) * n)
    d.qpos.copy_(torch.tensor(q, dtype=torch.float32))
    groups.append((m, d))
  torch.cuda.synchronize()
  return groups, t_compile


def time_heterogeneous_stated(args, nworld: int = 16384, steps: int = 100):
  """World-steps/s of the heterogeneous workload at its stated size: 16,384
  worlds, each a hinge chain of U{2..32} 0.1 m capsule links (folded inside
  the box) plus U{1..64} free spheres (radius U(0.03, 0.1), density 1000) in
  a 1 m box of 5 planes, Newton -- one model per (links, spheres) pair
  present, nv 8 to 416 -- advanced together by the grouped step (with and
  without buckets by LDS class, the faster kept), `steps` steps captured as
  one HIP graph and replayed (the value), and eagerly."""
  import torch

  import mjx_warp_amd as mjw

  t0 = time.perf_counter()
  groups, t_compile = stated_groups(nworld)
  t_setup = time.perf_counter() - t0
  log(f"stated-size heterogeneous: {len(groups)} models, setup {t_setup:.1f} s (compile {t_compile:.1f} s)")
  state = [{f: getattr(d, f).clone() for f in ("qpos", "qvel", "qacc_warmstart")} for _, d in groups]

  def restore():
    for (_, d), st in zip(groups, state):
      for f, v in st.items():
        getattr(d, f).copy_(v)

  out = {"nworld": nworld, "models": len(groups), "steps": steps, "setup_s": t_setup,
         "nv_range": [min(int(m.nv) for m, _ in groups), max(int(m.nv) for m, _ in groups)]}
  best = None
  for buckets, queue in ((False, False), (True, False), (True, True)):
    ms = mjw.MultiStep(groups, buckets=buckets, queue=queue)
    restore()
    ms.step()
    torch.cuda.synchronize()
    restore()
    t1 = time.perf_counter()
    for _ in range(10):
      ms.step()
    torch.cuda.synchronize()
    rate = nworld * 10 / (time.perf_counter() - t1)
    out["grouped" + ("_buckets" if buckets else "") + ("_queue" if queue else "")] = rate
    if best is None or rate > best[0]:
      best = (rate, buckets, queue)
    ms.close()
  ms = mjw.MultiStep(groups, buckets=best[1], queue=best[2])
  restore()
  g = torch.cuda.CUDAGraph()
  side = torch.cuda.Stream()
  side.wait_stream(torch.cuda.current_stream())
  with torch.cuda.graph(g, stream=side):
    for _ in range(steps):
      ms.step()
  torch.cuda.current_stream().wait_stream(side)
  restore()
  torch.cuda.synchronize()
  t1 = time.perf_counter()
  g.replay()
  torch.cuda.synchronize()
  out["graph_rollout"] = nworld * steps / (time.perf_counter() - t1)
  out["buckets"] = best[1]
  out["queue"] = best[2]
  out["value"] = out["graph_rollout"]
  out["unit"] = "world-steps/s"
  out["finite"] = bool(all(torch.isfinite(d.qpos).all().item() for _, d in groups))
  out["rows_per_world"] = float(sum(int(d.nefc[0].item()) for _, d in groups)) / nworld
  ms.close()
  del g
  return out


def cpu_baseline(args, target_s: float = 15.0):
  """The oracle (FP32 CPU restatement, OpenMP over worlds) on this host."""
  import numpy as np
  import torch

  from oracle import oracle
  from mjx_warp_amd import _native

  sys.path.insert(0, os.path.join(ROOT, "tests"))
  import helpers

  # the GPU box's CPU share is 16 cores per GPU: the oracle's OpenMP pool
  # is sized to it whatever the host shows
  visible = len(os.sched_getaffinity(0))
  cores = min(16, visible)
  B = min(args.batch, 1024)
  mjm, m, d = setup(argparse.Namespace(**{**vars(args), "batch": B}), device="cpu")
  hm = helpers.host_model(m, np.float32)
  hd = helpers.host_data(m, d, np.float32)
  rng = np.random.default_rng(1)
  # probe
  t0 = time.perf_counter()
  hd.arrays["ctrl"][:] = rng.uniform(-1, 1, hd.arrays["ctrl"].shape)
  oracle.run("step", hm, hd, cores)
  probe = time.perf_counter() - t0
  nstep = max(2, min(200, int(target_s / max(probe, 1e-4))))
  t0 = time.perf_counter()
  for _ in range(nstep):
    hd.arrays["ctrl"][:] = rng.uniform(-1, 1, hd.arrays["ctrl"].shape)
    oracle.run("step", hm, hd, cores)
  dt = time.perf_counter() - t0
  cpu = ""
  try:
    with open("/proc/cpuinfo") as f:
      for line in f:
        if line.startswith("model name"):
          cpu = line.split(":", 1)[1].strip()
          break
  except OSError:
    pass
  return dict(value=B * nstep / dt, unit="world-steps/s", cores=cores, kind="port",
              sample=f"{B} humanoid worlds x {nstep} steps of the oracle (FP32, OpenMP over worlds) "
                     f"after 1 warm-up step, on {cpu or 'host'}",
              threads_used=cores, host_cores_visible=visible, cpu_model=cpu,
              threads_note="the OpenMP pool is the GPU box's CPU share of 16 cores per GPU")


def copy_probe(mib: int, reps: int = 5):
  """A plain device-to-device copy of known size: the reference point that
  scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes."""
  import torch

  n = mib * (1 << 20) // 4
  src = torch.ones(n, dtype=torch.float32, device="cuda")
  dst = torch.empty_like(src)
  for _ in range(reps):
    dst.copy_(src)
  torch.cuda.synchronize()
  log(f"copy probe: {reps} copies of {mib} MiB ({4 * n} bytes read and written each)")


def hetero_stated_probe(nworld: int, warm: int = 40, steps: int = 5):
  """A short run of the stated-shape heterogeneous workload at `nworld`
  worlds for a kernel trace: the grouped step with buckets by LDS class,
  `warm` untimed steps, then `steps` timed; prints one line and returns."""
  import torch

  import mjx_warp_amd as mjw

  groups, _ = stated_groups(nworld, processes=1)
  log(f"stated-shape heterogeneous probe: {len(groups)} models, {nworld} worlds")
  ms = mjw.MultiStep(groups, buckets=True)
  for k in range(warm):
    ms.step()
    if k % 10 == 0:
      torch.cuda.synchronize()
      log(f"step {k}")
  torch.cuda.synchronize()
  t = time.perf_counter()
  for _ in range(steps):
    ms.step()
  torch.cuda.synchronize()
  dt = time.perf_counter() - t
  rows = sum(int(d.nefc[0]) for _, d in groups) / nworld
  log(f"{len(groups)} models, {nworld} worlds: {nworld * steps / dt:.4g} world-steps/s, {rows:.1f} rows per world")
  ms.close()


def collision_probe(nworld: int, calls: int = 20):
  """The contact-heavy scene (128 spheres, 128 capsules, 5 planes: 261
  geoms, so collision takes sweep-and-prune) at `nworld` worlds: times
  sap_broadphase and collision over `calls` calls each; prints and returns."""
  import numpy as np
  import torch

  import mjx_warp_amd as mjw
  from mjx_warp_amd import models

  if nworld % 64:
    raise SystemExit("--collision-probe takes a multiple of 64 worlds (64 distinct drops, tiled)")
  mjm = models.contact_heavy()
  q = np.tile(models.contact_heavy_states(mjm, 64, np.random.default_rng(0)), (nworld // 64, 1))
  m = mjw.put_model(mjm)
  d = mjw.make_data(mjm, nworld=nworld, nconmax=2048 * nworld, njmax=4 * nworld)
  d.qpos.copy_(torch.tensor(q, dtype=torch.float32))
  mjw.kinematics(m, d)
  for _ in range(3):
    mjw.collision(m, d)
  torch.cuda.synchronize()
  for name, fn in (("sap_broadphase", mjw.sap_broadphase), ("collision", mjw.collision)):
    t = time.perf_counter()
    for _ in range(calls):
      fn(m, d)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / calls
    log(f"{name}: {nworld} worlds, {dt * 1e3:.3f} ms per call, {nworld / dt:.4g} worlds/s, "
        f"pairs {int(d.ncollision[0])} contacts {int(d.ncon[0])}")


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=100)
  ap.add_argument("--warmup", type=int, default=3)
  ap.add_argument("--batch", type=int, default=8192)
  ap.add_argument("--iterations", type=int, default=10)
  ap.add_argument("--ls_iterations", type=int, default=20)
  ap.add_argument("--no-plain", action="store_true", help="skip timing the plain per-stage step")
  ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
  ap.add_argument("--no-adjoint", action="store_true", help="skip the forward + adjoint rollout")
  ap.add_argument("--no-rollout-graph", action="store_true", help="skip the graph-replayed rollout")
  ap.add_argument("--no-hetero", action="store_true", help="skip the heterogeneous (many-model) batch")
  ap.add_argument("--no-hetero-stated", action="store_true",
                  help="skip the heterogeneous batch at its stated size (chains to 32 links, to 64 spheres)")
  ap.add_argument("--no-fallen", action="store_true", help="skip the fallen-humanoid leg")
  ap.add_argument("--fallen-preroll", type=int, default=1000, help="untimed steps before the fallen leg")
  ap.add_argument("--graph", action="store_true", help="replay a HIP graph of one captured step")
  ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                  help="only run a plain copy of MIB MiB (counter calibration) and exit")
  ap.add_argument("--hetero-stated-probe", type=int, default=0, metavar="NWORLD",
                  help="only run a short grouped rollout of the stated-shape heterogeneous workload "
                       "at NWORLD worlds (kernel traces) and exit")
  ap.add_argument("--collision-probe", type=int, default=0, metavar="NWORLD",
                  help="only time the contact-heavy scene's collision at NWORLD worlds and exit")
  args = ap.parse_args()
  if args.copy_probe:
    copy_probe(args.copy_probe)
    return
  if args.hetero_stated_probe:
    hetero_stated_probe(args.hetero_stated_probe)
    return
  if args.collision_probe:
    collision_probe(args.collision_probe)
    return
  if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
    print("bench.py measures one GPU: every box here has one MI355X and the step needs no "
          "communication between GPUs (worlds are independent); run with --gpus 1.", file=sys.stderr)
    sys.exit(2)

  import torch

  assert torch.cuda.is_available(), "bench.py needs the HIP device"
  log(f"fused step: B={args.batch} steps={args.steps} warmup={args.warmup} "
      f"iterations={args.iterations}/{args.ls_iterations}")
  mjm, fused = time_form(args, fused=True)
  log(f"fused: {fused['value']:.4g} world-steps/s, {fused['ms_per_step']:.3f} ms/step")
  plain = None
  if not args.no_plain:
    _, plain = time_form(args, fused=False)
    log(f"plain: {plain['value']:.4g} world-steps/s, {plain['ms_per_step']:.3f} ms/step")
  fallen = None
  if not args.no_fallen:
    # the same workload after 1000 untimed steps: the humanoids have fallen
    # and lie on the floor with 100-150 rows each (the reference benchmark's
    # 1000-step horizon, testspeed.py:42)
    _, ff = time_form(args, fused=True, preroll=args.fallen_preroll)
    fallen = {"preroll_steps": args.fallen_preroll, "value": ff["value"], "ms_per_step": ff["ms_per_step"],
              "kernel_ms": ff["kernel_ms"], "mean_ncon_per_world": ff["ncon_w"],
              "mean_nefc_per_world": ff["nefc_w"], "state_finite": ff["finite"]}
    if not args.no_plain:
      _, fp = time_form(args, fused=False, preroll=args.fallen_preroll)
      fallen["plain_value"] = fp["value"]
      fallen["plain_ms_per_step"] = fp["ms_per_step"]
      fallen["fused_over_plain"] = ff["value"] / fp["value"]
    log(f"fallen (after {args.fallen_preroll} steps): fused {ff['value']:.4g} world-steps/s, "
        f"{ff['nefc_w']:.1f} rows per world" + (f", plain {fallen['plain_value']:.4g}" if "plain_value" in fallen else ""))
  rgraph = None
  if not args.no_rollout_graph:
    rgraph = time_rollout_graph(args)
    log(f"rollout graph: {rgraph['value']:.4g} world-steps/s")
  het = None
  if not args.no_hetero:
    het = time_heterogeneous(args)
    log(f"heterogeneous: grouped {het['grouped']:.4g}, per model {het['per_group']:.4g}, "
        f"grouped graph {het['grouped_graph']:.4g}, grouped work queue {het['grouped_queue']:.4g}, "
        f"autotuned {het['autotuned']:.4g} world-steps/s "
        f"({het['autotuned_choice']}, tuning {het['autotune_s']:.1f} s)")
  het_stated = None
  if not args.no_hetero_stated:
    het_stated = time_heterogeneous_stated(args)
    log(f"heterogeneous at stated size: {het_stated['models']} models, nv {het_stated['nv_range']}, "
        f"{het_stated['rows_per_world']:.1f} rows per world: grouped {het_stated['grouped']:.4g}, "
        f"buckets {het_stated['grouped_buckets']:.4g}, buckets + work queue {het_stated['grouped_buckets_queue']:.4g}, "
        f"{het_stated['steps']}-step graph rollout "
        f"{het_stated['graph_rollout']:.4g} world-steps/s")
  adj = None
  if not args.no_adjoint:
    adj = time_adjoint(args)
    log(f"forward+adjoint: {adj['value']:.4g} world-steps/s (cost ratio {adj['cost_ratio']:.2f}, analytic, "
        f"taped; recomputing {adj['recompute']['value']:.4g}, cost ratio {adj['recompute']['cost_ratio']:.2f})"
        + (f"; finite differences {adj['finite_difference']['value']:.4g} (cost ratio "
           f"{adj['finite_difference']['cost_ratio']:.1f})" if "finite_difference" in adj else ""))
  bytes_ws = algorithmic_bytes(mjm, fused["ncon_w"], fused["nefc_w"])
  achieved = bytes_ws * args.batch / (fused["kernel_ms"] * 1e-3) / 1e9
  base = None
  if not args.no_cpu:
    log("cpu baseline ...")
    base = cpu_baseline(args)
    log(f"cpu: {base['value']:.4g} world-steps/s on {base['cores']} threads")
  out = {
    "metric": METRIC,
    "value": fused["value"],
    "unit": "world-steps/s",
    "n_gpus": 1,
    "steps": args.steps,
    "warmup": args.warmup,
    "ms_per_step": fused["ms_per_step"],
    "higher_is_better": True,
    "scaling": "weak",
    "vs_baseline": None,
    "dtype": "f32",
    "data": "synthetic",
    "config": {"workload": "humanoid 27-dof, fused step, ctrl ~ U(-1,1) resampled per step",
               "nworld": args.batch, "solver": "newton", "iterations": args.iterations,
               "ls_iterations": args.ls_iterations, "integrator": "implicitfast", "timestep": 0.005,
               "gear": 100, "nconmax_per_world": 64, "njmax_per_world": 277,
               "launch": "hip graph replay" if args.graph else "eager"},
    "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args),
                 "kernel": "fused step: description write (with the counter reset), forward launch, "
                           "solve+integrate launch with the big-world launch beside it, tail",
                 "limiter": "latency and instruction issue, not HBM: the step runs at a few percent of the "
                            "HBM peak (dependent per-world chains, pool reservations, the solver's VALU work)",
                 "algorithmic_bytes_per_world_step": bytes_ws, "worlds_per_launch": args.batch,
                 "kernel_ms": fused["kernel_ms"]},
    "cpu_baseline": base,
    "mean_ncon_per_world": fused["ncon_w"],
    "mean_nefc_per_world": fused["nefc_w"],
    "state_finite": fused["finite"],
    "plain_step": None if plain is None else {"value": plain["value"], "ms_per_step": plain["ms_per_step"],
                                              "kernel_ms": plain["kernel_ms"],
                                              "fused_over_plain": fused["value"] / plain["value"]},
    "fallen": fallen,
    "rollout_graph": rgraph,
    "forward_adjoint": adj,
    "heterogeneous": het,
    "heterogeneous_stated_size": het_stated,
  }
  print(json.dumps(out), flush=True)


if __name__ == "__main__":
  main()
