"""The testspeed command line and `benchmark` (testspeed.py, test_util.py:88-120).

CPU: flags and defaults are the reference's, the function choices are the
package's (m, d) functions, traces sum element-wise.  GPU: `benchmark`
returns the reference's tuple from a graph replay (and from eager traced
calls), and the command line prints the reference's report, event trace
and allocation tables.
"""

import pytest

import mjx_warp_amd as mjw
from mjx_warp_amd import mjcf, models, test_util, testspeed


def test_flags_and_defaults():
  a = testspeed._parser().parse_args(["--mjcf=humanoid.xml"])
  assert (a.function, a.nstep, a.batch_size, a.solver, a.iterations, a.ls_iterations) == (
    "step", 1000, 4096, "newton", 1, 4)
  assert (a.ls_parallel, a.is_sparse, a.nconmax, a.njmax, a.keyframe, a.output) == (False, False, -1, -1, 0, "text")
  assert not a.event_trace and not a.measure_alloc
  a = testspeed._parser().parse_args(["--mjcf", "x.xml", "--event_trace", "--measure_alloc=false", "--solver=cg"])
  assert a.event_trace and not a.measure_alloc and a.solver == "cg"
  with pytest.raises(SystemExit):
    testspeed._parser().parse_args(["--mjcf=x.xml", "--function=put_model"])


def test_function_choices():
  names = testspeed._functions()
  for n in ("step", "forward", "step_fused", "kinematics", "collision", "solve", "euler", "implicit"):
    assert n in names
  for n in ("put_model", "make_data", "benchmark", "mul_m", "solve_m"):
    assert n not in names


def test_trace_sum():
  a = {"step": ([1.0, 2.0], {"forward": ([0.5, 1.0], {})})}
  b = {"step": ([3.0, 4.0], {"forward": ([1.5, 2.0], {})})}
  assert test_util._sum(a, b) == {"step": ([4.0, 6.0], {"forward": ([2.0, 3.0], {})})}


def _humanoid_data(mjm):
  d = mjcf.MjData(mjm)
  mjcf.mj_resetDataKeyframe(mjm, d, 0)
  test_util.forward_host(mjm, d)
  return d


@pytest.mark.gpu
@pytest.mark.parametrize("event_trace", [False, True])
def test_benchmark_tuple(event_trace):
  mjm = models.humanoid()
  d = _humanoid_data(mjm)
  assert d.ncon > 0 and d.nefc >= 4 * d.ncon
  jit, run, trace, steps, ncon, nefc = mjw.benchmark(mjw.step, mjm, d, nstep=3, batch_size=64, solver="newton",
                                                       iterations=10, ls_iterations=20, event_trace=event_trace,
                                                       measure_alloc=True)
  assert jit > 0 and run > 0 and steps == 3 * 64
  assert len(ncon) == 3 and len(nefc) == 3 and all(c > 0 for c in ncon)
  assert mjm.opt.iterations == 10 and mjm.opt.ls_iterations == 20
  if event_trace:
    times, sub = trace["step"]
    assert len(times) == 1 and times[0] > 0
    assert set(sub) == {"forward", "implicit"}
  else:
    assert trace == {}


@pytest.mark.gpu
def test_command_line_report(capsys):
  assert testspeed.main(["--mjcf=humanoid.xml", "--batch_size=32", "--nstep=4", "--event_trace",
                         "--measure_alloc", "--iterations=10", "--ls_iterations=20"]) == 0
  out = capsys.readouterr().out
  for s in ("Model nbody: 17 nv: 27 ngeom: 20", "Summary for 32 parallel rollouts", "Total steps per second",
            "Event trace:", "step:", "  forward:", "    fwd_position:", "ncon alloc:", "nefc alloc:"):
    assert s in out, s
  assert testspeed.main(["--mjcf=humanoid.xml", "--batch_size=32", "--nstep=4", "--output=tsv",
                         "--function=step_fused"]) == 0
  assert "steps/second:" in capsys.readouterr().out
