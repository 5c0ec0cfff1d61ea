// Stage functors and the two kernel templates every stage launch uses,
// shared by the kernel translation units (kernels_*.hip), which instantiate
// them, and api.hip, which launches them.
//
// k_stage<TEAM, S>: one stage functor S over teams of TEAM lanes, the model
// and data descriptions passed as kernel arguments.
// k_stage_groups<TEAM, S, TI>: the same over a device table of (model,
// data) groups; TI selects the group's tuning record.
#pragma once

#include <hip/hip_runtime.h>

#include "collision.hpp"
#include "constraint.hpp"
#include "derivative.hpp"
#include "forward.hpp"
#include "smooth.hpp"
#include "solver.hpp"
#include "solver_lds.hpp"
#include "solver_reg.hpp"

namespace mjh {

constexpr int kTeam = 64;

// launch-time tuning of the solvers: the LDS solver's row capacity and
// staged J rows; wide: the Newton Hessians of the 64-lane solvers live in
// efc_h (global) instead of LDS, for models whose nv^2 does not fit;
// tail_rows: the row capacity of the LDS solver in the tail launch (0: the
// tail has only the global-memory solver)
struct Tune {
  int rmax;
  int rj;
  int wide;
  int tail_rows;
};

#define STAGE(NAME, CALL)                                                                              \
  struct NAME {                                                                                        \
    template <int T>                                                                                   \
    MJH_DEV void operator()(const mjhModel& m, const mjhData& d, float* l, const Team<T>& t,          \
                            const Tune& tune) const {                                                  \
      (void)l;                                                                                         \
      (void)tune;                                                                                      \
      CALL;                                                                                            \
    }                                                                                                  \
  };

STAGE(StKinematics, kinematics(m, d, l, t))
STAGE(StComPos, com_pos(m, d, l, t))
STAGE(StCrb, crb(m, d, l, t))
STAGE(StFactorM, factor_m(m, d, l, t))
STAGE(StRne, rne(m, d, l, t))
STAGE(StTransmission, transmission(m, d, t))
STAGE(StComVel, com_vel(m, d, l, t))
STAGE(StPassive, passive(m, d, t))
STAGE(StNxn, nxn_broadphase(m, d, l, t))
STAGE(StSap, sap_broadphase(m, d, l, t))
STAGE(StBroadphase, if (m.ngeom <= 100) nxn_broadphase(m, d, l, t); else sap_broadphase(m, d, l, t))
STAGE(StNarrowphase, narrowphase(m, d, t))
STAGE(StMakeConstraint, make_constraint(m, d, l, t))
STAGE(StFwdVelocity, fwd_velocity(m, d, l, t))
STAGE(StFwdActuation, fwd_actuation(m, d, t))
STAGE(StFwdAcceleration, fwd_acceleration(m, d, l, t))
STAGE(StEuler, euler(m, d, l, t))
STAGE(StImplicit, implicit(m, d, l, t))
// forward.py:676-677 (njmax == 0)
STAGE(StQaccSmooth, if (t.valid) for (int i = t.lane; i < m.nv; i += T) d.qacc[t.world * m.nv + i] =
                        d.qacc_smooth[t.world * m.nv + i])
// Solver of one world: the register solver on 32-lane teams (nv <= 32),
// else the LDS solver; false when the world is left for the tail.
template <int T>
MJH_DEV bool solve_team(const mjhModel& m, const mjhData& d, float* l, const Team<T>& t, const Tune& tune) {
  if constexpr (T == kRegTeam) return solve_reg<2>(m, d, l, tune.rj, t);
  else return solve_lds(m, d, l, tune.rmax, tune.rj, tune.wide != 0, t);
}
// Worlds the 32-lane solve launches leave to the big-world launch: more
// rows than two row sets hold, at most kRegRows.
MJH_DEV bool big_world(const mjhData& d, int w) {
  if (d.njmax == 0) return false;
  const int nr = d.world_efc[2 * w + 1];
  return nr > 2 * kRegTeam && nr <= kRegRows;
}
template <int T>
MJH_DEV bool left_to_big(const mjhData& d, const Team<T>& t) {
  if constexpr (T == kRegTeam) return t.valid && big_world(d, t.world);
  else return false;
}
STAGE(StSolveTeam, if (left_to_big(d, t)) return;
      if (!solve_team(m, d, l, t, tune) && t.lane == 0) d.world_flags[t.world] |= kFlagDeferred)
// big worlds (32-lane teams, big_world): the register solver with kRegSets
// row sets; with `step` set also what the fused solve launch does around it
// (factorisation, smooth acceleration, integrator).  It touches no other
// world, so it can run beside the solve launch, which skips these worlds.
template <bool step>
struct StBigWorlds {
  template <int T>
  MJH_DEV void operator()(const mjhModel& m, const mjhData& d, float* l, const Team<T>& t, const Tune& tune) const {
    static_assert(T == kRegTeam, "big worlds run on 32-lane teams");
    if (!t.valid || !big_world(d, t.world)) return;
    if (step) {
      factor_m(m, d, l, t);
      sync();
      fwd_acceleration_solve(m, d, l, t);
    }
    solve_reg<kRegSets>(m, d, l, tune.rj, t);
    sync();
    if (step) {
      if (m.opt.integrator == mjhINT_EULER) euler(m, d, l, t);
      else implicit(m, d, l, t);
    }
  }
};

// tail: flagged worlds get the LDS solver with a large row capacity or,
// past it, the global-memory solver (and, when `integrate` is set, the
// integrator), then the flag is cleared
template <bool integrate>
struct StTail {
  template <int T>
  MJH_DEV void operator()(const mjhModel& m, const mjhData& d, float* l, const Team<T>& t, const Tune& tune) const {
    if (!t.valid || !(d.world_flags[t.world] & kFlagDeferred)) return;
    // the LDS solver with the tail's row capacity, else the global one
    bool solved = false;
    if constexpr (T == kTeam) solved = tune.tail_rows > 0 && solve_lds(m, d, l, tune.tail_rows, 0, tune.wide != 0, t);
    if (!solved) solve(m, d, l, t, tune.wide != 0);
    sync();
    if (integrate) {
      if (m.opt.integrator == mjhINT_EULER) euler(m, d, l, t);
      else implicit(m, d, l, t);
    }
    if (t.lane == 0) d.world_flags[t.world] &= ~kFlagDeferred;
  }
};

// Collision and constraint rows inside a fused launch: for nxn models the
// three pool reservations are in flight together (collide_and_reserve),
// SAP models take the stage functions one after the other.
template <int T>
MJH_DEV void fused_collision_rows(const mjhModel& m, const mjhData& d, float* l, const Team<T>& t) {
  if (m.ngeom <= 100) {
    int rbase, rtotal, nlim;
    collide_and_reserve(m, d, l, t, &rbase, &rtotal, &nlim);
    if (!(m.opt.disableflags & mjhDSBL_CONSTRAINT)) make_constraint_rows(m, d, l, t, rbase, rtotal, nlim);
  } else {
    collision(m, d, l, t);
    make_constraint(m, d, l, t);
  }
}

// The whole step of one world (forward.py:682-695 / 664-679).
struct StFusedStep {
  template <int T>
  MJH_DEV void operator()(const mjhModel& m, const mjhData& d, float* l, const Team<T>& t, const Tune& tune) const {
    kinematics(m, d, l, t);
    com_pos(m, d, l, t);
    crb(m, d, l, t);
    factor_m(m, d, l, t);
    sync();
    fused_collision_rows(m, d, l, t);
    sync();
    transmission(m, d, t);
    fwd_velocity(m, d, l, t);
    fwd_actuation(m, d, t);
    fwd_acceleration(m, d, l, t);
    if (d.njmax == 0) {
      if (t.valid)
        for (int i = t.lane; i < m.nv; i += T) d.qacc[t.world * m.nv + i] = d.qacc_smooth[t.world * m.nv + i];
      sync();
    } else if (!solve_team(m, d, l, t, tune)) {
      if (t.lane == 0) d.world_flags[t.world] |= kFlagDeferred;
      return;  // solved and integrated by the tail launch
    }
    sync();
    if (m.opt.integrator == mjhINT_EULER) euler(m, d, l, t);
    else implicit(m, d, l, t);
  }
};

// The fused step in two launches (32-lane teams, and 64-lane teams through
// the group kernels): the forward stages
// that walk the body tree, collide and build constraint rows (up to
// qfrc_smooth), then everything that factors or solves with the mass
// matrix: factor_m, qacc_smooth, the solver and the integrator.  The second
// group keeps a world's matrix rows in VGPRs; one kernel holding both
// groups would impose that register budget on the whole step.
struct StFusedForward {
  template <int T>
  MJH_DEV void operator()(const mjhModel& m, const mjhData& d, float* l, const Team<T>& t, const Tune&) const {
    kinematics(m, d, l, t);
    com_pos(m, d, l, t);
    crb(m, d, l, t);
    sync();
    fused_collision_rows(m, d, l, t);
    sync();
    transmission(m, d, t);
    fwd_velocity(m, d, l, t);
    fwd_actuation(m, d, t);
    fwd_acceleration_force(m, d, t);
  }
};
struct StFusedSolveIntegrate {
  template <int T>
  MJH_DEV void operator()(const mjhModel& m, const mjhData& d, float* l, const Team<T>& t, const Tune& tune) const {
    if (left_to_big(d, t)) return;  // StBigWorlds<true> steps it
    factor_m(m, d, l, t);
    sync();
    fwd_acceleration_solve(m, d, l, t);
    if (d.njmax == 0) {
      if (t.valid)
        for (int i = t.lane; i < m.nv; i += T) d.qacc[t.world * m.nv + i] = d.qacc_smooth[t.world * m.nv + i];
      sync();
    } else if (!solve_team(m, d, l, t, tune)) {
      if (t.lane == 0) d.world_flags[t.world] |= kFlagDeferred;
      return;  // solved and integrated by the tail launch
    }
    sync();
    if (m.opt.integrator == mjhINT_EULER) euler(m, d, l, t);
    else implicit(m, d, l, t);
  }
};

// The fused step on 64-lane teams in three launches: the forward stages
// with the mass-matrix factor and qacc_smooth, the solver, the integrator.
// Each launch takes the LDS its own stages need; the solver's LDS is what
// limits its occupancy, and the stages around it keep theirs (one launch
// for the whole step ran them at the solver's: the three attached
// humanoids stepped 20% slower than per-stage launches).  One world per
// workgroup, so the per-world early returns are uniform per workgroup.
struct StForwardFactor {
  template <int T>
  MJH_DEV void operator()(const mjhModel& m, const mjhData& d, float* l, const Team<T>& t, const Tune& tune) const {
    StFusedForward()(m, d, l, t, tune);
    sync();
    factor_m(m, d, l, t);
    sync();
    fwd_acceleration_solve(m, d, l, t);
  }
};
struct StSolveStep {
  template <int T>
  MJH_DEV void operator()(const mjhModel& m, const mjhData& d, float* l, const Team<T>& t, const Tune& tune) const {
    if (d.njmax == 0) {
      if (t.valid)
        for (int i = t.lane; i < m.nv; i += T) d.qacc[t.world * m.nv + i] = d.qacc_smooth[t.world * m.nv + i];
      return;
    }
    if (!solve_team(m, d, l, t, tune) && t.lane == 0) d.world_flags[t.world] |= kFlagDeferred;
  }
};
// worlds the solve launch deferred are integrated by the tail launch
struct StIntegrateSolved {
  template <int T>
  MJH_DEV void operator()(const mjhModel& m, const mjhData& d, float* l, const Team<T>& t, const Tune&) const {
    static_assert(T == 64, "one world per workgroup");
    if (t.valid && (d.world_flags[t.world] & kFlagDeferred)) return;
    if (m.opt.integrator == mjhINT_EULER) euler(m, d, l, t);
    else implicit(m, d, l, t);
  }
};

template <int TEAM, class S>
__global__ void __launch_bounds__(64) k_stage(mjhModel m, mjhData d, int lds_per_world, Tune tune) {
  static_assert(64 % TEAM == 0, "teams tile the wavefront");
  extern __shared__ float smem[];
  Team<TEAM> t(d.nworld);
  S()(m, d, smem + t.slot * lds_per_world, t, tune);
}

// counter reset without memset nodes (capture-safe):
// which bit 0: ncollision + ncon, bit 1: nefc
MJH_DEV void reset_counters(const mjhData& d, int which) {
  if (which & 1) {
    d.ncollision[0] = 0;
    d.ncon[0] = 0;
  }
  if (which & 2) d.nefc[0] = 0;
}

// ---- heterogeneous batches: several (model, data) groups, each with its
// own model and worlds, advanced by one launch per stage over all groups ----

// One group as a launch sees it: the group's workgroups are
// [first_block, first_block + its block count) of the grid.
struct GroupDesc {
  mjhModel m;
  mjhData d;
  int first_block;
  Tune tune[2];  // [0] the solve / tail launches, [1] the big-world launch
};

// The group of workgroup b: the last with first_block <= b (uniform per
// workgroup; first_block ascends and the first is 0).
MJH_DEV const GroupDesc& group_of(const GroupDesc* __restrict__ g, int ngroups, int b) {
  int lo = 0, hi = ngroups - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (g[mid].first_block <= b) lo = mid;
    else hi = mid - 1;
  }
  return g[lo];
}

template <int TEAM, class S, int TI>
__global__ void __launch_bounds__(64) k_stage_groups(const GroupDesc* __restrict__ g, int ngroups, int lds_per_world) {
  extern __shared__ float smem[];
  const int b = blockIdx.x;
  const GroupDesc& G = group_of(g, ngroups, b);
  Team<TEAM> t(G.d.nworld, b - G.first_block);
  S()(G.m, G.d, smem + t.slot * lds_per_world, t, G.tune[TI]);
}

// The work-queue form of k_stage_groups: a grid of at most a few
// workgroups per SIMD (a multiple of kQueueLanes), each taking the next
// block index from one of kQueueLanes counters in device memory (queue,
// kQueueStride ints apart, zeroed before the launch) until all nblocks are
// taken.  Workgroup w draws from counter c = w % kQueueLanes, whose tickets
// i name blocks c + kQueueLanes * i: every block once, each counter's in
// index order, largest models first.  Atomics on one address serialise
// (about 16 ns each here), so the tickets are spread over counters on
// separate 256-byte lines.  Every workgroup exits once its counter passes
// nblocks.  One wavefront per workgroup, all lanes active at the read:
// lane 0's ticket is taken with readfirstlane, which puts it in a scalar
// register, so the group lookup and the model description are scalar
// loads as with blockIdx.x (a shuffle leaves it in a vector register and
// every model pointer with it).
constexpr int kQueueLanes = 8;
constexpr int kQueueStride = 64;
template <int TEAM, class S, int TI>
__global__ void __launch_bounds__(64) k_stage_queue(const GroupDesc* __restrict__ g, int ngroups, int lds_per_world,
                                                    int* queue, int nblocks) {
  extern __shared__ float smem[];
  const int c = blockIdx.x % kQueueLanes;
  int* counter = queue + c * kQueueStride;
  for (;;) {
    int ticket = 0;
    if (threadIdx.x == 0) ticket = atomicAdd(counter, 1);
    const int b = c + kQueueLanes * __builtin_amdgcn_readfirstlane(ticket);
    if (b >= nblocks) return;
    const GroupDesc& G = group_of(g, ngroups, b);
    Team<TEAM> t(G.d.nworld, b - G.first_block);
    S()(G.m, G.d, smem + t.slot * lds_per_world, t, G.tune[TI]);
    __syncthreads();
  }
}

using StBigWorldsSolve = StBigWorlds<false>;
using StBigWorldsStep = StBigWorlds<true>;
using StTailSolve = StTail<false>;
using StTailStep = StTail<true>;

// ---- explicit instantiations ----
// kernels.hip instantiates one list below per compilation (-DMJH_KERNEL_LIST
// = the list's name; build.py compiles it once per list, in parallel) and
// api.hip declares every list extern.  X(TEAM, S) for k_stage,
// G(TEAM, S, TI) for k_stage_groups.  The register solver's instantiations
// take the longest to compile, so each sits in a list of its own.

// tree passes and smooth forces, both team sizes
#define MJH_KERNELS_TREE(X, G)                                                                    \
  X(32, StKinematics) X(64, StKinematics) X(32, StComPos) X(64, StComPos) X(32, StCrb) X(64, StCrb) \
  X(32, StFactorM) X(64, StFactorM) X(32, StRne) X(64, StRne) X(32, StComVel) X(64, StComVel)       \
  X(32, StTransmission) X(64, StTransmission) X(32, StPassive) X(64, StPassive)                     \
  X(32, StFwdVelocity) X(64, StFwdVelocity) X(32, StFwdActuation) X(64, StFwdActuation)             \
  X(32, StFwdAcceleration) X(64, StFwdAcceleration) X(32, StQaccSmooth) X(64, StQaccSmooth)
// broadphase, narrowphase, constraint rows
#define MJH_KERNELS_COLLISION(X, G)                                                              \
  X(32, StNxn) X(64, StNxn) X(32, StSap) X(64, StSap) X(32, StBroadphase) X(64, StBroadphase)     \
  X(32, StNarrowphase) X(64, StNarrowphase) X(32, StMakeConstraint) X(64, StMakeConstraint)
// integrators, the LDS solver and the global-memory tail
#define MJH_KERNELS_INTEGRATE(X, G) \
  X(32, StEuler) X(64, StEuler) X(32, StImplicit) X(64, StImplicit) X(64, StSolveTeam) X(64, StTailSolve)
// the register solver launches of the plain step
#define MJH_KERNELS_SOLVE(X, G) X(32, StSolveTeam) G(32, StSolveTeam, 0)
#define MJH_KERNELS_BIG_SOLVE(X, G) X(32, StBigWorldsSolve)
#define MJH_KERNELS_BIG_SOLVE_G(X, G) G(32, StBigWorldsSolve, 1)
// the fused step's launches
#define MJH_KERNELS_FUSED_FORWARD(X, G) X(32, StFusedForward) G(32, StFusedForward, 0)
#define MJH_KERNELS_FUSED_SOLVE(X, G) X(32, StFusedSolveIntegrate)
#define MJH_KERNELS_FUSED_SOLVE_G(X, G) G(32, StFusedSolveIntegrate, 0)
#define MJH_KERNELS_BIG_STEP(X, G) X(32, StBigWorldsStep)
#define MJH_KERNELS_BIG_STEP_G(X, G) G(32, StBigWorldsStep, 1)
#define MJH_KERNELS_TAIL_STEP(X, G) X(64, StTailStep) G(64, StTailStep, 0)
#define MJH_KERNELS_FUSED64(X, G) X(64, StFusedStep) G(64, StFusedStep, 0)
#define MJH_KERNELS_SPLIT64(X, G) G(64, StForwardFactor, 0) G(64, StSolveStep, 0) G(64, StIntegrateSolved, 0)

#define MJH_KERNEL_LISTS(X, G)                                                                          \
  MJH_KERNELS_TREE(X, G) MJH_KERNELS_COLLISION(X, G) MJH_KERNELS_INTEGRATE(X, G) MJH_KERNELS_SOLVE(X, G) \
  MJH_KERNELS_BIG_SOLVE(X, G) MJH_KERNELS_BIG_SOLVE_G(X, G) MJH_KERNELS_FUSED_FORWARD(X, G)             \
  MJH_KERNELS_FUSED_SOLVE(X, G) MJH_KERNELS_FUSED_SOLVE_G(X, G) MJH_KERNELS_BIG_STEP(X, G)              \
  MJH_KERNELS_BIG_STEP_G(X, G) MJH_KERNELS_TAIL_STEP(X, G) MJH_KERNELS_FUSED64(X, G) MJH_KERNELS_SPLIT64(X, G)

#define MJH_INSTANTIATE_STAGE(TEAM, S) template __global__ void k_stage<TEAM, S>(mjhModel, mjhData, int, Tune);
#define MJH_INSTANTIATE_GROUPS(TEAM, S, TI)                                                   \
  template __global__ void k_stage_groups<TEAM, S, TI>(const GroupDesc* __restrict__, int, int); \
  template __global__ void k_stage_queue<TEAM, S, TI>(const GroupDesc* __restrict__, int, int, int*, int);
#define MJH_EXTERN_STAGE(TEAM, S) extern MJH_INSTANTIATE_STAGE(TEAM, S)
#define MJH_EXTERN_GROUPS(TEAM, S, TI)                                                               \
  extern template __global__ void k_stage_groups<TEAM, S, TI>(const GroupDesc* __restrict__, int, int); \
  extern template __global__ void k_stage_queue<TEAM, S, TI>(const GroupDesc* __restrict__, int, int, int*, int);

}  // namespace mjh
