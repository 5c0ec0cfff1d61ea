// Kernels and extern "C" entry points of libmjx_warp_amd.so.
//
// Every stage kernel is launched with 64-lane workgroups (one wavefront),
// each holding 64/TEAM worlds; the grid covers nworld.  The plain step
// (mjh_step) launches one kernel per reference stage; the fused step
// (mjh_step_fused) runs every stage of a world inside one launch (two on
// 32-lane teams).  Both use the register (nv <= 32) or LDS solver; on
// 32-lane teams worlds with 65-256 rows are solved by a second launch beside
// the main solve launch, and worlds over every capacity are flagged and
// finished by a tail launch of the global-memory solver (and, for the fused
// step, the integrator).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>
#include "adjoint.hpp"
#include "stages.hpp"

namespace mjh {
MJH_KERNEL_LISTS(MJH_EXTERN_STAGE, MJH_EXTERN_GROUPS)
}  // namespace mjh

using namespace mjh;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
__global__ void k_reset(mjhData d, int which) {
  if (threadIdx.x == 0) {
    reset_counters(d, which);
  }
}

// res = M vec / x = inv(M) y / qfrc += J' xfrc for caller-provided vectors
__global__ void __launch_bounds__(64) k_mul_m(mjhModel m, mjhData d, float* res, const float* vec,
                                              const uint8_t* skip) {
  Team<kTeam> t(d.nworld);
  if (!t.valid) return;
  if (skip && skip[t.world]) return;
  const size_t o = (size_t)t.world * m.nv;
  mul_m(m, d.qMd + o * m.nv, res + o, vec + o, t);
}
__global__ void __launch_bounds__(64) k_solve_m(mjhModel m, mjhData d, float* x, const float* y, int ldsw) {
  extern __shared__ float smem[];
  Team<kTeam> t(d.nworld);
  if (!t.valid) return;
  const size_t o = (size_t)t.world * m.nv;
  solve_blocks(m, d.qLDd + o * m.nv, x + o, y + o, smem + t.slot * ldsw, t);
}
__global__ void __launch_bounds__(64) k_xfrc(mjhModel m, mjhData d, float* qfrc) {
  Team<kTeam> t(d.nworld);
  xfrc_accumulate(m, d, qfrc, t);
}

// ---- constraint-row pools laid out by the caller ----
// The reference appends rows to its pool with atomics (constraint.py:119,
// 162), so a world's rows may sit anywhere in [0, nefc), tagged by
// efc_worldid (-1 for unused rows); the solver kernels here read a world's
// rows as the block world_efc[w] = (first, count).  mjh_solve derives the
// blocks from efc_worldid before solving: a world whose rows are contiguous
// gets (first, count); a world whose rows are interleaved with others' is
// gathered, in row order, into the pool's unused capacity [nefc, njmax),
// solved there, and its efc_force / efc_active / efc_Jaref are scattered
// back to its own rows afterwards.  world_flags bit 1 marks gathered
// worlds; the bits above bit 1 serve as a per-world counter meanwhile.
constexpr int kFlagGathered = 2;
constexpr int kFlagCount = 4;

MJH_DEV int pool_rows(const mjhData& d) { return min(max(d.nefc[0], 0), d.njmax); }

// per world: clear the block and the counter bits (keeps the flags)
__global__ void k_blocks_zero(mjhData d) {
  for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < d.nworld; w += gridDim.x * blockDim.x) {
    d.world_efc[2 * w] = 0;
    d.world_efc[2 * w + 1] = 0;
    d.world_flags[w] &= kFlagDeferred;
  }
}
// per row: count each world's rows and its runs of consecutive rows
__global__ void k_blocks_count(mjhData d) {
  const int n = pool_rows(d);
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const int w = d.efc_worldid[r];
    if (w < 0 || w >= d.nworld) continue;
    atomicAdd(d.world_efc + 2 * w + 1, 1);
    if (r == 0 || d.efc_worldid[r - 1] != w) {
      atomicAdd(d.world_flags + w, kFlagCount);
      d.world_efc[2 * w] = r;  // the run's first row (the only run of a contiguous world)
    }
  }
}
// one workgroup of 1024: worlds with more than one run are gathered; their
// blocks are laid out back to back after the pool's rows (an exclusive scan
// over worlds in chunks of 1024), as far as the capacity goes
__global__ void __launch_bounds__(1024) k_blocks_place(mjhData d) {
  __shared__ int part[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = pool_rows(d);
  __syncthreads();
  for (int w0 = 0; w0 < d.nworld; w0 += 1024) {
    const int w = w0 + (int)threadIdx.x;
    int f = 0, n = 0;
    bool gather = false;
    if (w < d.nworld) {
      f = d.world_flags[w];
      gather = f / kFlagCount > 1;
      n = gather ? d.world_efc[2 * w + 1] : 0;
    }
    part[threadIdx.x] = n;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
      const int v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
      __syncthreads();
      part[threadIdx.x] += v;
      __syncthreads();
    }
    const int start = carry + part[threadIdx.x] - n;
    if (w < d.nworld) {
      f &= kFlagDeferred;
      if (gather) {
        const int fit = max(0, min(n, d.njmax - start));
        if (fit < n) atomicAdd(d.overflow + 2, n - fit);
        d.world_efc[2 * w] = start;
        d.world_efc[2 * w + 1] = fit;
        f |= kFlagGathered;
      }
      d.world_flags[w] = f;
    }
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
}
// per row of a gathered world: its source index into the world's new
// block (efc_worldid of the unused capacity), in arrival order
__global__ void k_blocks_slots(mjhData d) {
  const int n = pool_rows(d);
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const int w = d.efc_worldid[r];
    if (w < 0 || w >= d.nworld || !(d.world_flags[w] & kFlagGathered)) continue;
    const int slot = atomicAdd(d.world_flags + w, kFlagCount) / kFlagCount;
    if (slot < d.world_efc[2 * w + 1]) d.efc_active[d.world_efc[2 * w] + slot] = r;
  }
}
// one wavefront per gathered world: the source rows in ascending order
// (rank sort from efc_active into efc_worldid of the new block), then the
// rows' J, D and aref copied in that order; efc_worldid of the new block
// keeps the source rows for the scatter
__global__ void __launch_bounds__(64) k_blocks_gather(mjhData d, int nv) {
  const int w = blockIdx.x;
  if (w >= d.nworld || !(d.world_flags[w] & kFlagGathered)) return;
  const int b = d.world_efc[2 * w], n = d.world_efc[2 * w + 1];
  const int* src = d.efc_active + b;
  for (int i = threadIdx.x; i < n; i += 64) {
    const int v = src[i];
    int rank = 0;
    for (int j = 0; j < n; j++) rank += src[j] < v;
    d.efc_worldid[b + rank] = v;
  }
  __syncthreads();
  for (int k = 0; k < n; k++) {
    const int r = d.efc_worldid[b + k];
    for (int i = threadIdx.x; i < nv; i += 64) d.efc_J[(size_t)(b + k) * nv + i] = d.efc_J[(size_t)r * nv + i];
    if (threadIdx.x == 0) {
      d.efc_D[b + k] = d.efc_D[r];
      d.efc_aref[b + k] = d.efc_aref[r];
      d.efc_pos[b + k] = d.efc_pos[r];
      d.efc_margin[b + k] = d.efc_margin[r];
    }
  }
  if (threadIdx.x == 0) d.world_flags[w] = (d.world_flags[w] & kFlagDeferred) | kFlagGathered;
}
// after the solve: the gathered rows' results back to the caller's rows
__global__ void __launch_bounds__(64) k_blocks_scatter(mjhData d) {
  const int w = blockIdx.x;
  if (w >= d.nworld || !(d.world_flags[w] & kFlagGathered)) return;
  const int b = d.world_efc[2 * w], n = d.world_efc[2 * w + 1];
  for (int k = threadIdx.x; k < n; k += 64) {
    const int r = d.efc_worldid[b + k];
    d.efc_force[r] = d.efc_force[b + k];
    d.efc_active[r] = d.efc_active[b + k];
    d.efc_Jaref[r] = d.efc_Jaref[b + k];
  }
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += 64) d.efc_worldid[b + k] = w;
  if (threadIdx.x == 0) d.world_flags[w] &= ~kFlagGathered;
}

// transition derivatives and the adjoint recursion
__global__ void __launch_bounds__(64) k_fd_perturb(mjhModel m, mjhData d, mjhData fd, float eps, int centered) {
  Team<kTeam> t(fd.nworld);
  fd_perturb(m, d, fd, eps, centered, t);
}
__global__ void __launch_bounds__(64) k_fd_difference(mjhModel m, mjhData fd, int nworld, float eps, int centered,
                                                      float* A, float* B) {
  Team<kTeam> t(nworld * (fd_nx(m) + m.nu));
  fd_difference(m, fd, nworld, eps, centered, A, B, t);
}
__global__ void __launch_bounds__(64) k_adjoint_step(mjhModel m, mjhData d, const float* A, const float* B,
                                                     const float* g_qpos, const float* g_qvel, const float* g_act,
                                                     const float* lam_next, float* lam, float* grad_ctrl) {
  Team<kTeam> t(d.nworld);
  adjoint_step(m, d, A, B, g_qpos, g_qvel, g_act, lam_next, lam, grad_ctrl, t);
}
__global__ void __launch_bounds__(64) k_adjoint_vjp(mjhModel m, mjhData d, const float* qpos_t, const float* qvel_t,
                                                    const float* lam_next, const float* gadd, float* lam,
                                                    float* grad_ctrl, int rcap) {
  extern __shared__ float smem[];
  Team<kTeam> t(d.nworld);
  adjoint_vjp(m, d, qpos_t, qvel_t, lam_next, gadd, lam, grad_ctrl, rcap, smem, t);
}
// The adjoint's tape: copies of the Data fields the step adjoint reads, one
// slot per step.  A segment is a field's words; a pool field copies only
// its first count (device counter, clamped to the capacity) entries.
struct TapeSeg {
  const char* src;
  unsigned long long dst;    // byte offset in the slot
  unsigned long long bytes;  // capacity
  const int* count;          // entries in use (pools), or null
  unsigned int row;          // bytes per entry (pools)
};
constexpr int kTapeSegsMax = 40;
struct TapeSegs {
  TapeSeg s[kTapeSegsMax];
  int n;
};
__global__ void __launch_bounds__(256) k_tape_copy(TapeSegs segs, char* slot) {
  const TapeSeg& g = segs.s[blockIdx.y];
  unsigned long long nb = g.bytes;
  if (g.count) {
    const long long c = *g.count;
    const unsigned long long used = c > 0 ? (unsigned long long)c * g.row : 0ull;
    if (used < nb) nb = used;
  }
  const unsigned int* src = reinterpret_cast<const unsigned int*>(g.src);
  unsigned int* dst = reinterpret_cast<unsigned int*>(slot + g.dst);
  const unsigned long long n4 = nb / 4;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (unsigned long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
__global__ void __launch_bounds__(64) k_adjoint_state_grad(mjhModel m, mjhData d, const float* lam, float* g_qpos,
                                                           float* g_qvel, float* g_act) {
  Team<kTeam> t(d.nworld);
  adjoint_state_grad(m, d, lam, g_qpos, g_qvel, g_act, t);
}

// per-block linear algebra (factor_m, solve_m, the integrators): the
// largest kinematic tree's block and right-hand side staged in LDS, and
// MuJoCo's sparse factor (nM) for sparse models
int lds_linalg(const mjhModel& m) {
  const int b = m.tree_dof_max > 0 ? m.tree_dof_max : m.nv;
  const int a = b * b + b;
  return m.opt.is_sparse && m.nM > a ? m.nM : a;
}
// the global-memory solver of the tail launches: per-block linear algebra
// and its Newton Hessian with right-hand side (nv^2 + nv), or when wide (H
// in efc_h) the right-hand side and the blocked Cholesky's panel buffer
int lds_tail(const mjhModel& m, bool wide) {
  const int a = lds_linalg(m), h = wide ? m.nv + chol_blocked_floats(m.nv) : m.nv * m.nv + m.nv;
  return a > h ? a : h;
}
int lds_sap(const mjhModel& m) { return lds_sap_words(m.ngeom); }

size_t lds_cap();
bool reg_path(const mjhModel& m);
int lds_forward(const mjhModel& m, bool reg);
constexpr int kLdsRowsMin = 256, kLdsRowsMax = 1024;

// Row capacity of the LDS solver (rows of one world) and how many of them
// get their Jacobian rows staged in LDS, for a model on 32-lane (reg) or
// 64-lane teams.  On the 64-lane path no rows are staged (the LDS solver
// stages a world's rows only when all of them fit, and the LDS that takes
// costs occupancy: the three attached humanoids step 1.25x faster without
// it); the Newton Hessian moves to efc_h (wide, factored there in LDS
// panels) when with it in LDS next to 256 rows a CU would hold fewer than
// five workgroups (three attached humanoids, nv 81: 1.03e6 world-steps/s
// wide, 0.89e6 in LDS; a 40-link chain, nv 40: 2.3e6 wide, 2.75e6 in LDS);
// and the row capacity fills what
// the solver's LDS leaves before a CU would hold one workgroup fewer (a CU
// holds lds_cap() bytes: 160 KiB on gfx950): rows cost 8 floats each, and a
// world with more rows than the capacity is solved by the slower
// global-memory tail.  MJH_SOLVER_RMAX / MJH_SOLVER_RJ override the
// defaults, MJH_SOLVER_WIDE=1 / 0 puts the Hessian of any 64-lane model in
// efc_h / in LDS (where it fits).
Tune tune_for(const mjhModel& m, bool reg) {
  Tune t{64, reg ? 32 : 0, 0};
  if (!reg) {
    t.rmax = kLdsRowsMin;
    const int cap = (int)(lds_cap() / sizeof(float));
    const bool fits = lds_solver_floats(m.nv, kLdsRowsMin, t.rj) <= cap && lds_tail(m, false) <= cap;
    t.wide = !fits || lds_solver_floats(m.nv, kLdsRowsMin, t.rj) > cap / 5;
    if (const char* e = std::getenv("MJH_SOLVER_WIDE")) {
      if (std::strcmp(e, "1") == 0) t.wide = 1;
      if (std::strcmp(e, "0") == 0 && fits) t.wide = 0;
    }
    int base = lds_solver_floats(m.nv, 0, t.rj, t.wide);
    const int fwd = lds_forward(m, false);
    if (fwd > base) base = fwd;
    if (base + 8 * kLdsRowsMin <= cap) {
      const int per_cu = cap / (base + 8 * kLdsRowsMin);
      const int rows = (cap / per_cu - lds_solver_floats(m.nv, 0, t.rj, t.wide)) / 8;
      t.rmax = rows < kLdsRowsMin ? kLdsRowsMin : rows > kLdsRowsMax ? kLdsRowsMax : rows;
    }
  }
  if (const char* e = std::getenv("MJH_SOLVER_RMAX")) t.rmax = std::atoi(e);
  if (const char* e = std::getenv("MJH_SOLVER_RJ")) t.rj = std::atoi(e);
  if (t.rmax < 0) t.rmax = 0;
  if (t.rj > t.rmax) t.rj = t.rmax;
  if (t.rj < 0) t.rj = 0;
  // on the 64-lane path the tail launch (the few worlds over rmax) takes the
  // LDS solver with kLdsRowsMax rows when that fits one workgroup; 32-lane
  // models reach the tail only past the big-world launch's 256 rows, rarely
  // enough to leave it the global-memory solver and its small LDS
  t.tail_rows = !reg && lds_solver_floats(m.nv, kLdsRowsMax, 0, t.wide != 0) <= (int)(lds_cap() / sizeof(float))
                    ? kLdsRowsMax
                    : 0;
  return t;
}
Tune tune_for(const mjhModel& m) { return tune_for(m, reg_path(m)); }
// the tail launch: the global-memory solver's LDS and the LDS solver's
// with the tail's row capacity
int lds_tail(const mjhModel& m, const Tune& t) {
  const int a = lds_tail(m, t.wide != 0);
  const int b = t.tail_rows > 0 ? lds_solver_floats(m.nv, t.tail_rows, 0, t.wide != 0) : 0;
  return a > b ? a : b;
}
int lds_solve(const mjhModel& m, const Tune& t) {
  const int a = lds_solver_floats(m.nv, t.rmax, t.rj, t.wide != 0), b = lds_linalg(m);
  return a > b ? a : b;
}
int lds_solve(const mjhModel& m) { return lds_solve(m, tune_for(m)); }
int lds_kin(const mjhModel& m) { return lds_kinematics_floats(m.nbody); }
int lds_crb(const mjhModel& m, bool reg) { return lds_crb_floats(m.nbody, m.nv, reg); }
int lds_vel(const mjhModel& m) { return lds_velocity_floats(m.nbody); }
int lds_broadphase(const mjhModel& m) { return m.ngeom > 100 ? lds_sap(m) : kNxnWords * m.ngeom; }
int lds_con(bool reg) { return reg ? lds_constraint_floats<kRegTeam>() : lds_constraint_floats<kTeam>(); }
// every forward stage of the fused forward launch, on 32-lane (reg) or
// 64-lane teams
int lds_forward(const mjhModel& m, bool reg) {
  int a = lds_kin(m);
  const int b = lds_crb(m, reg), c = lds_vel(m), g = lds_broadphase(m), k = lds_con(reg);
  a = a > b ? a : b;
  a = a > c ? a : c;
  a = a > k ? a : k;
  return a > g ? a : g;
}
int lds_forward(const mjhModel& m) { return lds_forward(m, reg_path(m)); }
// the 64-lane fused step (forward, LDS solver, integrator in one launch)
int lds_step(const mjhModel& m, const Tune& t) {
  const int a = lds_solve(m, t), b = lds_forward(m, false);
  return a > b ? a : b;
}
int lds_step(const mjhModel& m) { return lds_step(m, tune_for(m, false)); }
// The 64-lane fused step as three launches (forward with the mass-matrix
// factor; solver; integrator), each with its own LDS, or as one launch
// (MJH_FUSED64=one).
bool split64() {
  static const bool one = [] {
    const char* e = std::getenv("MJH_FUSED64");
    return e && std::strcmp(e, "one") == 0;
  }();
  return !one;
}

// Worlds of models with nv <= 32 run on 32-lane teams (two per wavefront)
// with the register solver; larger models, or every model when
// MJH_SOLVER_PATH=lds, on 64-lane teams with the LDS solver.
bool reg_path(const mjhModel& m) {
  static const bool force_lds = [] {
    const char* e = std::getenv("MJH_SOLVER_PATH");
    return e && std::strcmp(e, "lds") == 0;
  }();
  return m.nv <= kRegNvMax && !force_lds;
}
// per-world LDS of the 32-lane kernels, a multiple of 4 floats so every
// world's broadcast buffers are float4-aligned
int lds_solve_reg(const mjhModel& m, const Tune& t) {
  const int a = lds_reg_floats(t.rj, (m.nv + 3) & ~3), b = lds_linalg(m);
  return ((a > b ? a : b) + 3) & ~3;
}
int lds_solve_reg(const mjhModel& m) { return lds_solve_reg(m, tune_for(m)); }
// The register tail stages as many Jacobian rows as two worlds' LDS fit in
// 64 KB (up to every row of kRegRows).
Tune tune_tail_reg(const mjhModel& m, Tune t) {
  const int nvc = (m.nv + 3) & ~3;
  const int fit = (16384 / 2 - lds_reg_floats(0, nvc) - 3) / (nvc + 1);
  t.rj = fit < kRegRows ? fit : kRegRows;
  return t;
}
Tune tune_tail_reg(const mjhModel& m) { return tune_tail_reg(m, tune_for(m)); }
int lds_tail_reg(const mjhModel& m, const Tune& tail) {
  const int a = lds_reg_floats(tail.rj, (m.nv + 3) & ~3), b = lds_linalg(m);
  return ((a > b ? a : b) + 3) & ~3;
}
int lds_tail_reg(const mjhModel& m) { return lds_tail_reg(m, tune_tail_reg(m)); }

// A second stream per device and host thread for work that may run beside
// the caller's stream (default priority; ordered against it with events).
// Null when it cannot be created: that work then runs in order on the
// caller's stream.  Kept until the process ends.
struct SideStream {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
SideStream* side_stream() {
  thread_local std::map<int, SideStream> per_device;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  SideStream& ss = per_device[dev];
  if (!ss.s) {
    // Creation may happen while the caller captures a graph: relaxed mode
    // keeps it from invalidating that capture.
    hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
    (void)hipThreadExchangeStreamCaptureMode(&mode);
    const bool ok = hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking) == hipSuccess &&
                    hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming) == hipSuccess &&
                    hipEventCreateWithFlags(&ss.join, hipEventDisableTiming) == hipSuccess;
    (void)hipThreadExchangeStreamCaptureMode(&mode);
    if (!ok) {
      (void)hipGetLastError();
      if (ss.join) (void)hipEventDestroy(ss.join);
      if (ss.fork) (void)hipEventDestroy(ss.fork);
      if (ss.s) (void)hipStreamDestroy(ss.s);
      ss = SideStream{};
      return nullptr;
    }
  }
  return &ss;
}

#define TRY(x)                     \
  do {                             \
    int _rc = (x);                 \
    if (_rc != MJH_OK) return _rc; \
  } while (0)

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MJH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return MJH_OK;
}

int check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) return fail(MJH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return MJH_OK;
}

// LDS bytes one workgroup may take on the current device (160 KiB on
// gfx950), read once per device.
size_t lds_cap() {
  static std::mutex mu;
  static std::map<int, size_t> per_device;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return 65536;
  }
  std::lock_guard<std::mutex> lock(mu);
  auto it = per_device.find(dev);
  if (it != per_device.end()) return it->second;
  int v = 0;
  size_t cap = 65536;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0) cap = v;
  else (void)hipGetLastError();
  per_device[dev] = cap;
  return cap;
}
// A launch of `kernel` with `bytes` of dynamic LDS: above 64 KiB the kernel
// is allowed that much first.
int allow_lds(const void* kernel, size_t bytes, const char* name) {
  if (bytes <= 65536) return MJH_OK;
  if (bytes > lds_cap())
    return fail(MJH_ERR_CAPACITY, std::string(name) + ": needs " + std::to_string(bytes) + " bytes of LDS per workgroup, the device has " +
                                      std::to_string(lds_cap()));
  return check_hip(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes), name);
}

// Runs side(stream) beside main(s): side on the side stream, forked after
// the work already queued on s and joined before what s runs next; in order
// on s when there is no side stream or use_side is false.
template <class Side, class Main>
int beside(hipStream_t s, Side side, Main main, bool use_side = true) {
  SideStream* ss = use_side ? side_stream() : nullptr;
  if (!ss) {
    TRY(main(s));
    return side(s);
  }
  TRY(check_hip(hipEventRecord(ss->fork, s), "fork"));
  TRY(check_hip(hipStreamWaitEvent(ss->s, ss->fork, 0), "fork"));
  TRY(side(ss->s));
  TRY(main(s));
  TRY(check_hip(hipEventRecord(ss->join, ss->s), "join"));
  return check_hip(hipStreamWaitEvent(s, ss->join, 0), "join");
}

template <class S, int TEAM>
int launch_team(const mjhModel& m, const mjhData& d, hipStream_t s, int ldsw, const char* name,
                const Tune* tune = nullptr) {
  if (d.nworld <= 0) return MJH_OK;
  constexpr int wpb = 64 / TEAM;
  if (wpb > 1 && ldsw % 4 != 0) return fail(MJH_ERR_INVALID, std::string(name) + ": LDS stride not 16-byte aligned");
  const int grid = (d.nworld + wpb - 1) / wpb;
  const size_t bytes = (size_t)ldsw * wpb * sizeof(float);
  if (bytes > lds_cap())
    return fail(MJH_ERR_CAPACITY, std::string(name) + ": model too large for the dense LDS kernels (nv = " +
                                      std::to_string(m.nv) + ")");
  TRY(allow_lds((const void*)k_stage<TEAM, S>, bytes, name));
  hipLaunchKernelGGL((k_stage<TEAM, S>), dim3(grid), dim3(64), bytes, s, m, d, ldsw, tune ? *tune : tune_for(m));
  return check_launch(name);
}

// Per-world LDS stride of a 32-lane launch: at least kRegLdsMin floats and
// a multiple of 4, so every world's slot (and the float4 buffers stages put
// at its start) is 16-byte aligned.
int reg_stride(int ldsw) { return ((ldsw > kRegLdsMin ? ldsw : kRegLdsMin) + 3) & ~3; }
// the fused forward launch of 32-lane teams
int lds_forward_reg(const mjhModel& m) { return reg_stride(lds_forward(m)); }

// A stage on the model's team size: 32-lane teams (two worlds per
// wavefront) when the register path applies, else 64-lane teams.
template <class S>
int launch(const mjhModel& m, const mjhData& d, hipStream_t s, int ldsw, const char* name) {
  if (reg_path(m)) return launch_team<S, kRegTeam>(m, d, s, reg_stride(ldsw), name);
  return launch_team<S, kTeam>(m, d, s, ldsw, name);
}

int reset(const mjhData& d, hipStream_t s, int which) {
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, d, which);
  return check_launch("reset");
}



int do_collision(const mjhModel& m, const mjhData& d, hipStream_t s) {
  TRY(launch<StBroadphase>(m, d, s, lds_broadphase(m), "broadphase"));
  return launch<StNarrowphase>(m, d, s, 0, "narrowphase");
}

int solve_reg_desc(const mjhModel& m, const mjhData& d, hipStream_t s);

int do_solve(const mjhModel& m, const mjhData& d, hipStream_t s) {
  if (reg_path(m) && d.step_desc && d.nworld > 0) {
    TRY(solve_reg_desc(m, d, s));
  } else if (reg_path(m)) {
    const Tune tt = tune_tail_reg(m);
    TRY(beside(
        s,
        [&](hipStream_t q) {
          return launch_team<StBigWorlds<false>, kRegTeam>(m, d, q, lds_tail_reg(m), "solve_big_worlds", &tt);
        },
        [&](hipStream_t q) { return launch_team<StSolveTeam, kRegTeam>(m, d, q, lds_solve_reg(m), "solve"); }));
  } else {
    TRY((launch_team<StSolveTeam, kTeam>(m, d, s, lds_solve(m), "solve")));
  }
  return launch_team<StTail<false>, kTeam>(m, d, s, lds_tail(m, tune_for(m)), "solve_tail");
}

int do_fwd_position(const mjhModel& m, const mjhData& d, hipStream_t s) {
  TRY(launch<StKinematics>(m, d, s, lds_kin(m), "kinematics"));
  TRY(launch<StComPos>(m, d, s, lds_kin(m), "com_pos"));
  TRY(launch<StCrb>(m, d, s, lds_crb(m, reg_path(m)), "crb"));
  TRY(launch<StFactorM>(m, d, s, lds_linalg(m), "factor_m"));
  TRY(reset(d, s, 1));
  TRY(do_collision(m, d, s));
  if (!(m.opt.disableflags & mjhDSBL_CONSTRAINT)) {
    TRY(reset(d, s, 2));
    TRY(launch<StMakeConstraint>(m, d, s, lds_con(reg_path(m)), "make_constraint"));
  }
  return launch<StTransmission>(m, d, s, 0, "transmission");
}

int do_forward(const mjhModel& m, const mjhData& d, hipStream_t s) {
  TRY(do_fwd_position(m, d, s));
  TRY(launch<StFwdVelocity>(m, d, s, lds_vel(m), "fwd_velocity"));
  TRY(launch<StFwdActuation>(m, d, s, 0, "fwd_actuation"));
  TRY(launch<StFwdAcceleration>(m, d, s, lds_linalg(m), "fwd_acceleration"));
  if (d.njmax == 0) return launch<StQaccSmooth>(m, d, s, 0, "qacc_smooth");
  return do_solve(m, d, s);
}

// forward.py:268: implicitfast needs the dense mass matrix
int check_implicit_dense(const mjhModel& m) {
  if (m.opt.integrator == mjhINT_IMPLICITFAST && m.opt.is_sparse)
    return fail(MJH_ERR_UNSUPPORTED, "implicit integration requires a dense mass matrix");
  return MJH_OK;
}

int do_integrate(const mjhModel& m, const mjhData& d, hipStream_t s) {
  TRY(check_implicit_dense(m));
  if (m.opt.integrator == mjhINT_EULER) return launch<StEuler>(m, d, s, lds_linalg(m), "euler");
  if (m.opt.integrator == mjhINT_IMPLICITFAST) return launch<StImplicit>(m, d, s, lds_linalg(m), "implicit");
  return fail(MJH_ERR_UNSUPPORTED, "integrator " + std::to_string(m.opt.integrator) + " not implemented.");
}

int do_step_fused_desc(const mjhModel& m, const mjhData& d, hipStream_t s);

int do_step_fused(const mjhModel& m, const mjhData& d, hipStream_t s) {
  if (m.opt.integrator != mjhINT_EULER && m.opt.integrator != mjhINT_IMPLICITFAST)
    return fail(MJH_ERR_UNSUPPORTED, "integrator " + std::to_string(m.opt.integrator) + " not implemented.");
  TRY(check_implicit_dense(m));
  if (d.step_desc && d.nworld > 0) return do_step_fused_desc(m, d, s);
  TRY(reset(d, s, (m.opt.disableflags & mjhDSBL_CONSTRAINT) ? 1 : 3));
  if (reg_path(m)) {
    TRY((launch_team<StFusedForward, kRegTeam>(m, d, s, lds_forward_reg(m), "step_fused_forward")));
    const Tune tt = tune_tail_reg(m);
    TRY(beside(
        s,
        [&](hipStream_t q) {
          return launch_team<StBigWorlds<true>, kRegTeam>(m, d, q, lds_tail_reg(m), "step_big_worlds", &tt);
        },
        [&](hipStream_t q) {
          return launch_team<StFusedSolveIntegrate, kRegTeam>(m, d, q, lds_solve_reg(m), "step_fused_solve");
        }));
  } else {
    TRY((launch_team<StFusedStep, kTeam>(m, d, s, lds_step(m), "step_fused")));
  }
  return launch_team<StTail<true>, kTeam>(m, d, s, lds_tail(m, tune_for(m)), "step_tail");
}

// one workgroup per group
__global__ void k_reset_groups(const GroupDesc* __restrict__ g) {
  if (threadIdx.x != 0) return;
  const GroupDesc& G = g[blockIdx.x];
  reset_counters(G.d, (G.m.opt.disableflags & mjhDSBL_CONSTRAINT) ? 1 : 3);
}

// A device table of groups for launches of one team size.
struct GroupTable {
  GroupDesc* dev = nullptr;
  int ngroups = 0;
  int blocks = 0;
  int* queue = nullptr;  // kQueueSlots launches x kQueueLanes counters (work-queue plans), or null
};
constexpr int kQueueSlots = 8;
constexpr int kQueueInts = kQueueSlots * kQueueLanes * kQueueStride;
// workgroups of a work-queue launch: 8 per CU of a 256-CU device (a
// multiple of kQueueLanes); more than fit at once only take an empty
// ticket and exit (a 32-per-CU grid measured slower: more workgroups
// contending for the counters)
constexpr int kQueueGrid = 2048;
static_assert(kQueueGrid % kQueueLanes == 0, "queue grid covers every counter");
static_assert(kQueueSlots * kQueueLanes == 64, "k_zero_queue clears one counter per lane");
__global__ void k_zero_queue(int* q) {
  q[threadIdx.x * kQueueStride] = 0;
}

}  // namespace

namespace {

// Groups launched together: device tables of its groups (32-lane groups,
// seen by the 32-lane launches and by the 64-lane tail; 64-lane groups) and
// each launch's per-world LDS stride, the largest over its groups.
struct Bucket {
  GroupTable t32, t32_64, t64;
  int lds_fwd = 0, lds_solve = 0, lds_big = 0, lds_tail32 = 0, lds_step = 0, lds_tail64 = 0;
  int lds_fwd64 = 0, lds_solve64 = 0;
};

}  // namespace

// The plan of a heterogeneous batch: one bucket, or (MJH_MULTI_BUCKETS) one
// per team size and LDS occupancy class, each with its own launches.
struct mjhMulti {
  std::vector<Bucket> buckets;
};

namespace {

int build_table(const std::vector<GroupDesc>& host, int wpb, GroupTable* t) {
  t->ngroups = (int)host.size();
  if (host.empty()) return MJH_OK;
  std::vector<GroupDesc> h = host;
  long long blocks = 0;
  for (GroupDesc& g : h) {
    g.first_block = (int)blocks;
    blocks += (g.d.nworld + wpb - 1) / wpb;
  }
  if (blocks > 0x7fffffffLL) return fail(MJH_ERR_CAPACITY, "multi: too many worlds");
  t->blocks = (int)blocks;
  TRY(check_hip(hipMalloc(&t->dev, h.size() * sizeof(GroupDesc)), "multi: table"));
  return check_hip(hipMemcpy(t->dev, h.data(), h.size() * sizeof(GroupDesc), hipMemcpyHostToDevice), "multi: table");
}

void free_tables(mjhMulti* p) {
  for (Bucket& b : p->buckets)
    for (GroupTable* t : {&b.t32, &b.t32_64, &b.t64}) {
      if (t->dev) (void)hipFree(t->dev);
      if (t->queue) (void)hipFree(t->queue);
    }
}

// One launch of stage S over a group table: a workgroup per block, or, for
// a table with queue counters, the work-queue form taking counter `slot`.
template <int TEAM, class S, int TI>
int launch_groups(const GroupTable& t, int ldsw, hipStream_t s, const char* name, int slot = -1) {
  if (t.ngroups == 0 || t.blocks == 0) return MJH_OK;
  constexpr int wpb = 64 / TEAM;
  if (wpb > 1 && ldsw % 4 != 0) return fail(MJH_ERR_INVALID, std::string(name) + ": LDS stride not 16-byte aligned");
  const size_t bytes = (size_t)ldsw * wpb * sizeof(float);
  if (t.queue && slot >= 0 && slot < kQueueSlots) {
    TRY(allow_lds((const void*)k_stage_queue<TEAM, S, TI>, bytes, name));
    // a multiple of kQueueLanes, so every counter has workgroups drawing on it
    const int want = t.blocks < kQueueGrid ? t.blocks : kQueueGrid;
    const int grid = (want + kQueueLanes - 1) / kQueueLanes * kQueueLanes;
    hipLaunchKernelGGL((k_stage_queue<TEAM, S, TI>), dim3(grid), dim3(64), bytes, s, t.dev, t.ngroups, ldsw,
                       t.queue + slot * kQueueLanes * kQueueStride, t.blocks);
    return check_launch(name);
  }
  TRY(allow_lds((const void*)k_stage_groups<TEAM, S, TI>, bytes, name));
  hipLaunchKernelGGL((k_stage_groups<TEAM, S, TI>), dim3(t.blocks), dim3(64), bytes, s, t.dev, t.ngroups, ldsw);
  return check_launch(name);
}

int max_of(int a, int b) { return a > b ? a : b; }

// The fused step through the group kernels with d as the only group, its
// description written to d.step_desc first: the kernels then read the
// model from device memory (the kernel-argument copy of the model compiles
// to slower code for these kernels: part of it in scratch memory, more
// registers, one wavefront per SIMD fewer).
static_assert(sizeof(GroupDesc) <= MJH_STEP_DESC_WORDS * sizeof(int), "step_desc too small");
static_assert(sizeof(GroupDesc) <= 4000, "the description must fit the kernel arguments");
// writes the description and clears the step's counters (which as in
// reset_counters) in one launch
__global__ void k_write_desc(GroupDesc g, GroupDesc* out, int which) {
  if (threadIdx.x == 0) {
    *out = g;
    reset_counters(g.d, which);
  }
}
// The description of d as the only group of a launch.
GroupDesc single_desc(const mjhModel& m, const mjhData& d) {
  GroupDesc g{};
  g.m = m;
  g.d = d;
  g.first_block = 0;
  g.tune[0] = tune_for(m);
  g.tune[1] = reg_path(m) ? tune_tail_reg(m) : g.tune[0];
  return g;
}

// The 32-lane solve launches of the plain step (mjh_solve) through the
// group kernels, as the fused step does; counters are left alone.
int solve_reg_desc(const mjhModel& m, const mjhData& d, hipStream_t s) {
  const int l_solve = lds_solve_reg(m), l_big = lds_tail_reg(m);
  if ((size_t)max_of(2 * l_solve, 2 * l_big) * sizeof(float) > lds_cap())
    return fail(MJH_ERR_CAPACITY, "solve: model too large for the dense LDS kernels (nv = " + std::to_string(m.nv) + ")");
  GroupDesc* dev = reinterpret_cast<GroupDesc*>(d.step_desc);
  hipLaunchKernelGGL(k_write_desc, dim3(1), dim3(64), 0, s, single_desc(m, d), dev, 0);
  TRY(check_launch("solve: description"));
  GroupTable t32;
  t32.dev = dev;
  t32.ngroups = 1;
  t32.blocks = (d.nworld + 1) / 2;
  return beside(
      s, [&](hipStream_t q) { return launch_groups<kRegTeam, StBigWorlds<false>, 1>(t32, l_big, q, "solve_big_worlds"); },
      [&](hipStream_t q) { return launch_groups<kRegTeam, StSolveTeam, 0>(t32, l_solve, q, "solve"); });
}

int do_step_fused_desc(const mjhModel& m, const mjhData& d, hipStream_t s) {
  GroupDesc g{};
  g.m = m;
  g.d = d;
  g.first_block = 0;
  g.tune[0] = tune_for(m);
  g.tune[1] = reg_path(m) ? tune_tail_reg(m) : g.tune[0];
  GroupDesc* dev = reinterpret_cast<GroupDesc*>(d.step_desc);
  const bool reg = reg_path(m);
  const int l_fwd = lds_forward_reg(m), l_solve = lds_solve_reg(m), l_big = lds_tail_reg(m), l_step = lds_step(m),
            l_tail = lds_tail(m, g.tune[0]), l_fwd64 = max_of(lds_forward(m, false), lds_linalg(m)),
            l_solve64 = lds_solve(m, g.tune[0]);
  const size_t need = reg ? (size_t)max_of(max_of(2 * l_fwd, 2 * l_solve), max_of(2 * l_big, l_tail))
                          : (size_t)max_of(split64() ? max_of(l_fwd64, l_solve64) : l_step, l_tail);
  if (need * sizeof(float) > lds_cap())
    return fail(MJH_ERR_CAPACITY,
                "step_fused: model too large for the dense LDS kernels (nv = " + std::to_string(m.nv) + ")");
  hipLaunchKernelGGL(k_write_desc, dim3(1), dim3(64), 0, s, g, dev, (m.opt.disableflags & mjhDSBL_CONSTRAINT) ? 1 : 3);
  TRY(check_launch("step_fused: description"));
  GroupTable t32, t64;
  t32.dev = t64.dev = dev;
  t32.ngroups = t64.ngroups = 1;
  t32.blocks = (d.nworld + 1) / 2;
  t64.blocks = d.nworld;
  if (reg) {
    TRY((launch_groups<kRegTeam, StFusedForward, 0>(t32, l_fwd, s, "step_fused_forward")));
    TRY(beside(
        s, [&](hipStream_t q) { return launch_groups<kRegTeam, StBigWorlds<true>, 1>(t32, l_big, q, "step_big_worlds"); },
        [&](hipStream_t q) {
          return launch_groups<kRegTeam, StFusedSolveIntegrate, 0>(t32, l_solve, q, "step_fused_solve");
        }));
  } else if (split64()) {
    TRY((launch_groups<kTeam, StForwardFactor, 0>(t64, l_fwd64, s, "step_fused_forward")));
    TRY((launch_groups<kTeam, StSolveStep, 0>(t64, l_solve64, s, "step_fused_solve")));
    TRY((launch_groups<kTeam, StIntegrateSolved, 0>(t64, lds_linalg(m), s, "step_fused_integrate")));
  } else {
    TRY((launch_groups<kTeam, StFusedStep, 0>(t64, l_step, s, "step_fused")));
  }
  return launch_groups<kTeam, StTail<true>, 0>(t64, l_tail, s, "step_tail");
}

// How one model's worlds are launched: on 32-lane teams with the register
// solver (reg) or on 64-lane teams with the LDS solver, and the solver's
// tuning for the solve / tail launches ([0]) and the big-world launch ([1]);
// the library's choice, or the caller's where o sets it.
struct Choice {
  bool reg = false;
  Tune tune[2] = {};
};
int choose(int i, const mjhModel& m, const mjhGroupTune* o, Choice* c) {
  const std::string g = "multi: group " + std::to_string(i) + ": ";
  c->reg = reg_path(m);
  if (o) {
    if (o->team != 0 && o->team != kRegTeam && o->team != kTeam) return fail(MJH_ERR_INVALID, g + "team must be 0, 32 or 64");
    if (o->team == kRegTeam && m.nv > kRegNvMax) return fail(MJH_ERR_INVALID, g + "32-lane teams need nv <= 32");
    if (o->team) c->reg = o->team == kRegTeam;
  }
  Tune t = tune_for(m, c->reg);
  if (o && o->rmax > 0) t.rmax = o->rmax;
  if (o && o->rj >= 0) t.rj = o->rj;
  if (t.rj > t.rmax) t.rj = t.rmax;
  c->tune[0] = t;
  c->tune[1] = c->reg ? tune_tail_reg(m, t) : t;
  return MJH_OK;
}

// per-world LDS strides of a group's launches
struct GroupLds {
  int fwd = 0, solve = 0, big = 0, tail32 = 0, step = 0, tail64 = 0, fwd64 = 0, solve64 = 0;
};
GroupLds group_lds(const mjhModel& m, const Choice& c) {
  GroupLds l;
  if (c.reg) {
    l.fwd = reg_stride(lds_forward(m, true));
    l.solve = lds_solve_reg(m, c.tune[0]);
    l.big = lds_tail_reg(m, c.tune[1]);
    l.tail32 = lds_tail(m, c.tune[0]);
  } else {
    l.step = lds_step(m, c.tune[0]);
    l.tail64 = lds_tail(m, c.tune[0]);
    l.fwd64 = max_of(lds_forward(m, false), lds_linalg(m));
    l.solve64 = lds_solve(m, c.tune[0]);
  }
  return l;
}
void fold(Bucket* b, const GroupLds& l) {
  b->lds_fwd = max_of(b->lds_fwd, l.fwd);
  b->lds_solve = max_of(b->lds_solve, l.solve);
  b->lds_big = max_of(b->lds_big, l.big);
  b->lds_tail32 = max_of(b->lds_tail32, l.tail32);
  b->lds_step = max_of(b->lds_step, l.step);
  b->lds_tail64 = max_of(b->lds_tail64, l.tail64);
  b->lds_fwd64 = max_of(b->lds_fwd64, l.fwd64);
  b->lds_solve64 = max_of(b->lds_solve64, l.solve64);
}
// Workgroups a CU holds by the LDS of the group's main launches (the
// forward and solve launches), at most 16.
int lds_class(const Choice& c, const GroupLds& l) {
  const size_t words = c.reg ? 2 * (size_t)max_of(l.fwd, max_of(l.solve, l.big))
                             : (size_t)(split64() ? max_of(l.fwd64, l.solve64) : l.step);
  const size_t bytes = words * sizeof(float);
  const size_t k = bytes ? lds_cap() / bytes : 16;
  return k < 16 ? (int)k : 16;
}

// Checks group i (as mjh_check and the fused step do) and fills its table
// entry (model, data, tuning; first_block is the caller's), its launch
// choice and its LDS strides.
int plan_group(int i, const mjhModel& m, const mjhData& d, const mjhGroupTune* o, GroupDesc* gd, Choice* c,
               GroupLds* l) {
  if (mjh_check(&m, &d) != MJH_OK) return fail(MJH_ERR_INVALID, "multi: group " + std::to_string(i) + ": " + g_last_error);
  if (m.opt.integrator != mjhINT_EULER && m.opt.integrator != mjhINT_IMPLICITFAST)
    return fail(MJH_ERR_UNSUPPORTED, "multi: group " + std::to_string(i) + ": integrator " +
                                         std::to_string(m.opt.integrator) + " not implemented.");
  TRY(check_implicit_dense(m));
  TRY(choose(i, m, o, c));
  gd->m = m;
  gd->d = d;
  gd->first_block = 0;
  gd->tune[0] = c->tune[0];
  gd->tune[1] = c->tune[1];
  *l = group_lds(m, *c);
  return MJH_OK;
}

// each launch's LDS: two worlds per workgroup on 32-lane teams, one on 64
int plan_capacity(const Bucket& plan) {
  const int need = max_of(max_of(2 * plan.lds_fwd, 2 * plan.lds_solve), max_of(2 * plan.lds_big, plan.lds_tail32));
  const int need64 = split64() ? max_of(plan.lds_fwd64, plan.lds_solve64) : plan.lds_step;
  if ((size_t)max_of(need, max_of(need64, plan.lds_tail64)) * sizeof(float) > lds_cap())
    return fail(MJH_ERR_CAPACITY, "multi: a group's model is too large for the dense LDS kernels");
  return MJH_OK;
}

// The launches of one fused step of every group of a plan; the big-world
// launch beside the solve launch when use_side, else in order on s.
int multi_launches(const Bucket& p, hipStream_t s, bool use_side) {
  for (const GroupTable* t : {&p.t32, &p.t64}) {
    if (t->ngroups == 0) continue;
    hipLaunchKernelGGL(k_reset_groups, dim3(t->ngroups), dim3(64), 0, s, t->dev);
    TRY(check_launch("multi: reset"));
  }
  for (const GroupTable* t : {&p.t32, &p.t32_64, &p.t64})
    if (t->ngroups && t->queue) {
      hipLaunchKernelGGL(k_zero_queue, dim3(1), dim3(64), 0, s, t->queue);
      TRY(check_launch("multi: queue counters"));
    }
  if (p.t32.ngroups) {
    TRY((launch_groups<kRegTeam, StFusedForward, 0>(p.t32, p.lds_fwd, s, "multi: forward", 0)));
    TRY(beside(
        s,
        [&](hipStream_t q) {
          return launch_groups<kRegTeam, StBigWorlds<true>, 1>(p.t32, p.lds_big, q, "multi: big worlds", 1);
        },
        [&](hipStream_t q) {
          return launch_groups<kRegTeam, StFusedSolveIntegrate, 0>(p.t32, p.lds_solve, q, "multi: solve", 2);
        },
        use_side));
    TRY((launch_groups<kTeam, StTail<true>, 0>(p.t32_64, p.lds_tail32, s, "multi: tail", 0)));
  }
  if (p.t64.ngroups) {
    if (split64()) {
      TRY((launch_groups<kTeam, StForwardFactor, 0>(p.t64, p.lds_fwd64, s, "multi: forward", 0)));
      TRY((launch_groups<kTeam, StSolveStep, 0>(p.t64, p.lds_solve64, s, "multi: solve", 1)));
      TRY((launch_groups<kTeam, StIntegrateSolved, 0>(p.t64, p.lds_tail64, s, "multi: integrate", 2)));
    } else {
      TRY((launch_groups<kTeam, StFusedStep, 0>(p.t64, p.lds_step, s, "multi: step", 3)));
    }
    TRY((launch_groups<kTeam, StTail<true>, 0>(p.t64, p.lds_tail64, s, "multi: tail", 4)));
  }
  return MJH_OK;
}

// The group table of mjh_multi_step_static: a fixed device array (each GPU
// holds its own copy at its own address), so the call allocates nothing.
// The three tables of a plan sit back to back: 2 n32 + n64 <= 2 ngroups
// entries.
static __device__ __align__(256) GroupDesc scratch[2 * MJH_MULTI_STATIC_MAX_GROUPS];
constexpr int kMaxDevices = 64;
GroupDesc* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// The device address of this GPU's table, looked up once per device with
// that device current; host code uses it, never the array's name.
int static_table(GroupDesc** out) {
  int dev = 0;
  TRY(check_hip(hipGetDevice(&dev), "multi_static: device"));
  if (dev < 0 || dev >= kMaxDevices) return fail(MJH_ERR_CAPACITY, "multi_static: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    TRY(check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "multi_static: table address"));
    g_scratch_ptr[dev] = static_cast<GroupDesc*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return MJH_OK;
}

}  // namespace

extern "C" {

int mjh_abi_version(void) { return MJH_ABI_VERSION; }

const char* mjh_last_error(void) { return g_last_error.c_str(); }

int mjh_check(const mjhModel* m, const mjhData* d) {
  if (!m || !d) return fail(MJH_ERR_INVALID, "null model or data");
  if (m->opt.cone != mjhCONE_PYRAMIDAL) return fail(MJH_ERR_UNSUPPORTED, "only the pyramidal cone is implemented");
  if (m->opt.solver != mjhSOL_CG && m->opt.solver != mjhSOL_NEWTON)
    return fail(MJH_ERR_UNSUPPORTED, "Solver: " + std::to_string(m->opt.solver) + " is unsupported.");
  if (m->nlevel < 1 || m->nbody < 1) return fail(MJH_ERR_INVALID, "model has no bodies");
  if (d->nworld < 0 || d->nconmax < 0 || d->njmax < 0) return fail(MJH_ERR_INVALID, "negative sizes");
  if (m->opt.is_sparse) {
    if (d->qM_rows != 1 || d->qM_cols != m->nM || d->qMd_n != m->nv)
      return fail(MJH_ERR_INVALID, "sparse model: qM must be [nworld, 1, nM] and qMd [nworld, nv, nv]");
  } else if (d->qM_rows != m->nv || d->qM_cols != m->nv) {
    return fail(MJH_ERR_INVALID, "qM must be dense [nworld, nv, nv]");
  }
  if (m->nv > 0 && d->nworld > 0 && (!d->qMd || !d->qLDd))
    return fail(MJH_ERR_INVALID, "qMd / qLDd must be set (qM / qLD for a dense model)");
  return MJH_OK;
}

#define STAGE_ENTRY(fname, S, ldsexpr)                                \
  int fname(const mjhModel* m, const mjhData* d, mjhStream stream) {  \
    return launch<S>(*m, *d, (hipStream_t)stream, (ldsexpr), #fname); \
  }

STAGE_ENTRY(mjh_kinematics, StKinematics, lds_kin(*m))
STAGE_ENTRY(mjh_com_pos, StComPos, lds_kin(*m))
STAGE_ENTRY(mjh_crb, StCrb, lds_crb(*m, reg_path(*m)))
STAGE_ENTRY(mjh_factor_m, StFactorM, lds_linalg(*m))
STAGE_ENTRY(mjh_rne, StRne, lds_vel(*m))
STAGE_ENTRY(mjh_transmission, StTransmission, 0)
STAGE_ENTRY(mjh_com_vel, StComVel, lds_vel(*m))
STAGE_ENTRY(mjh_passive, StPassive, 0)
STAGE_ENTRY(mjh_fwd_velocity, StFwdVelocity, lds_vel(*m))
STAGE_ENTRY(mjh_fwd_actuation, StFwdActuation, 0)
STAGE_ENTRY(mjh_fwd_acceleration, StFwdAcceleration, lds_linalg(*m))

int mjh_solve(const mjhModel* m, const mjhData* d, mjhStream stream) {
  hipStream_t s = (hipStream_t)stream;
  if (d->nworld > 0 && d->njmax > 0) {
    // each world's rows from efc_worldid (the pool may be laid out by the caller)
    const int wg = (d->nworld + 255) / 256, rg = (d->njmax + 255) / 256;
    hipLaunchKernelGGL(k_blocks_zero, dim3(wg), dim3(256), 0, s, *d);
    hipLaunchKernelGGL(k_blocks_count, dim3(rg < 4096 ? rg : 4096), dim3(256), 0, s, *d);
    hipLaunchKernelGGL(k_blocks_place, dim3(1), dim3(1024), 0, s, *d);
    hipLaunchKernelGGL(k_blocks_slots, dim3(rg < 4096 ? rg : 4096), dim3(256), 0, s, *d);
    hipLaunchKernelGGL(k_blocks_gather, dim3(d->nworld), dim3(64), 0, s, *d, m->nv);
    TRY(check_launch("solve: row blocks"));
  }
  TRY(do_solve(*m, *d, s));
  if (d->nworld > 0 && d->njmax > 0) {
    hipLaunchKernelGGL(k_blocks_scatter, dim3(d->nworld), dim3(64), 0, s, *d);
    TRY(check_launch("solve: scatter"));
  }
  return MJH_OK;
}
int mjh_nxn_broadphase(const mjhModel* m, const mjhData* d, mjhStream stream) {
  TRY(reset(*d, (hipStream_t)stream, 1));
  return launch<StNxn>(*m, *d, (hipStream_t)stream, kNxnWords * m->ngeom, "nxn_broadphase");
}
int mjh_sap_broadphase(const mjhModel* m, const mjhData* d, mjhStream stream) {
  TRY(reset(*d, (hipStream_t)stream, 1));
  return launch<StSap>(*m, *d, (hipStream_t)stream, lds_sap(*m), "sap_broadphase");
}
int mjh_primitive_narrowphase(const mjhModel* m, const mjhData* d, mjhStream stream) {
  return launch<StNarrowphase>(*m, *d, (hipStream_t)stream, 0, "primitive_narrowphase");
}
int mjh_collision(const mjhModel* m, const mjhData* d, mjhStream stream) {
  TRY(reset(*d, (hipStream_t)stream, 1));
  return do_collision(*m, *d, (hipStream_t)stream);
}
int mjh_make_constraint(const mjhModel* m, const mjhData* d, mjhStream stream) {
  if (m->opt.disableflags & mjhDSBL_CONSTRAINT) return MJH_OK;
  TRY(reset(*d, (hipStream_t)stream, 2));
  return launch<StMakeConstraint>(*m, *d, (hipStream_t)stream, lds_con(reg_path(*m)), "make_constraint");
}
int mjh_fwd_position(const mjhModel* m, const mjhData* d, mjhStream stream) {
  return do_fwd_position(*m, *d, (hipStream_t)stream);
}
int mjh_euler(const mjhModel* m, const mjhData* d, mjhStream stream) {
  return launch<StEuler>(*m, *d, (hipStream_t)stream, lds_linalg(*m), "euler");
}
int mjh_implicit(const mjhModel* m, const mjhData* d, mjhStream stream) {
  if (m->opt.is_sparse) return fail(MJH_ERR_UNSUPPORTED, "implicit integration requires a dense mass matrix");
  return launch<StImplicit>(*m, *d, (hipStream_t)stream, lds_linalg(*m), "implicit");
}
int mjh_forward(const mjhModel* m, const mjhData* d, mjhStream stream) {
  return do_forward(*m, *d, (hipStream_t)stream);
}
int mjh_step(const mjhModel* m, const mjhData* d, mjhStream stream) {
  if (m->opt.integrator != mjhINT_EULER && m->opt.integrator != mjhINT_IMPLICITFAST)
    return fail(MJH_ERR_UNSUPPORTED, "integrator " + std::to_string(m->opt.integrator) + " not implemented.");
  TRY(check_implicit_dense(*m));
  TRY(do_forward(*m, *d, (hipStream_t)stream));
  return do_integrate(*m, *d, (hipStream_t)stream);
}
int mjh_step_fused(const mjhModel* m, const mjhData* d, mjhStream stream) {
  return do_step_fused(*m, *d, (hipStream_t)stream);
}
int mjh_rollout(const mjhModel* m, const mjhData* d, int nstep, int fused, mjhStream stream) {
  for (int i = 0; i < nstep; i++) TRY(fused ? mjh_step_fused(m, d, stream) : mjh_step(m, d, stream));
  return MJH_OK;
}
int mjh_solve_m(const mjhModel* m, const mjhData* d, float* x, const float* y, mjhStream stream) {
  if (d->nworld <= 0) return MJH_OK;
  const int ldsw = lds_linalg(*m);
  if ((size_t)ldsw * sizeof(float) > lds_cap()) return fail(MJH_ERR_CAPACITY, "solve_m: nv too large");
  TRY(allow_lds((const void*)k_solve_m, (size_t)ldsw * sizeof(float), "solve_m"));
  hipLaunchKernelGGL(k_solve_m, dim3(d->nworld), dim3(64), ldsw * sizeof(float), (hipStream_t)stream, *m, *d, x, y,
                     ldsw);
  return check_launch("solve_m");
}
int mjh_mul_m(const mjhModel* m, const mjhData* d, float* res, const float* vec, const uint8_t* skip,
              mjhStream stream) {
  if (d->nworld <= 0) return MJH_OK;
  hipLaunchKernelGGL(k_mul_m, dim3(d->nworld), dim3(64), 0, (hipStream_t)stream, *m, *d, res, vec, skip);
  return check_launch("mul_m");
}
int mjh_xfrc_accumulate(const mjhModel* m, const mjhData* d, float* qfrc, mjhStream stream) {
  if (d->nworld <= 0) return MJH_OK;
  hipLaunchKernelGGL(k_xfrc, dim3(d->nworld), dim3(64), 0, (hipStream_t)stream, *m, *d, qfrc);
  return check_launch("xfrc_accumulate");
}

int mjh_multi_create(int ngroups, const mjhModel* models, const mjhData* datas, mjhMulti** out) {
  return mjh_multi_create_tuned(ngroups, models, datas, nullptr, 0, out);
}
int mjh_multi_create_tuned(int ngroups, const mjhModel* models, const mjhData* datas, const mjhGroupTune* tunes,
                           int flags, mjhMulti** out) {
  if (!out) return fail(MJH_ERR_INVALID, "multi: null output");
  *out = nullptr;
  if (ngroups <= 0 || !models || !datas) return fail(MJH_ERR_INVALID, "multi: no groups");
  if (flags & ~(MJH_MULTI_BUCKETS | MJH_MULTI_QUEUE)) return fail(MJH_ERR_INVALID, "multi: unknown flags");
  std::vector<Bucket> buckets;
  std::vector<std::vector<GroupDesc>> g32, g64;
  std::map<std::pair<int, int>, int> index;
  for (int i = 0; i < ngroups; i++) {
    GroupDesc gd{};
    Choice c;
    GroupLds l;
    TRY(plan_group(i, models[i], datas[i], tunes ? tunes + i : nullptr, &gd, &c, &l));
    const std::pair<int, int> key = (flags & MJH_MULTI_BUCKETS) ? std::make_pair(c.reg ? 1 : 0, lds_class(c, l))
                                                                : std::make_pair(0, 0);
    auto it = index.find(key);
    if (it == index.end()) {
      it = index.emplace(key, (int)buckets.size()).first;
      buckets.emplace_back();
      g32.emplace_back();
      g64.emplace_back();
    }
    fold(&buckets[it->second], l);
    (c.reg ? g32 : g64)[it->second].push_back(gd);
  }
  for (const Bucket& b : buckets) TRY(plan_capacity(b));
  // largest models first: the hardware hands out workgroups in index
  // order, so the longest worlds start first and the short ones fill the
  // end of the launch (results do not depend on the order)
  const auto larger = [](const GroupDesc& a, const GroupDesc& b) { return a.m.nv > b.m.nv; };
  for (auto& v : g32) std::stable_sort(v.begin(), v.end(), larger);
  for (auto& v : g64) std::stable_sort(v.begin(), v.end(), larger);
  mjhMulti* p = new mjhMulti{buckets};
  int rc = MJH_OK;
  for (size_t k = 0; k < p->buckets.size() && rc == MJH_OK; k++) {
    Bucket& b = p->buckets[k];
    rc = build_table(g32[k], 2, &b.t32);
    if (rc == MJH_OK) rc = build_table(g32[k], 1, &b.t32_64);
    if (rc == MJH_OK) rc = build_table(g64[k], 1, &b.t64);
    if (flags & MJH_MULTI_QUEUE)
      for (GroupTable* t : {&b.t32, &b.t32_64, &b.t64})
        if (rc == MJH_OK && t->ngroups)
          rc = check_hip(hipMalloc(&t->queue, kQueueInts * sizeof(int)), "multi: queue counters");
  }
  if (rc != MJH_OK) {
    free_tables(p);
    delete p;
    return rc;
  }
  *out = p;
  return MJH_OK;
}
int mjh_multi_step(const mjhMulti* p, mjhStream stream) {
  if (!p) return fail(MJH_ERR_INVALID, "multi: null plan");
  for (const Bucket& b : p->buckets) TRY(multi_launches(b, (hipStream_t)stream, true));
  return MJH_OK;
}
int mjh_multi_step_static(int ngroups, const mjhModel* models, const mjhData* datas, mjhStream stream) {
  if (ngroups <= 0 || !models || !datas) return fail(MJH_ERR_INVALID, "multi_static: no groups");
  if (ngroups > MJH_MULTI_STATIC_MAX_GROUPS)
    return fail(MJH_ERR_CAPACITY, "multi_static: " + std::to_string(ngroups) + " groups; at most " +
                                      std::to_string(MJH_MULTI_STATIC_MAX_GROUPS) + " fit the library's table");
  Bucket plan;
  int n32 = 0, n64 = 0;
  for (int i = 0; i < ngroups; i++) {
    GroupDesc gd{};
    Choice c;
    GroupLds l;
    TRY(plan_group(i, models[i], datas[i], nullptr, &gd, &c, &l));
    fold(&plan, l);
    (c.reg ? n32 : n64)++;
  }
  TRY(plan_capacity(plan));
  GroupDesc* table = nullptr;
  TRY(static_table(&table));
  hipStream_t s = (hipStream_t)stream;
  // the three tables of a plan, back to back: 32-lane launches (two worlds
  // per workgroup), the 64-lane tail over the same groups, 64-lane groups
  plan.t32.dev = table;
  plan.t32_64.dev = table + n32;
  plan.t64.dev = table + 2 * n32;
  plan.t32.ngroups = plan.t32_64.ngroups = n32;
  plan.t64.ngroups = n64;
  long long b32 = 0, b32_64 = 0, b64 = 0;
  int i32 = 0, i64 = 0;
  for (int i = 0; i < ngroups; i++) {
    GroupDesc gd{};
    Choice c;
    GroupLds l;
    TRY(plan_group(i, models[i], datas[i], nullptr, &gd, &c, &l));
    const int nw = datas[i].nworld;
    if (c.reg) {
      gd.first_block = (int)b32;
      hipLaunchKernelGGL(k_write_desc, dim3(1), dim3(64), 0, s, gd, plan.t32.dev + i32, 0);
      gd.first_block = (int)b32_64;
      hipLaunchKernelGGL(k_write_desc, dim3(1), dim3(64), 0, s, gd, plan.t32_64.dev + i32, 0);
      b32 += (nw + 1) / 2;
      b32_64 += nw;
      i32++;
    } else {
      gd.first_block = (int)b64;
      hipLaunchKernelGGL(k_write_desc, dim3(1), dim3(64), 0, s, gd, plan.t64.dev + i64, 0);
      b64 += nw;
      i64++;
    }
    TRY(check_launch("multi_static: table"));
    if (b32_64 > 0x7fffffffLL || b64 > 0x7fffffffLL) return fail(MJH_ERR_CAPACITY, "multi_static: too many worlds");
  }
  plan.t32.blocks = (int)b32;
  plan.t32_64.blocks = (int)b32_64;
  plan.t64.blocks = (int)b64;
  return multi_launches(plan, s, false);
}
int mjh_multi_destroy(mjhMulti* p) {
  if (!p) return MJH_OK;
  free_tables(p);
  delete p;
  return MJH_OK;
}

int mjh_transition_fd(const mjhModel* m, const mjhData* d, const mjhData* fd, float eps, int centered, float* A,
                      float* B, mjhStream stream) {
  TRY(mjh_check(m, d));
  TRY(mjh_check(m, fd));
  if (!(eps > 0.0f) || !std::isfinite(eps)) return fail(MJH_ERR_INVALID, "transition_fd: eps must be positive");
  const int nx = 2 * m->nv + m->na, nc = nx + m->nu;
  const long long K = 1 + (centered ? 2LL : 1LL) * nc;
  if ((long long)fd->nworld != K * d->nworld)
    return fail(MJH_ERR_INVALID, "transition_fd: fd must hold " + std::to_string(K) + " worlds per world of d");
  if ((long long)d->nworld * nc > 0x7fffffffLL) return fail(MJH_ERR_CAPACITY, "transition_fd: too many columns");
  if (!A || (m->nu > 0 && !B)) return fail(MJH_ERR_INVALID, "transition_fd: A and B must be set");
  if (d->nworld <= 0) return MJH_OK;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fd_perturb, dim3(fd->nworld), dim3(64), 0, s, *m, *d, *fd, eps, centered ? 1 : 0);
  TRY(check_launch("transition_fd: perturb"));
  TRY(do_step_fused(*m, *fd, s));
  hipLaunchKernelGGL(k_fd_difference, dim3(d->nworld * nc), dim3(64), 0, s, *m, *fd, d->nworld, eps,
                     centered ? 1 : 0, A, B);
  return check_launch("transition_fd: difference");
}
int mjh_adjoint_step(const mjhModel* m, const mjhData* d, const float* A, const float* B, const float* g_qpos,
                     const float* g_qvel, const float* g_act, const float* lam_next, float* lam, float* grad_ctrl,
                     mjhStream stream) {
  if (!m || !d) return fail(MJH_ERR_INVALID, "null model or data");
  if (!lam || (lam_next && (!A || (m->nu > 0 && grad_ctrl && !B))))
    return fail(MJH_ERR_INVALID, "adjoint_step: lam (and A, B with lam_next) must be set");
  if (d->nworld <= 0) return MJH_OK;
  hipLaunchKernelGGL(k_adjoint_step, dim3(d->nworld), dim3(64), 0, (hipStream_t)stream, *m, *d, A, B, g_qpos, g_qvel,
                     g_act, lam_next, lam, grad_ctrl);
  return check_launch("adjoint_step");
}
}  // extern "C"

namespace {

// The slot layout: every field the adjoint reads, 256-byte aligned, pools
// at full capacity (their indices stay valid).  Fills the segments to copy
// (segs) and the Data whose fields point into the slot (out); returns the
// slot's bytes.
size_t tape_plan(const mjhModel& m, const mjhData& d, char* slot, mjhData* out, TapeSegs* segs) {
  mjhData t = d;
  size_t off = 0;
  int k = 0;
  auto seg = [&](auto& field, const void* src, size_t bytes, const int* count, unsigned row) {
    off = (off + 255) & ~(size_t)255;
    if (segs && k < kTapeSegsMax) segs->s[k] = TapeSeg{(const char*)src, off, bytes, count, row};
    k++;
    field = reinterpret_cast<std::remove_reference_t<decltype(field)>>(slot + off);
    off += bytes;
  };
  const size_t W = d.nworld, nv = m.nv, nb = m.nbody, nu = m.nu, ng = m.ngeom, F = 4;
  const size_t J = d.njmax, C = d.nconmax;
  seg(t.qvel, d.qvel, W * nv * F, nullptr, 0);
  seg(t.cdof, d.cdof, W * nv * 6 * F, nullptr, 0);
  seg(t.cdof_dot, d.cdof_dot, W * nv * 6 * F, nullptr, 0);
  seg(t.cinert, d.cinert, W * nb * 10 * F, nullptr, 0);
  seg(t.subtree_com, d.subtree_com, W * nb * 3 * F, nullptr, 0);
  seg(t.qMd, d.qMd, W * nv * nv * F, nullptr, 0);
  seg(t.qacc, d.qacc, W * nv * F, nullptr, 0);
  seg(t.qacc_integration, d.qacc_integration, W * nv * F, nullptr, 0);
  seg(t.ctrl, d.ctrl, W * nu * F, nullptr, 0);
  seg(t.actuator_moment, d.actuator_moment, W * nu * nv * F, nullptr, 0);
  seg(t.actuator_length, d.actuator_length, W * nu * F, nullptr, 0);
  seg(t.actuator_velocity, d.actuator_velocity, W * nu * F, nullptr, 0);
  seg(t.geom_xpos, d.geom_xpos, W * ng * 3 * F, nullptr, 0);
  seg(t.geom_xmat, d.geom_xmat, W * ng * 9 * F, nullptr, 0);
  seg(t.world_efc, d.world_efc, W * 2 * F, nullptr, 0);
  seg(t.world_con, d.world_con, W * 2 * F, nullptr, 0);
  seg(t.efc_J, d.efc_J, J * nv * F, d.nefc, (unsigned)(nv * F));
  seg(t.efc_D, d.efc_D, J * F, d.nefc, F);
  seg(t.efc_force, d.efc_force, J * F, d.nefc, F);
  seg(t.efc_pos, d.efc_pos, J * F, d.nefc, F);
  seg(t.efc_margin, d.efc_margin, J * F, d.nefc, F);
  seg(t.contact_geom, d.contact_geom, C * 2 * F, d.ncon, 2 * F);
  seg(t.contact_dim, d.contact_dim, C * F, d.ncon, F);
  seg(t.contact_pos, d.contact_pos, C * 3 * F, d.ncon, 3 * F);
  seg(t.contact_frame, d.contact_frame, C * 9 * F, d.ncon, 9 * F);
  seg(t.contact_friction, d.contact_friction, C * 5 * F, d.ncon, 5 * F);
  seg(t.contact_dist, d.contact_dist, C * F, d.ncon, F);
  seg(t.contact_includemargin, d.contact_includemargin, C * F, d.ncon, F);
  seg(t.contact_solref, d.contact_solref, C * 2 * F, d.ncon, 2 * F);
  seg(t.contact_solimp, d.contact_solimp, C * 5 * F, d.ncon, 5 * F);
  if (segs) segs->n = k < kTapeSegsMax ? k : kTapeSegsMax;
  if (out) *out = t;
  return (off + 255) & ~(size_t)255;
}

int check_adjoint(const mjhModel* m, const mjhData* d) {
  if (!m || !d) return fail(MJH_ERR_INVALID, "null model or data");
  if (m->na > 0) return fail(MJH_ERR_UNSUPPORTED, "adjoint_vjp: activation states (na > 0) are not differentiated");
  if (m->opt.integrator != mjhINT_EULER && m->opt.integrator != mjhINT_IMPLICITFAST)
    return fail(MJH_ERR_UNSUPPORTED, "adjoint_vjp: integrator must be Euler or implicitfast");
  if (m->actuator_affine_bias_gain && !(m->opt.disableflags & mjhDSBL_ACTUATION))
    return fail(MJH_ERR_UNSUPPORTED, "adjoint_vjp: affine actuator gains / biases are not differentiated");
  return MJH_OK;
}

int launch_vjp(const mjhModel& m, const mjhData& d, const float* qpos, const float* qvel, const float* lam_next,
               const float* gadd, float* lam, float* grad_ctrl, hipStream_t s) {
  if (d.nworld <= 0) return MJH_OK;
  const int fixed = lds_adjoint_floats(m.nbody, m.nv, 0);
  const int cap = (int)(lds_cap() / sizeof(float));
  // rows per world the LDS holds: the pool's per-world share, at most what fits
  int rcap = d.njmax / d.nworld + 1;
  if (rcap > (cap - fixed) / 8) rcap = (cap - fixed) / 8;
  if (rcap < 4) return fail(MJH_ERR_CAPACITY, "adjoint_vjp: nv too large for the LDS of one workgroup");
  const size_t bytes = (size_t)lds_adjoint_floats(m.nbody, m.nv, rcap) * sizeof(float);
  TRY(allow_lds((const void*)k_adjoint_vjp, bytes, "adjoint_vjp"));
  hipLaunchKernelGGL(k_adjoint_vjp, dim3(d.nworld), dim3(64), bytes, s, m, d, qpos, qvel, lam_next, gadd, lam,
                     grad_ctrl, rcap);
  return check_launch("adjoint_vjp");
}

}  // namespace

extern "C" {

int mjh_adjoint_tape_bytes(const mjhModel* m, const mjhData* d, unsigned long long* bytes) {
  if (!m || !d || !bytes) return fail(MJH_ERR_INVALID, "adjoint_tape_bytes: null argument");
  *bytes = tape_plan(*m, *d, nullptr, nullptr, nullptr);
  return MJH_OK;
}
int mjh_adjoint_tape_save(const mjhModel* m, const mjhData* d, void* slot, mjhStream stream) {
  if (!m || !d || !slot) return fail(MJH_ERR_INVALID, "adjoint_tape_save: null argument");
  TapeSegs segs;
  tape_plan(*m, *d, (char*)slot, nullptr, &segs);
  if (segs.n >= kTapeSegsMax) return fail(MJH_ERR_CAPACITY, "adjoint_tape_save: too many fields");
  hipLaunchKernelGGL(k_tape_copy, dim3(64, segs.n), dim3(256), 0, (hipStream_t)stream, segs, (char*)slot);
  return check_launch("adjoint_tape_save");
}
int mjh_adjoint_vjp_taped(const mjhModel* m, const mjhData* d, const void* slot, const float* qpos, const float* qvel,
                          const float* lam_next, const float* gadd, float* lam, float* grad_ctrl, mjhStream stream) {
  TRY(check_adjoint(m, d));
  if (!slot || !qpos || !qvel || !lam_next || !lam || (m->nu > 0 && !grad_ctrl))
    return fail(MJH_ERR_INVALID, "adjoint_vjp_taped: slot, qpos, qvel, lam_next, lam (and grad_ctrl) must be set");
  mjhData taped;
  tape_plan(*m, *d, (char*)slot, &taped, nullptr);
  return launch_vjp(*m, taped, qpos, qvel, lam_next, gadd, lam, grad_ctrl, (hipStream_t)stream);
}
int mjh_adjoint_vjp(const mjhModel* m, const mjhData* d, const float* qpos, const float* qvel, const float* lam_next,
                    const float* gadd, float* lam, float* grad_ctrl, mjhStream stream) {
  TRY(check_adjoint(m, d));
  if (!qpos || !qvel || !lam_next || !lam || (m->nu > 0 && !grad_ctrl))
    return fail(MJH_ERR_INVALID, "adjoint_vjp: qpos, qvel, lam_next, lam (and grad_ctrl with actuators) must be set");
  return launch_vjp(*m, *d, qpos, qvel, lam_next, gadd, lam, grad_ctrl, (hipStream_t)stream);
}
int mjh_adjoint_state_grad(const mjhModel* m, const mjhData* d, const float* lam, float* g_qpos, float* g_qvel,
                           float* g_act, mjhStream stream) {
  if (!m || !d) return fail(MJH_ERR_INVALID, "null model or data");
  if (!lam || !g_qpos || !g_qvel || (m->na > 0 && !g_act))
    return fail(MJH_ERR_INVALID, "adjoint_state_grad: lam and the gradients must be set");
  if (d->nworld <= 0) return MJH_OK;
  hipLaunchKernelGGL(k_adjoint_state_grad, dim3(d->nworld), dim3(64), 0, (hipStream_t)stream, *m, *d, lam, g_qpos,
                     g_qvel, g_act);
  return check_launch("adjoint_state_grad");
}

}  // extern "C"
