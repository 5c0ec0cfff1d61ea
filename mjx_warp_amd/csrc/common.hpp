// Device-side helpers shared by every stage kernel: small fixed-size vector
// algebra, spatial algebra in MuJoCo's (rotation, translation) convention,
// and the per-world "team" abstraction.
//
// Execution model: a workgroup is one 64-lane wavefront and processes
// WPB = 64 / TEAM worlds; each world is handled by TEAM consecutive lanes.
// Stage functions loop `for (i = t.lane; i < n; i += TEAM)` over the
// world's bodies / dofs / rows and use sync() between dependent phases.
// Because a workgroup is a single wavefront, sync() is cheap and legal in
// world-divergent control flow.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mjx_warp_amd.h"

#define MJH_DEV __device__ __forceinline__

// Per-world base pointer into a world-major Data array.
#define WPTR(arr, w, stride) ((arr) + (size_t)(w) * (size_t)(stride))

namespace mjh {

struct V3 {
  float x, y, z;
};
struct Q4 {
  float w, x, y, z;
};
struct M33 {
  float m[9];  // row-major
};

MJH_DEV V3 v3(float x, float y, float z) { return V3{x, y, z}; }
MJH_DEV V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
MJH_DEV V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
MJH_DEV V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
MJH_DEV V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
MJH_DEV V3 operator*(float s, V3 a) { return V3{a.x * s, a.y * s, a.z * s}; }
MJH_DEV float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MJH_DEV V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
MJH_DEV float length(V3 a) { return sqrtf(dot(a, a)); }
MJH_DEV V3 normalize(V3 a) {
  float l = length(a);
  return l > 0.0f ? a * (1.0f / l) : a;
}
MJH_DEV float vget(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

MJH_DEV V3 ld3(const float* p) { return V3{p[0], p[1], p[2]}; }
MJH_DEV void st3(float* p, V3 a) {
  p[0] = a.x;
  p[1] = a.y;
  p[2] = a.z;
}
MJH_DEV Q4 ld4(const float* p) { return Q4{p[0], p[1], p[2], p[3]}; }
MJH_DEV void st4(float* p, Q4 q) {
  p[0] = q.w;
  p[1] = q.x;
  p[2] = q.y;
  p[3] = q.z;
}
MJH_DEV M33 ld9(const float* p) {
  M33 r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.m[i] = p[i];
  return r;
}
MJH_DEV void st9(float* p, const M33& a) {
#pragma unroll
  for (int i = 0; i < 9; i++) p[i] = a.m[i];
}

// matrix * vector
MJH_DEV V3 mulmv(const M33& a, V3 v) {
  return V3{a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z,
            a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
            a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z};
}
// transpose(matrix) * vector
MJH_DEV V3 mulmtv(const M33& a, V3 v) {
  return V3{a.m[0] * v.x + a.m[3] * v.y + a.m[6] * v.z,
            a.m[1] * v.x + a.m[4] * v.y + a.m[7] * v.z,
            a.m[2] * v.x + a.m[5] * v.y + a.m[8] * v.z};
}
MJH_DEV V3 mcol(const M33& a, int c) { return V3{a.m[c], a.m[3 + c], a.m[6 + c]}; }

// Quaternion product u*v (math.py:23-30 conventions, w first).
MJH_DEV Q4 mul_quat(Q4 u, Q4 v) {
  return Q4{u.w * v.w - u.x * v.x - u.y * v.y - u.z * v.z,
            u.w * v.x + u.x * v.w + u.y * v.z - u.z * v.y,
            u.w * v.y - u.x * v.z + u.y * v.w + u.z * v.x,
            u.w * v.z + u.x * v.y - u.y * v.x + u.z * v.w};
}
// Rotate vec by quat: 2(u.v)u + (s^2 - u.u)v + 2s(u x v)   (math.py:33-38)
MJH_DEV V3 rot_vec_quat(V3 vec, Q4 q) {
  V3 u = V3{q.x, q.y, q.z};
  float s = q.w;
  V3 r = (2.0f * dot(u, vec)) * u + (s * s - dot(u, u)) * vec;
  return r + (2.0f * s) * cross(u, vec);
}
MJH_DEV Q4 axis_angle_to_quat(V3 axis, float angle) {
  float s = sinf(angle * 0.5f), c = cosf(angle * 0.5f);
  return Q4{c, axis.x * s, axis.y * s, axis.z * s};
}
// 3x3 rotation matrix of a unit quaternion (math.py:48-64)
MJH_DEV M33 quat_to_mat(Q4 q) {
  float q00 = q.w * q.w, q11 = q.x * q.x, q22 = q.y * q.y, q33 = q.z * q.z;
  float q01 = q.w * q.x, q02 = q.w * q.y, q03 = q.w * q.z;
  float q12 = q.x * q.y, q13 = q.x * q.z, q23 = q.y * q.z;
  M33 r;
  r.m[0] = q00 + q11 - q22 - q33;
  r.m[1] = 2.0f * (q12 - q03);
  r.m[2] = 2.0f * (q13 + q02);
  r.m[3] = 2.0f * (q12 + q03);
  r.m[4] = q00 - q11 + q22 - q33;
  r.m[5] = 2.0f * (q23 - q01);
  r.m[6] = 2.0f * (q13 - q02);
  r.m[7] = 2.0f * (q23 + q01);
  r.m[8] = q00 - q11 - q22 + q33;
  return r;
}
MJH_DEV Q4 quat_inv(Q4 q) { return Q4{q.w, -q.x, -q.y, -q.z}; }
MJH_DEV Q4 qnormalize(Q4 q) {
  float l = sqrtf(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  if (l > 0.0f) {
    float s = 1.0f / l;
    return Q4{q.w * s, q.x * s, q.y * s, q.z * s};
  }
  return q;
}
// quaternion -> 3D rotation vector (math.py:115-128)
MJH_DEV V3 quat_to_vel(Q4 q) {
  V3 axis = V3{q.x, q.y, q.z};
  float s = length(axis);
  if (s == 0.0f) return V3{0.0f, 0.0f, 0.0f};
  float speed = 2.0f * atan2f(s, q.w);
  if (speed > 3.14159265358979f) speed -= 2.0f * 3.14159265358979f;
  return axis * (speed / s);
}
// qb * quat(res) = qa  (math.py:131-139)
MJH_DEV V3 quat_sub(Q4 qa, Q4 qb) { return quat_to_vel(mul_quat(quat_inv(qb), qa)); }
// q * exp(v dt) with local-frame angular velocity, renormalised (math.py:142-152)
MJH_DEV Q4 quat_integrate(Q4 q, V3 v, float dt) {
  float n = length(v);
  V3 axis = normalize(v);
  Q4 r = mul_quat(q, axis_angle_to_quat(axis, dt * n));
  return qnormalize(r);
}

// ---- spatial algebra: 6-vectors (rot[3], lin[3]); inertia 10-vectors
// (Ixx Iyy Izz Ixy Ixz Iyz, m*c[3], m) ----
struct S6 {
  float v[6];
};
MJH_DEV S6 ld6(const float* p) {
  S6 r;
#pragma unroll
  for (int i = 0; i < 6; i++) r.v[i] = p[i];
  return r;
}
MJH_DEV void st6(float* p, const S6& a) {
#pragma unroll
  for (int i = 0; i < 6; i++) p[i] = a.v[i];
}
MJH_DEV float dot6(const S6& a, const S6& b) {
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < 6; i++) s += a.v[i] * b.v[i];
  return s;
}
// inertia * motion (mju_mulInertVec; math.py:72-82)
MJH_DEV S6 inert_vec(const float* i, const S6& v) {
  S6 r;
  r.v[0] = i[0] * v.v[0] + i[3] * v.v[1] + i[4] * v.v[2] - i[8] * v.v[4] + i[7] * v.v[5];
  r.v[1] = i[3] * v.v[0] + i[1] * v.v[1] + i[5] * v.v[2] + i[8] * v.v[3] - i[6] * v.v[5];
  r.v[2] = i[4] * v.v[0] + i[5] * v.v[1] + i[2] * v.v[2] - i[7] * v.v[3] + i[6] * v.v[4];
  r.v[3] = i[8] * v.v[1] - i[7] * v.v[2] + i[9] * v.v[3];
  r.v[4] = i[6] * v.v[2] - i[8] * v.v[0] + i[9] * v.v[4];
  r.v[5] = i[7] * v.v[0] - i[6] * v.v[1] + i[9] * v.v[5];
  return r;
}
// motion x motion (math.py:85-97)
MJH_DEV S6 motion_cross(const S6& u, const S6& v) {
  V3 u0{u.v[0], u.v[1], u.v[2]}, u1{u.v[3], u.v[4], u.v[5]};
  V3 v0{v.v[0], v.v[1], v.v[2]}, v1{v.v[3], v.v[4], v.v[5]};
  V3 a = cross(u0, v0);
  V3 l = cross(u1, v0) + cross(u0, v1);
  return S6{{a.x, a.y, a.z, l.x, l.y, l.z}};
}
// motion x force (math.py:100-112)
MJH_DEV S6 motion_cross_force(const S6& v, const S6& f) {
  V3 v0{v.v[0], v.v[1], v.v[2]}, v1{v.v[3], v.v[4], v.v[5]};
  V3 f0{f.v[0], f.v[1], f.v[2]}, f1{f.v[3], f.v[4], f.v[5]};
  V3 a = cross(v0, f0) + cross(v1, f1);
  V3 l = cross(v0, f1);
  return S6{{a.x, a.y, a.z, l.x, l.y, l.z}};
}

// ---- contact frame helpers (math.py:155-188) ----
MJH_DEV M33 make_frame(V3 a) {
  a = normalize(a);
  V3 b = (-0.5f < a.y && a.y < 0.5f) ? V3{0.0f, 1.0f, 0.0f} : V3{0.0f, 0.0f, 1.0f};
  b = b - a * dot(a, b);
  b = normalize(b);
  if (length(a) == 0.0f) b = V3{0.0f, 0.0f, 0.0f};
  V3 c = cross(a, b);
  M33 r;
  r.m[0] = a.x; r.m[1] = a.y; r.m[2] = a.z;
  r.m[3] = b.x; r.m[4] = b.y; r.m[5] = b.z;
  r.m[6] = c.x; r.m[7] = c.y; r.m[8] = c.z;
  return r;
}

// ---- team of lanes working on one world ----
template <int TEAM>
struct Team {
  static constexpr int size = TEAM;
  static constexpr int wpb = 64 / TEAM;
  int lane;   // 0..TEAM-1
  int slot;   // world slot inside the workgroup
  int world;  // global world index (may be >= nworld for padding slots)
  bool valid;
  MJH_DEV Team(int nworld) : Team(nworld, blockIdx.x) {}
  // block: the workgroup's index among those covering these nworld worlds
  MJH_DEV Team(int nworld, int block) {
    lane = threadIdx.x % TEAM;
    slot = threadIdx.x / TEAM;
    world = block * wpb + slot;
    valid = world < nworld;
  }
};

// Barrier for the workgroup (= one wavefront): orders LDS and global memory
// accesses of all lanes.
MJH_DEV void sync() { __syncthreads(); }

template <int CTRL>
MJH_DEV float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
MJH_DEV int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}

// Sum over each 32-lane half of the wavefront (all lanes of a half get
// identical bits: every step adds the same two rounded values in both lanes
// of a pair).  The input is materialised first, so a product passed in is
// rounded before the butterfly: contracting it into the first add as
// fma(a, b, partner) would give the two lanes of a pair different bits, and
// team-wide decisions taken on the sum would diverge.
MJH_DEV float sum32(float v) {
  asm volatile("" : "+v"(v));
  v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
  v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
  v += dpp_f<0x141>(v);  // row_half_mirror: 8-lane halves of a row
  v += dpp_f<0x140>(v);  // row_mirror: the two 8-lane halves of a row
  // bitmask swizzle inside 32-lane groups: and 0x1f, or 0, xor 0x10
  v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));
  return v;
}
MJH_DEV int sum32_int(int v) {
  v += dpp_i<0xB1>(v);
  v += dpp_i<0x4E>(v);
  v += dpp_i<0x141>(v);
  v += dpp_i<0x140>(v);
  v += __builtin_amdgcn_ds_swizzle(v, 0x401F);
  return v;
}

// Sum of v over the TEAM lanes of a team; every lane gets the same bits
// (64-lane teams: the two half sums, read as wave-uniform values, added in
// one order).
template <int TEAM>
MJH_DEV float team_sum(float v) {
  static_assert(TEAM == 32 || TEAM == 64, "team size");
  const float s = sum32(v);
  if constexpr (TEAM == 32) {
    return s;
  } else {
    const float lo = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s), 0));
    const float hi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s), 32));
    return lo + hi;
  }
}
template <int TEAM>
MJH_DEV int team_sum_int(int v) {
  static_assert(TEAM == 32 || TEAM == 64, "team size");
  const int s = sum32_int(v);
  if constexpr (TEAM == 32) return s;
  else return __builtin_amdgcn_readlane(s, 0) + __builtin_amdgcn_readlane(s, 32);
}

// Reserves `total` consecutive entries of a global pool for this team and
// returns the first: one atomicAdd on `counter` per wavefront (the teams of
// a wavefront take adjacent blocks).  total is team-uniform (0 for teams
// without a world); every lane of the wavefront must call it.
template <int TEAM>
MJH_DEV int team_reserve(int* counter, int total) {
  const int t0 = __builtin_amdgcn_readlane(total, 0);
  const int t1 = TEAM == 64 ? 0 : __builtin_amdgcn_readlane(total, 32);
  int base = 0;
  if ((threadIdx.x & 63) == 0 && t0 + t1 > 0) base = atomicAdd(counter, t0 + t1);
  base = __builtin_amdgcn_readlane(base, 0);
  return (TEAM == 32 && (threadIdx.x & 32)) ? base + t0 : base;
}

// Three blocks of three pools in one round trip: the wavefront's first lane
// issues the three atomics back to back and waits once; each team gets its
// bases in *b0..*b2 (as team_reserve).  Every lane of the wavefront must
// call it.
template <int TEAM>
MJH_DEV void team_reserve3(int* c0, int n0, int* c1, int n1, int* c2, int n2, int* b0, int* b1, int* b2) {
  const int a0 = __builtin_amdgcn_readlane(n0, 0), a1 = __builtin_amdgcn_readlane(n1, 0),
            a2 = __builtin_amdgcn_readlane(n2, 0);
  const int z0 = TEAM == 64 ? 0 : __builtin_amdgcn_readlane(n0, 32);
  const int z1 = TEAM == 64 ? 0 : __builtin_amdgcn_readlane(n1, 32);
  const int z2 = TEAM == 64 ? 0 : __builtin_amdgcn_readlane(n2, 32);
  int r0 = 0, r1 = 0, r2 = 0;
  if ((threadIdx.x & 63) == 0) {
    r0 = atomicAdd(c0, a0 + z0);
    r1 = atomicAdd(c1, a1 + z1);
    r2 = atomicAdd(c2, a2 + z2);
  }
  r0 = __builtin_amdgcn_readlane(r0, 0);
  r1 = __builtin_amdgcn_readlane(r1, 0);
  r2 = __builtin_amdgcn_readlane(r2, 0);
  const bool upper = TEAM == 32 && (threadIdx.x & 32);
  *b0 = upper ? r0 + a0 : r0;
  *b1 = upper ? r1 + a1 : r1;
  *b2 = upper ? r2 + a2 : r2;
}

// This lane's team as a mask over the wavefront's lanes.
template <int TEAM>
MJH_DEV unsigned long long team_lanes() {
  if constexpr (TEAM == 64) return ~0ull;
  else return 0xffffffffull << (threadIdx.x & 32);
}

// Exclusive prefix sum over the team (lane order) of v in [0, 2^BITS) and
// the team's total, from one ballot per bit.
template <int TEAM, int BITS>
MJH_DEV int team_exscan_bits(int v, int* total) {
  const unsigned long long team = team_lanes<TEAM>();
  const unsigned long long below = __lanemask_lt() & team;
  int off = 0, tot = 0;
#pragma unroll
  for (int b = 0; b < BITS; b++) {
    const unsigned long long m = __ballot((v >> b) & 1) & team;
    off += __popcll(m & below) << b;
    tot += __popcll(m) << b;
  }
  *total = tot;
  return off;
}
// Exclusive prefix sum over the TEAM lanes (lane order); *total gets the sum.
template <int TEAM>
MJH_DEV int team_exscan_int(int v, int lane, int* total) {
  int incl = v;
#pragma unroll
  for (int o = 1; o < TEAM; o <<= 1) {
    int t = __shfl_up(incl, o, TEAM);
    if (lane >= o) incl += t;
  }
  *total = __shfl(incl, TEAM - 1, TEAM);
  return incl - v;
}

// Iterates the dense per-tree blocks (qLD_tile / qLD_tileadr / qLD_tilesize,
// io.py:150-162) and calls f(dofid, size) for each.
template <class F>
MJH_DEV void for_each_tree_block(const mjhModel& m, F f) {
  for (int g = 0; g < m.nqLD_group; g++) {
    const int beg = m.qLD_tileadr[g];
    const int end = (g == m.nqLD_group - 1) ? m.nqLD_tile : m.qLD_tileadr[g + 1];
    const int size = m.qLD_tilesize[g];
    for (int k = beg; k < end; k++) f(m.qLD_tile[k], size);
  }
}

}  // namespace mjh
