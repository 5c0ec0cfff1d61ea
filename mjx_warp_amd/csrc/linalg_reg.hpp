// Register-resident dense linear algebra for 32-lane teams (nv <= 32).
//
// A world's nv x nv symmetric positive definite matrix is held one row per
// lane (lane i owns row i in VGPRs), padded to NVC (a multiple of 4) with
// identity rows so that every loop has a compile-time trip count.  Block-
// diagonal matrices (one block per kinematic tree, as the reference's dense
// tiles, io.py:150-162) are factored as a whole: entries outside a lane's
// block are zero, so the factor's blocks equal the per-block factors.
//
// Team primitives used here and by the register solver: DPP sums, lane
// broadcasts and LDS broadcast buffers.  LDS instructions of one wavefront
// execute in issue order, so a write followed by reads of the same wavefront
// needs no barrier, only a compiler fence.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace mjh {

constexpr int kRegTeam = 32;
constexpr int kRegNvMax = 32;
// LDS floats a 32-lane team needs at least: RegChol's column broadcast
// buffer (one float per lane)
constexpr int kRegLdsMin = 32;

MJH_DEV void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  asm volatile("" ::: "memory");
}

// Value of v in lane k of this lane's team (k wave-uniform).
MJH_DEV float bcast32(float v, int k) {
  const float lo = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
  const float hi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k + 32));
  return (threadIdx.x & 32) ? hi : lo;
}

// p, opaque to the optimiser: address arithmetic on it is redone at each use
// instead of being hoisted out of a loop into live VGPRs.
typedef __attribute__((address_space(1))) const float gfloat;
MJH_DEV gfloat* opaque(const float* p) {
  gfloat* q = (gfloat*)p;
  asm volatile("" : "+v"(q));
  return q;
}
MJH_DEV int opaque(int i) {
  asm volatile("" : "+v"(i));
  return i;
}

// Calls f(std::integral_constant<int, NVC>) with NVC = nv rounded up to 4.
template <class F>
MJH_DEV auto with_nvc(int nv, F f) {
  switch ((nv + 3) / 4) {
    case 1: return f(std::integral_constant<int, 4>{});
    case 2: return f(std::integral_constant<int, 8>{});
    case 3: return f(std::integral_constant<int, 12>{});
    case 4: return f(std::integral_constant<int, 16>{});
    case 5: return f(std::integral_constant<int, 20>{});
    case 6: return f(std::integral_constant<int, 24>{});
    case 7: return f(std::integral_constant<int, 28>{});
    default: return f(std::integral_constant<int, 32>{});
  }
}

// The tree block [beg, end) holding `dof` (padding lanes: [dof, dof + 1)).
MJH_DEV void tree_block(const mjhModel& m, int dof, int* beg, int* end) {
  *beg = dof;
  *end = dof + 1;
  for_each_tree_block(m, [&](int d0, int n) {
    if (dof >= d0 && dof < d0 + n) {
      *beg = d0;
      *end = d0 + n;
    }
  });
}

template <int NVC>
struct RegChol {
  static_assert(NVC % 4 == 0 && NVC <= kRegNvMax, "padded dof count");
  int lane;     // 0..31: the row this lane owns
  float* cbuf;  // [kRegLdsMin] LDS column broadcast buffer of the team
  float h[NVC];  // matrix row, then factor row (lower part)
  float rdiag;   // 1 / L[lane][lane]

  // Row `lane` of a symmetric matrix read column-wise from global memory
  // (A[j * nv + lane], coalesced across the team), restricted to the lane's
  // tree block [beg, end); padding rows are identity rows.
  // `diag` is added to the lane's diagonal entry.
  MJH_DEV void load_sym(const float* A, int nv, int beg, int end, float diag = 0.0f) {
    const int ln = opaque(lane);
    gfloat* Ac = opaque(A) + (ln < nv ? ln : 0);
#pragma unroll
    for (int j = 0; j < NVC; j++) {
      const float v = Ac[(j < nv ? j : 0) * nv] + (j == ln ? diag : 0.0f);
      h[j] = (ln < nv && j >= beg && j < end) ? v : (j == ln ? 1.0f : 0.0f);
    }
    rdiag = 1.0f;
  }

  // A lower factor row read from global memory (L[lane * nv + j], j within
  // the lane's block, up to the diagonal).
  MJH_DEV void load_lower(const float* L, int nv, int beg) {
    const int ln = opaque(lane);
    const bool dof = ln < nv;
    gfloat* Lr = opaque(L) + (dof ? ln * nv : 0);
#pragma unroll
    for (int j = 0; j < NVC; j++) {
      const float v = Lr[j < nv ? j : 0];
      h[j] = (dof && j >= beg && j <= ln) ? v : (j == ln ? 1.0f : 0.0f);
    }
    rdiag = dof ? 1.0f / Lr[ln] : 1.0f;
  }

  // Cholesky H = L L' of the lane-owned rows (right-looking, one column
  // broadcast through LDS per step).  Lane i keeps L[i][0..i] in h[0..i]
  // and 1 / L[i][i] in rdiag.
  MJH_DEV void factor() {
#pragma unroll
    for (int k = 0; k < NVC; k++) {
      wave_fence();
      cbuf[lane] = h[k];
      wave_fence();
      const float dkk = cbuf[k];
      const float lkk = __builtin_amdgcn_sqrtf(dkk);
      const float inv = __builtin_amdgcn_rcpf(lkk);
      // L[i][k] L[j][k] = H[i][k] H[j][k] / dkk
      const float g = lane > k ? h[k] * (inv * inv) : 0.0f;
#pragma unroll
      for (int j = k + 1; j < NVC; j++) h[j] = fmaf(-g, cbuf[j], h[j]);
      if (lane == k) rdiag = inv;
      h[k] = lane > k ? h[k] * inv : (lane == k ? lkk : h[k]);
    }
  }

  // x = inv(L L') b with L from factor() / load_lower(); b, x one entry per
  // lane (padding lanes carry 0)
  MJH_DEV float solve(float b) const {
    // forward: y = inv(L) b, column by column
    float y = b;
#pragma unroll
    for (int k = 0; k < NVC; k++) {
      const float yk = bcast32(y * rdiag, k);
      y = lane > k ? fmaf(-h[k], yk, y) : y;
    }
    y *= rdiag;
    // backward: x = inv(L') y, x[k] = (y[k] - sum_{j>k} L[j][k] x[j]) / L[k][k]
    float x = 0.0f;
#pragma unroll
    for (int k = NVC - 1; k >= 0; k--) {
      const float s = sum32(lane > k ? h[k] * x : 0.0f);
      if (lane == k) x = (y - s) * rdiag;
    }
    return x;
  }

  // Writes the factor's rows into L (global, [nv, nv]): the lower triangle of
  // the lane's block [beg, end) and zeros above the diagonal inside the
  // block, as the reference's tile factor (smooth.py:353-379).
  MJH_DEV void store_block(float* L, int nv, int beg, int end) const {
    if (lane >= nv) return;
    float* Lr = L + lane * nv;
#pragma unroll
    for (int j = 0; j < NVC; j++)
      if (j >= beg && j < end) Lr[j] = j <= lane ? h[j] : 0.0f;
  }
};

// qLD = tree-block Cholesky factors of qM, in registers (32-lane teams)
MJH_DEV void reg_factor_blocks(const mjhModel& m, const float* M, float* L, float* lds, const Team<kRegTeam>& t) {
  const int nv = m.nv;
  int beg, end;
  tree_block(m, t.lane, &beg, &end);
  with_nvc(nv, [&](auto nvc) {
    RegChol<decltype(nvc)::value> c{t.lane, lds};
    c.load_sym(M, nv, beg, end);
    c.factor();
    c.store_block(L, nv, beg, end);
    return 0;
  });
}

// x = inv(M) y from the tree-block factor L (qLD), in registers; x may
// alias y
MJH_DEV void reg_solve_blocks(const mjhModel& m, const float* L, float* x, const float* y, const Team<kRegTeam>& t) {
  const int nv = m.nv;
  int beg, end;
  tree_block(m, t.lane, &beg, &end);
  const float b = t.lane < nv ? y[t.lane] : 0.0f;
  const float r = with_nvc(nv, [&](auto nvc) {
    RegChol<decltype(nvc)::value> c{t.lane, nullptr};
    c.load_lower(L, nv, beg);
    return c.solve(b);
  });
  if (t.lane < nv) x[t.lane] = r;
}

}  // namespace mjh
