// Smooth dynamics of one world, executed by a team of lanes: forward
// kinematics, subtree COM / com-frame inertias / motion axes, composite
// rigid-body inertia and the dense joint-space mass matrix, its Cholesky
// factor and solves, RNE bias forces, body velocities, passive forces and
// actuator transmission.  Each function reproduces the arithmetic of the
// reference stage cited in its comment; the reference's per-level launches
// become per-level loops separated by sync(), and its atomic accumulations
// become per-lane sequential sums (no two lanes write the same word).
//
// The tree passes keep the quantities they read back (parent poses, subtree
// sums, composite inertias, body accelerations and forces) in the team's
// LDS: each level or accumulation step then waits on an LDS access instead
// of a global-memory round trip.  Data fields are still written to global
// memory once each; the arithmetic and its order are unchanged.
#pragma once

#include "common.hpp"
#include "linalg_reg.hpp"

namespace mjh {

MJH_DEV int level_end(const mjhModel& m, int l) {
  return (l == m.nlevel - 1) ? m.nbody : m.body_treeadr[l + 1];
}

// LDS floats per world of the tree passes below: kinematics (poses; also
// enough for com_pos), crb (composite inertias, motion axes, dof parents
// and the mass matrix), com_vel / rne (velocities, accelerations, forces)
__host__ __device__ inline int lds_kinematics_floats(int nbody) { return 16 * nbody; }
// 32-lane teams build the mass matrix in LDS; 64-lane teams write it to
// qMd directly (global: no nv^2 of LDS)
__host__ __device__ inline int lds_crb_floats(int nbody, int nv, bool lds_matrix = true) {
  return 10 * nbody + 7 * nv + (lds_matrix ? nv * nv : 0);
}
__host__ __device__ inline int lds_velocity_floats(int nbody) { return 12 * nbody; }

// ---------------------------------------------------------------------------
// smooth.py:32-145 kinematics
// lds: lds_kinematics_floats; holds xpos / xquat / xmat of the bodies done so far
template <int TEAM>
MJH_DEV void kinematics(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  const int w = t.world;
  const int nbody = m.nbody;
  float* xpos = WPTR(d.xpos, w, nbody * 3);
  float* xquat = WPTR(d.xquat, w, nbody * 4);
  float* xmat = WPTR(d.xmat, w, nbody * 9);
  float* xipos = WPTR(d.xipos, w, nbody * 3);
  float* ximat = WPTR(d.ximat, w, nbody * 9);
  float* xanchor = WPTR(d.xanchor, w, m.njnt * 3);
  float* xaxis = WPTR(d.xaxis, w, m.njnt * 3);
  const float* qpos = WPTR(d.qpos, w, m.nq);
  float* lpos = lds;
  float* lquat = lds + 3 * nbody;
  float* lmat = lds + 7 * nbody;

  if (t.valid && t.lane == 0) {
    st3(xpos, V3{0.f, 0.f, 0.f});
    st4(xquat, Q4{1.f, 0.f, 0.f, 0.f});
    st3(xipos, V3{0.f, 0.f, 0.f});
    M33 I = {{1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f}};
    st9(xmat, I);
    st9(ximat, I);
    st3(lpos, V3{0.f, 0.f, 0.f});
    st4(lquat, Q4{1.f, 0.f, 0.f, 0.f});
    st9(lmat, I);
  }
  sync();
  for (int l = 1; l < m.nlevel; l++) {
    const int beg = m.body_treeadr[l], end = level_end(m, l);
    for (int k = beg + t.lane; k < end && t.valid; k += TEAM) {
      const int b = m.body_tree[k];
      int jntadr = m.body_jntadr[b];
      const int jntnum = m.body_jntnum[b];
      V3 pos;
      Q4 quat;
      if (jntnum == 1 && m.jnt_type[jntadr] == mjhJNT_FREE) {
        const int qadr = m.jnt_qposadr[jntadr];
        pos = ld3(qpos + qadr);
        quat = ld4(qpos + qadr + 3);
        st3(xanchor + 3 * jntadr, pos);
        st3(xaxis + 3 * jntadr, ld3(m.jnt_axis + 3 * jntadr));
      } else {
        const int pid = m.body_parentid[b];
        pos = mulmv(ld9(lmat + 9 * pid), ld3(m.body_pos + 3 * b)) + ld3(lpos + 3 * pid);
        quat = mul_quat(ld4(lquat + 4 * pid), ld4(m.body_quat + 4 * b));
        for (int j = 0; j < jntnum; j++, jntadr++) {
          const int qadr = m.jnt_qposadr[jntadr];
          const int jt = m.jnt_type[jntadr];
          const V3 jaxis = ld3(m.jnt_axis + 3 * jntadr);
          const V3 jpos = ld3(m.jnt_pos + 3 * jntadr);
          const V3 anchor = rot_vec_quat(jpos, quat) + pos;
          const V3 axis = rot_vec_quat(jaxis, quat);
          if (jt == mjhJNT_BALL) {
            quat = mul_quat(quat, ld4(qpos + qadr));
            pos = anchor - rot_vec_quat(jpos, quat);
          } else if (jt == mjhJNT_SLIDE) {
            pos = pos + axis * (qpos[qadr] - m.qpos0[qadr]);
          } else if (jt == mjhJNT_HINGE) {
            quat = mul_quat(quat, axis_angle_to_quat(jaxis, qpos[qadr] - m.qpos0[qadr]));
            pos = anchor - rot_vec_quat(jpos, quat);
          }
          st3(xanchor + 3 * jntadr, anchor);
          st3(xaxis + 3 * jntadr, axis);
        }
      }
      st3(xpos + 3 * b, pos);
      st3(lpos + 3 * b, pos);
      quat = qnormalize(quat);
      st4(xquat + 4 * b, quat);
      st4(lquat + 4 * b, quat);
      const M33 mat = quat_to_mat(quat);
      st9(xmat + 9 * b, mat);
      st9(lmat + 9 * b, mat);
      st3(xipos + 3 * b, pos + rot_vec_quat(ld3(m.body_ipos + 3 * b), quat));
      st9(ximat + 9 * b, quat_to_mat(mul_quat(quat, ld4(m.body_iquat + 4 * b))));
    }
    sync();
  }
  if (t.valid) {
    float* gxpos = WPTR(d.geom_xpos, w, m.ngeom * 3);
    float* gxmat = WPTR(d.geom_xmat, w, m.ngeom * 9);
    for (int g = t.lane; g < m.ngeom; g += TEAM) {
      const int b = m.geom_bodyid[g];
      const Q4 q = ld4(lquat + 4 * b);
      st3(gxpos + 3 * g, ld3(lpos + 3 * b) + rot_vec_quat(ld3(m.geom_pos + 3 * g), q));
      st9(gxmat + 9 * g, quat_to_mat(mul_quat(q, ld4(m.geom_quat + 4 * g))));
    }
    float* sxpos = WPTR(d.site_xpos, w, m.nsite * 3);
    float* sxmat = WPTR(d.site_xmat, w, m.nsite * 9);
    for (int s = t.lane; s < m.nsite; s += TEAM) {
      const int b = m.site_bodyid[s];
      const Q4 q = ld4(lquat + 4 * b);
      st3(sxpos + 3 * s, ld3(lpos + 3 * b) + rot_vec_quat(ld3(m.site_pos + 3 * s), q));
      st9(sxmat + 9 * s, quat_to_mat(mul_quat(q, ld4(m.site_quat + 4 * s))));
    }
  }
  sync();
}

// ---------------------------------------------------------------------------
// smooth.py:148-245 com_pos
// lds: 3 * nbody floats; holds the subtree sums / centres of mass
template <int TEAM>
MJH_DEV void com_pos(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  const int w = t.world;
  const int nbody = m.nbody;
  float* scw = WPTR(d.subtree_com, w, nbody * 3);
  float* sc = lds;
  const float* xipos = WPTR(d.xipos, w, nbody * 3);
  if (t.valid)
    for (int b = t.lane; b < nbody; b += TEAM) st3(sc + 3 * b, ld3(xipos + 3 * b) * m.body_mass[b]);
  sync();
  // accumulate subtrees into parents; the world body adds itself once more,
  // which the doubled subtree_mass[0] cancels (io.py:249-252, smooth.py:238-241)
  if (t.valid && t.lane < 3) {
    const int k = t.lane;
    for (int b = nbody - 1; b >= 0; b--) sc[3 * m.body_parentid[b] + k] += sc[3 * b + k];
  }
  sync();
  if (t.valid)
    for (int b = t.lane; b < nbody; b += TEAM) {
      const V3 c = ld3(sc + 3 * b) * (1.0f / m.subtree_mass[b]);
      st3(sc + 3 * b, c);
      st3(scw + 3 * b, c);
    }
  sync();
  if (t.valid) {
    const float* ximat = WPTR(d.ximat, w, nbody * 9);
    float* cinert = WPTR(d.cinert, w, nbody * 10);
    for (int b = t.lane; b < nbody; b += TEAM) {
      const M33 mat = ld9(ximat + 9 * b);
      const V3 inert = ld3(m.body_inertia + 3 * b);
      const float mass = m.body_mass[b];
      const V3 dif = ld3(xipos + 3 * b) - ld3(sc + 3 * m.body_rootid[b]);
      // mat * diag(inert) * mat'
      float tmp[3][3];
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
          tmp[i][j] = mat.m[3 * i] * inert.x * mat.m[3 * j] + mat.m[3 * i + 1] * inert.y * mat.m[3 * j + 1] +
                      mat.m[3 * i + 2] * inert.z * mat.m[3 * j + 2];
      float* r = cinert + 10 * b;
      r[0] = tmp[0][0] + mass * (dif.y * dif.y + dif.z * dif.z);
      r[1] = tmp[1][1] + mass * (dif.x * dif.x + dif.z * dif.z);
      r[2] = tmp[2][2] + mass * (dif.x * dif.x + dif.y * dif.y);
      r[3] = tmp[0][1] - mass * dif.x * dif.y;
      r[4] = tmp[0][2] - mass * dif.x * dif.z;
      r[5] = tmp[1][2] - mass * dif.y * dif.z;
      r[6] = mass * dif.x;
      r[7] = mass * dif.y;
      r[8] = mass * dif.z;
      r[9] = mass;
    }
    const float* xaxis = WPTR(d.xaxis, w, m.njnt * 3);
    const float* xanchor = WPTR(d.xanchor, w, m.njnt * 3);
    const float* xmat = WPTR(d.xmat, w, nbody * 9);
    float* cdof = WPTR(d.cdof, w, m.nv * 6);
    for (int j = t.lane; j < m.njnt; j += TEAM) {
      const int b = m.jnt_bodyid[j];
      const int da = m.jnt_dofadr[j];
      const int jt = m.jnt_type[j];
      const V3 axis = ld3(xaxis + 3 * j);
      const M33 xm = ld9(xmat + 9 * b);
      const V3 offset = ld3(sc + 3 * m.body_rootid[b]) - ld3(xanchor + 3 * j);
      if (jt == mjhJNT_FREE || jt == mjhJNT_BALL) {
        int r0 = da;
        if (jt == mjhJNT_FREE) {
          for (int i = 0; i < 3; i++) {
            float* c = cdof + 6 * (da + i);
            for (int k = 0; k < 6; k++) c[k] = 0.0f;
            c[3 + i] = 1.0f;
          }
          r0 = da + 3;
        }
        for (int i = 0; i < 3; i++) {
          const V3 ax = mcol(xm, i);  // row i of xmat' = column i of xmat
          const V3 lin = cross(ax, offset);
          float* c = cdof + 6 * (r0 + i);
          st3(c, ax);
          st3(c + 3, lin);
        }
      } else if (jt == mjhJNT_SLIDE) {
        float* c = cdof + 6 * da;
        st3(c, V3{0.f, 0.f, 0.f});
        st3(c + 3, axis);
      } else {
        float* c = cdof + 6 * da;
        st3(c, axis);
        st3(c + 3, cross(axis, offset));
      }
    }
  }
  sync();
}

// ---------------------------------------------------------------------------
// smooth.py:248-314 crb (dense qM; the sparse layout is not built)
// MuJoCo's sparse layout of the dense nv x nv matrix A into S (smooth.py:
// 281-294 qM_sparse): row i holds A(i, j) for j = i, parent(i), ... from
// dof_Madr[i]; lanes over rows
template <int TEAM>
MJH_DEV void dense_to_sparse(const mjhModel& m, const float* A, float* S, const Team<TEAM>& t) {
  for (int i = t.lane; i < m.nv; i += TEAM) {
    int a = m.dof_Madr[i];
    for (int j = i; j >= 0; j = m.dof_parentid[j]) S[a++] = A[i * m.nv + j];
  }
}

// smooth.py:317-350 _factor_i_sparse: the L'DL factor of the sparse matrix S
// (global) into L (unit L below the diagonal, D on it) and Dinv = 1 / D,
// leaves first as MuJoCo; worked in LDS (nM floats), lanes over the entries
// of each updated row.  Every world runs the same loop (model-only bounds).
template <int TEAM>
MJH_DEV void factor_sparse(const mjhModel& m, const float* S, float* L, float* Dinv, float* lds,
                           const Team<TEAM>& t) {
  const int nv = m.nv, nM = m.nM;
  if (t.valid)
    for (int k = t.lane; k < nM; k += TEAM) lds[k] = S[k];
  sync();
  for (int k = nv - 1; k >= 0; k--) {
    const int ak = m.dof_Madr[k];
    int p = 1;
    for (int i = m.dof_parentid[k]; i >= 0; i = m.dof_parentid[i], p++) {
      const float tmp = lds[ak + p] / lds[ak];
      const int ai = m.dof_Madr[i];
      const int len = m.dof_Madr[i + 1] - ai;  // i < k <= nv - 1
      for (int q = t.lane; q < len; q += TEAM) lds[ai + q] -= lds[ak + p + q] * tmp;
      sync();
      if (t.lane == 0) lds[ak + p] = tmp;
      sync();
    }
  }
  if (t.valid) {
    for (int k = t.lane; k < nM; k += TEAM) L[k] = lds[k];
    for (int k = t.lane; k < nv; k += TEAM) Dinv[k] = 1.0f / lds[m.dof_Madr[k]];
  }
  sync();
}

// lds: lds_crb_floats; holds crb, cdof, dof_parentid and, on 32-lane
// teams, the mass matrix, which is written to qMd row by row at the end (and,
// for sparse models, to qM in MuJoCo's sparse layout).  64-lane teams write
// the entries straight into qMd: the tree blocks are cleared first (entries
// of dofs in different trees are zero for good and never written), then
// every dof writes its row's ancestor entries and their mirror images -- the
// same values, with no nv^2 of LDS.
template <int TEAM>
MJH_DEV void crb(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  constexpr bool in_lds = TEAM != 64;
  const int w = t.world;
  const int nbody = m.nbody, nv = m.nv;
  float* crbw = WPTR(d.crb, w, nbody * 10);
  const float* cinert = WPTR(d.cinert, w, nbody * 10);
  const float* cdofw = WPTR(d.cdof, w, nv * 6);
  float* lcrb = lds;
  float* lcdof = lds + 10 * nbody;
  int* lpar = (int*)(lcdof + 6 * nv);
  float* const qMg = WPTR(d.qMd, w, nv * nv);
  float* lqM = in_lds ? lcdof + 7 * nv : qMg;
  if (t.valid) {
    for (int i = t.lane; i < nbody * 10; i += TEAM) lcrb[i] = cinert[i];
    for (int i = t.lane; i < nv * 6; i += TEAM) lcdof[i] = cdofw[i];
    for (int i = t.lane; i < nv; i += TEAM) lpar[i] = m.dof_parentid[i];
    if (in_lds) {
      for (int i = t.lane; i < nv * nv; i += TEAM) lqM[i] = 0.0f;
    } else {
      for_each_tree_block(m, [&](int dof, int n) {
        for (int p = t.lane; p < n * n; p += TEAM) lqM[(dof + p / n) * nv + dof + p % n] = 0.0f;
      });
    }
  }
  sync();
  // composite inertia: children into parents, bodies attached to the world
  // are not added to the world body (smooth.py:255-261)
  if (t.valid && t.lane < 10) {
    const int k = t.lane;
    for (int b = nbody - 1; b >= 1; b--) {
      const int p = m.body_parentid[b];
      if (p != 0) lcrb[10 * p + k] += lcrb[10 * b + k];
    }
  }
  sync();
  if (t.valid) {
    for (int i = t.lane; i < nbody * 10; i += TEAM) crbw[i] = lcrb[i];
    for (int i = t.lane; i < nv; i += TEAM) {
      const S6 ci = ld6(lcdof + 6 * i);
      const S6 buf = inert_vec(lcrb + 10 * m.dof_bodyid[i], ci);
      lqM[i * nv + i] = m.dof_armature[i] + dot6(ci, buf);
      for (int j = lpar[i]; j >= 0; j = lpar[j]) {
        const float v = dot6(ld6(lcdof + 6 * j), buf);
        lqM[i * nv + j] = v;
        lqM[j * nv + i] = v;
      }
    }
  }
  sync();
  if (t.valid) {
    if (in_lds)
      for (int i = t.lane; i < nv * nv; i += TEAM) qMg[i] = lqM[i];
    if (m.opt.is_sparse) dense_to_sparse(m, lqM, WPTR(d.qM, w, m.nM), t);
  }
  sync();
}

// ---------------------------------------------------------------------------
// Dense Cholesky of an n x n block (row stride ld) held in LDS (or, for a
// wide Newton Hessian, in global memory): A = L L', L overwrites the lower
// triangle and the strict upper triangle is zeroed, matching
// wp.tile_cholesky as used by smooth.py:361-368.
template <int TEAM>
MJH_DEV void chol_lds(float* A, int n, int ld, const mjhModel& m, const Team<TEAM>& t) {
  (void)m;
  for (int k = 0; k < n; k++) {
    const float dkk = sqrtf(A[k * ld + k]);
    const float inv = 1.0f / dkk;
    sync();
    for (int i = k + 1 + t.lane; i < n; i += TEAM) A[i * ld + k] *= inv;
    if (t.lane == 0) A[k * ld + k] = dkk;
    sync();
    // trailing update of the lower triangle of rows/cols k+1..n-1: lane
    // per column c, down the rows r >= c, the scaled column k read as a
    // broadcast (same entry for every lane) and the lanes' entries of a row
    // adjacent (coalesced / conflict-free); each entry takes the same single
    // update per k as in any order of the pairs
    for (int c = k + 1 + t.lane; c < n; c += TEAM) {
      const float lck = A[c * ld + k];
      int r = c;
      // eight rows' loads issued before their stores: the compiler cannot
      // tell that column c's stores miss column k's loads
      for (; r + 8 <= n; r += 8) {
        float a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          a[u] = A[(r + u) * ld + k];
          b[u] = A[(r + u) * ld + c];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) A[(r + u) * ld + c] = b[u] - a[u] * lck;
      }
      for (; r < n; r++) A[r * ld + c] -= A[r * ld + k] * lck;
    }
    sync();
  }
  for (int p = t.lane; p < n * n; p += TEAM) {
    const int r = p / n, c = p % n;
    if (c > r) A[r * ld + c] = 0.0f;
  }
  sync();
}

// chol_lds for a matrix in global memory (a wide Newton Hessian), in
// panels of kCholPanel columns: the panel's diagonal block is factored in
// LDS by chol_lds, the rows below it are solved against that block (lane
// per row, the row's panel entries in registers) and kept in LDS, and the
// trailing lower triangle takes their products lane per column.  Every
// entry gets the same operations in the same order as in chol_lds (so the
// factor is the same bit for bit) with 1/kCholPanel of the global-memory
// traffic.  lds: chol_blocked_floats(n).
constexpr int kCholPanel = 16;
__host__ __device__ inline int chol_blocked_floats(int n) { return kCholPanel * kCholPanel + kCholPanel * n; }

template <int TEAM>
MJH_DEV void chol_blocked(float* A, int n, int ld, float* lds, const mjhModel& m, const Team<TEAM>& t) {
  constexpr int B = kCholPanel;
  float* D = lds;          // the diagonal block, row stride nb
  float* P = lds + B * B;  // the panel rows below it, row stride B
  for (int p = 0; p < n; p += B) {
    const int nb = n - p < B ? n - p : B;
    for (int e = t.lane; e < nb * nb; e += TEAM) D[e] = A[(p + e / nb) * ld + p + e % nb];
    sync();
    chol_lds(D, nb, nb, m, t);
    for (int e = t.lane; e < nb * nb; e += TEAM) A[(p + e / nb) * ld + p + e % nb] = D[e];
    for (int i = p + nb + t.lane; i < n; i += TEAM) {
      float x[B];
#pragma unroll
      for (int k = 0; k < B; k++) x[k] = k < nb ? A[i * ld + p + k] : 0.0f;
#pragma unroll
      for (int k = 0; k < B; k++) {
        if (k < nb) {
#pragma unroll
          for (int j = 0; j < k; j++) x[k] -= x[j] * D[k * nb + j];
          x[k] *= 1.0f / D[k * nb + k];
        }
      }
#pragma unroll
      for (int k = 0; k < B; k++)
        if (k < nb) {
          A[i * ld + p + k] = x[k];
          P[(i - p - nb) * B + k] = x[k];
        }
    }
    sync();
    for (int c = p + nb + t.lane; c < n; c += TEAM) {
      float pc[B];
#pragma unroll
      for (int k = 0; k < B; k++) pc[k] = k < nb ? P[(c - p - nb) * B + k] : 0.0f;
      int i = c;
      for (; i + 4 <= n; i += 4) {
        float a[4];
#pragma unroll
        for (int u = 0; u < 4; u++) a[u] = A[(i + u) * ld + c];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const float* pi = P + (i + u - p - nb) * B;
#pragma unroll
          for (int k = 0; k < B; k++)
            if (k < nb) a[u] -= pi[k] * pc[k];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) A[(i + u) * ld + c] = a[u];
      }
      for (; i < n; i++) {
        float a = A[i * ld + c];
        const float* pi = P + (i - p - nb) * B;
#pragma unroll
        for (int k = 0; k < B; k++)
          if (k < nb) a -= pi[k] * pc[k];
        A[i * ld + c] = a;
      }
    }
    sync();
  }
  for (int e = t.lane; e < n * n; e += TEAM) {
    const int r = e / n, c = e % n;
    if (c > r) A[r * ld + c] = 0.0f;
  }
  sync();
}

// cho_solve_lds on one 64-lane team: the rows in blocks of 64, one per
// lane.  A block first takes the terms of the entries already solved (lane
// i sums over its row of L, in the same order as the column sweep below
// would), then solves its diagonal block column by column, each solved
// entry passed to the other lanes with readlane instead of through memory
// and a barrier: one barrier per block instead of two per entry.  Every
// entry gets the same operations in the same order as the column sweep.
MJH_DEV void cho_solve_wave(const float* L, int n, int ld, float* x, int lane) {
  // forward: y = inv(L) x
  for (int b0 = 0; b0 < n; b0 += 64) {
    const int i = b0 + lane, nb = n - b0 < 64 ? n - b0 : 64;
    const bool own = i < n;
    float y = own ? x[i] : 0.0f;
    if (own)
      for (int j = 0; j < b0; j++) y -= L[i * ld + j] * x[j];
    for (int k = 0; k < nb; k++) {
      const float yk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(y / L[(b0 + k) * ld + b0 + k]), k));
      if (lane == k) y = yk;
      else if (lane > k && own) y -= L[i * ld + b0 + k] * yk;
    }
    sync();
    if (own) x[i] = y;
    sync();
  }
  // backward: x = inv(L') y
  for (int b0 = ((n - 1) >> 6) << 6; b0 >= 0; b0 -= 64) {
    const int i = b0 + lane, nb = n - b0 < 64 ? n - b0 : 64;
    const bool own = i < n;
    float v = own ? x[i] : 0.0f;
    if (own)
      for (int j = n - 1; j >= b0 + 64; j--) v -= L[j * ld + i] * x[j];
    for (int k = nb - 1; k >= 0; k--) {
      const float xk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v / L[(b0 + k) * ld + b0 + k]), k));
      if (lane == k) v = xk;
      else if (lane < k && own) v -= L[(b0 + k) * ld + i] * xk;
    }
    sync();
    if (own) x[i] = v;
    sync();
  }
}

// x = inv(L L') x for an n-vector x in LDS, L lower (row stride ld) in LDS.
template <int TEAM>
MJH_DEV void cho_solve_lds(const float* L, int n, int ld, float* x, const Team<TEAM>& t) {
  if constexpr (TEAM == 64) {
    cho_solve_wave(L, n, ld, x, t.lane);
    return;
  }
  for (int k = 0; k < n; k++) {
    const float xk = x[k] / L[k * ld + k];
    sync();
    for (int i = k + 1 + t.lane; i < n; i += TEAM) x[i] -= L[i * ld + k] * xk;
    if (t.lane == 0) x[k] = xk;
    sync();
  }
  for (int k = n - 1; k >= 0; k--) {
    const float xk = x[k] / L[k * ld + k];
    sync();
    for (int i = t.lane; i < k; i += TEAM) x[i] -= L[k * ld + i] * xk;
    if (t.lane == 0) x[k] = xk;
    sync();
  }
}

// Factor the tree blocks of M (global, [nv, nv]) into L (global, may alias
// a different array) through an LDS staging buffer of nv*nv floats.
template <int TEAM>
MJH_DEV void factor_blocks(const mjhModel& m, const float* M, float* L, float* lds, const Team<TEAM>& t) {
  if constexpr (TEAM == kRegTeam) {
    reg_factor_blocks(m, M, L, lds, t);
    return;
  }
  const int nv = m.nv;
  for_each_tree_block(m, [&](int dof, int n) {
    for (int p = t.lane; p < n * n; p += TEAM) lds[p] = M[(dof + p / n) * nv + dof + p % n];
    sync();
    chol_lds(lds, n, n, m, t);
    for (int p = t.lane; p < n * n; p += TEAM) L[(dof + p / n) * nv + dof + p % n] = lds[p];
    sync();
  });
}

// x = inv(M) y with M = L L' given by its tree blocks (smooth.py:654-681).
// x may alias y.  lds needs nv*nv + nv floats.
template <int TEAM>
MJH_DEV void solve_blocks(const mjhModel& m, const float* L, float* x, const float* y, float* lds,
                          const Team<TEAM>& t) {
  if constexpr (TEAM == kRegTeam) {
    reg_solve_blocks(m, L, x, y, t);
    return;
  }
  const int nv = m.nv;
  for_each_tree_block(m, [&](int dof, int n) {
    float* Ls = lds;
    float* xs = lds + n * n;
    for (int p = t.lane; p < n * n; p += TEAM) Ls[p] = L[(dof + p / n) * nv + dof + p % n];
    for (int p = t.lane; p < n; p += TEAM) xs[p] = y[dof + p];
    sync();
    cho_solve_lds(Ls, n, n, xs, t);
    for (int p = t.lane; p < n; p += TEAM) x[dof + p] = xs[p];
    sync();
  });
}

// smooth.py:383-395 factor_m: the dense Cholesky factor qLDd of qMd, and
// for sparse models MuJoCo's L'DL factor of qM into qLD / qLDiagInv.
// lds: nv * nv floats (at least nM).
template <int TEAM>
MJH_DEV void factor_m(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  const int nv = m.nv, w = t.world;
  if (t.valid) factor_blocks(m, WPTR(d.qMd, w, nv * nv), WPTR(d.qLDd, w, nv * nv), lds, t);
  if (m.opt.is_sparse) {
    sync();
    factor_sparse(m, WPTR(d.qM, w, m.nM), WPTR(d.qLD, w, m.nM), WPTR(d.qLDiagInv, w, nv), lds, t);
  }
}

// ---------------------------------------------------------------------------
// smooth.py:398-466 rne
// lds: lds_velocity_floats; holds the body accelerations and forces
template <int TEAM>
MJH_DEV void rne(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  const int w = t.world;
  const int nbody = m.nbody, nv = m.nv;
  float* caccw = WPTR(d.rne_cacc, w, nbody * 6);
  float* cfrcw = WPTR(d.rne_cfrc, w, nbody * 6);
  float* cacc = lds;
  float* cfrc = lds + 6 * nbody;
  const float* cdof_dot = WPTR(d.cdof_dot, w, nv * 6);
  const float* qvel = WPTR(d.qvel, w, nv);
  if (t.valid && t.lane < 6) {
    const bool nograv = m.opt.disableflags & mjhDSBL_GRAVITY;
    // constant indices only: a lane-indexed read of the model argument's
    // array would make the compiler copy the argument to scratch memory
    const float gl = t.lane == 3 ? m.opt.gravity[0] : (t.lane == 4 ? m.opt.gravity[1] : m.opt.gravity[2]);
    const float g = (t.lane < 3 || nograv) ? 0.0f : -gl;
    cacc[t.lane] = g;
    caccw[t.lane] = g;
  }
  sync();
  for (int l = 0; l < m.nlevel; l++) {
    const int beg = m.body_treeadr[l], end = level_end(m, l);
    for (int k = beg + t.lane; k < end && t.valid; k += TEAM) {
      const int b = m.body_tree[k];
      const int pid = m.body_parentid[b];
      const int da = m.body_dofadr[b];
      S6 a = ld6(cacc + 6 * pid);
      for (int i = 0; i < m.body_dofnum[b]; i++) {
        const float qv = qvel[da + i];
#pragma unroll
        for (int c = 0; c < 6; c++) a.v[c] += cdof_dot[6 * (da + i) + c] * qv;
      }
      st6(cacc + 6 * b, a);
      st6(caccw + 6 * b, a);
    }
    sync();
  }
  if (t.valid) {
    const float* cinert = WPTR(d.cinert, w, nbody * 10);
    const float* cvel = WPTR(d.cvel, w, nbody * 6);
    for (int b = t.lane; b < nbody; b += TEAM) {
      const float* ci = cinert + 10 * b;
      S6 f = inert_vec(ci, ld6(cacc + 6 * b));
      const S6 v = ld6(cvel + 6 * b);
      const S6 g = motion_cross_force(v, inert_vec(ci, v));
#pragma unroll
      for (int c = 0; c < 6; c++) f.v[c] += g.v[c];
      st6(cfrc + 6 * b, f);
    }
  }
  sync();
  if (t.valid && t.lane < 6) {
    const int k = t.lane;
    for (int b = nbody - 1; b >= 0; b--) cfrc[6 * m.body_parentid[b] + k] += cfrc[6 * b + k];
  }
  sync();
  if (t.valid) {
    for (int i = t.lane; i < nbody * 6; i += TEAM) cfrcw[i] = cfrc[i];
    const float* cdof = WPTR(d.cdof, w, nv * 6);
    float* bias = WPTR(d.qfrc_bias, w, nv);
    for (int i = t.lane; i < nv; i += TEAM) bias[i] = dot6(ld6(cdof + 6 * i), ld6(cfrc + 6 * m.dof_bodyid[i]));
  }
  sync();
}

// ---------------------------------------------------------------------------
// smooth.py:469-536 transmission
template <int TEAM>
MJH_DEV void transmission(const mjhModel& m, const mjhData& d, const Team<TEAM>& t) {
  if (m.nu == 0) return;
  const int w = t.world;
  if (t.valid) {
    const float* qpos = WPTR(d.qpos, w, m.nq);
    float* length = WPTR(d.actuator_length, w, m.nu);
    float* moment = WPTR(d.actuator_moment, w, m.nu * m.nv);
    for (int a = t.lane; a < m.nu; a += TEAM) {
      const int j = m.actuator_trnid[2 * a];
      const int jt = m.jnt_type[j];
      const int qadr = m.jnt_qposadr[j];
      const int vadr = m.jnt_dofadr[j];
      const int trn = m.actuator_trntype[a];
      const float* gear = m.actuator_gear + 6 * a;
      float* mrow = moment + a * m.nv;
      if (trn != mjhTRN_JOINT && trn != mjhTRN_JOINTINPARENT) continue;
      if (jt == mjhJNT_FREE) {
        length[a] = 0.0f;
        if (trn == mjhTRN_JOINTINPARENT) {
          const Q4 qneg = quat_inv(ld4(qpos + qadr + 3));
          const V3 ga = rot_vec_quat(V3{gear[3], gear[4], gear[5]}, qneg);
          mrow[vadr + 0] = gear[0];
          mrow[vadr + 1] = gear[1];
          mrow[vadr + 2] = gear[2];
          mrow[vadr + 3] = ga.x;
          mrow[vadr + 4] = ga.y;
          mrow[vadr + 5] = ga.z;
        } else {
          for (int i = 0; i < 6; i++) mrow[vadr + i] = gear[i];
        }
      } else if (jt == mjhJNT_BALL) {
        const Q4 q = ld4(qpos + qadr);
        const V3 aa = quat_to_vel(q);
        V3 ga = V3{gear[0], gear[1], gear[2]};
        if (trn == mjhTRN_JOINTINPARENT) ga = rot_vec_quat(ga, quat_inv(q));
        length[a] = dot(aa, ga);
        mrow[vadr + 0] = ga.x;
        mrow[vadr + 1] = ga.y;
        mrow[vadr + 2] = ga.z;
      } else {
        length[a] = qpos[qadr] * gear[0];
        mrow[vadr] = gear[0];
      }
    }
  }
  sync();
}

// ---------------------------------------------------------------------------
// smooth.py:539-606 com_vel
// lds: 6 * nbody floats; holds the body velocities
template <int TEAM>
MJH_DEV void com_vel(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  const int w = t.world;
  const int nbody = m.nbody, nv = m.nv;
  float* cvelw = WPTR(d.cvel, w, nbody * 6);
  float* cvel = lds;
  float* cdof_dot = WPTR(d.cdof_dot, w, nv * 6);
  const float* cdof = WPTR(d.cdof, w, nv * 6);
  const float* qvel = WPTR(d.qvel, w, nv);
  if (t.valid && t.lane < 6) {
    cvel[t.lane] = 0.0f;
    cvelw[t.lane] = 0.0f;
  }
  sync();
  for (int l = 1; l < m.nlevel; l++) {
    const int beg = m.body_treeadr[l], end = level_end(m, l);
    for (int k = beg + t.lane; k < end && t.valid; k += TEAM) {
      const int b = m.body_tree[k];
      const int pid = m.body_parentid[b];
      int da = m.body_dofadr[b];
      const int ja = m.body_jntadr[b], jn = m.body_jntnum[b];
      S6 v = ld6(cvel + 6 * pid);
      for (int j = ja; j < ja + jn; j++) {
        const int jt = m.jnt_type[j];
        if (jt == mjhJNT_FREE) {
          for (int i = 0; i < 3; i++) {
            const S6 c = ld6(cdof + 6 * (da + i));
#pragma unroll
            for (int q = 0; q < 6; q++) v.v[q] += c.v[q] * qvel[da + i];
          }
          for (int i = 3; i < 6; i++) st6(cdof_dot + 6 * (da + i), motion_cross(v, ld6(cdof + 6 * (da + i))));
          for (int i = 3; i < 6; i++) {
            const S6 c = ld6(cdof + 6 * (da + i));
#pragma unroll
            for (int q = 0; q < 6; q++) v.v[q] += c.v[q] * qvel[da + i];
          }
          da += 6;
        } else if (jt == mjhJNT_BALL) {
          for (int i = 0; i < 3; i++) st6(cdof_dot + 6 * (da + i), motion_cross(v, ld6(cdof + 6 * (da + i))));
          for (int i = 0; i < 3; i++) {
            const S6 c = ld6(cdof + 6 * (da + i));
#pragma unroll
            for (int q = 0; q < 6; q++) v.v[q] += c.v[q] * qvel[da + i];
          }
          da += 3;
        } else {
          const S6 c = ld6(cdof + 6 * da);
          st6(cdof_dot + 6 * da, motion_cross(v, c));
#pragma unroll
          for (int q = 0; q < 6; q++) v.v[q] += c.v[q] * qvel[da];
          da += 1;
        }
      }
      st6(cvel + 6 * b, v);
      st6(cvelw + 6 * b, v);
    }
    sync();
  }
}

// ---------------------------------------------------------------------------
// passive.py:27-108 passive
template <int TEAM>
MJH_DEV void passive(const mjhModel& m, const mjhData& d, const Team<TEAM>& t) {
  const int w = t.world, nv = m.nv;
  float* qfrc_passive = WPTR(d.qfrc_passive, w, nv);
  if (m.opt.disableflags & mjhDSBL_PASSIVE) {
    if (t.valid)
      for (int i = t.lane; i < nv; i += TEAM) qfrc_passive[i] = 0.0f;
    sync();
    return;
  }
  float* spring = WPTR(d.qfrc_spring, w, nv);
  float* damper = WPTR(d.qfrc_damper, w, nv);
  const float* qpos = WPTR(d.qpos, w, m.nq);
  const float* qvel = WPTR(d.qvel, w, nv);
  if (t.valid)
    for (int i = t.lane; i < nv; i += TEAM) spring[i] = 0.0f;
  sync();
  if (t.valid) {
    for (int j = t.lane; j < m.njnt; j += TEAM) {
      const float k = m.jnt_stiffness[j];
      if (k == 0.0f) continue;
      const int da = m.jnt_dofadr[j], qa = m.jnt_qposadr[j], jt = m.jnt_type[j];
      if (jt == mjhJNT_FREE) {
        for (int i = 0; i < 3; i++) spring[da + i] = -k * (qpos[qa + i] - m.qpos_spring[qa + i]);
        const V3 dif = quat_sub(ld4(qpos + qa + 3), ld4(m.qpos_spring + qa + 3));
        spring[da + 3] = -k * dif.x;
        spring[da + 4] = -k * dif.y;
        spring[da + 5] = -k * dif.z;
      } else if (jt == mjhJNT_BALL) {
        const V3 dif = quat_sub(ld4(qpos + qa), ld4(m.qpos_spring + qa));
        spring[da + 0] = -k * dif.x;
        spring[da + 1] = -k * dif.y;
        spring[da + 2] = -k * dif.z;
      } else {
        spring[da] = -k * (qpos[qa] - m.qpos_spring[qa]);
      }
    }
  }
  sync();
  if (t.valid)
    for (int i = t.lane; i < nv; i += TEAM) {
      const float f = -m.dof_damping[i] * qvel[i];
      damper[i] = f;
      qfrc_passive[i] = f + spring[i];
    }
  sync();
}

}  // namespace mjh
