// CPU restatement of mujoco_warp's physics step (the oracle).
//
// TEST INFRASTRUCTURE ONLY.  This file restates, on the CPU, the algorithm of
// the reference's step (mujoco_warp/_src/{smooth,passive,support,forward,
// collision_driver,collision_primitive,constraint,solver,math}.py) so the
// HIP implementation can be checked against it.  Only tests/, the smoke
// check and the benchmark's CPU-baseline leg load it; the product never
// does.  It is instantiated in float (parity with the FP32 kernels) and in
// double (the accuracy anchor).
//
// Structure: one function per reference stage, each processing one world
// sequentially (no lanes, no atomics inside a world).  Worlds write their
// contacts / rows as one contiguous block of the global pools, reserved in
// world order when run single-threaded.  Each function cites the reference
// lines it restates.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "../include/mjx_warp_amd.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

// ---- structs: the C ABI field lists with the real type as a parameter ----
template <class T>
struct Opt {
  T timestep, impratio, tolerance, ls_tolerance;
  T gravity[3];
  int integrator, cone, solver, iterations, ls_iterations, disableflags, is_sparse, ls_parallel;
};

#define O_R(name) T* name;
#define O_I(name) int32_t* name;
#define O_B(name) uint8_t* name;
#define O_FIELD(kind, name, dims) O_##kind(name)
#define O_SIZE(name) int32_t name;

template <class T>
struct Model {
  MJH_MODEL_SIZES(O_SIZE)
  Opt<T> opt;
  T meaninertia;
  MJH_MODEL_ARRAYS(O_FIELD)
};

template <class T>
struct Data {
  MJH_DATA_SIZES(O_SIZE)
  MJH_DATA_ARRAYS(O_FIELD)
};

// ---- small math (math.py) ----
template <class T>
struct V3 {
  T x, y, z;
};
template <class T> V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> T norm(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> V3<T> normalize(V3<T> a) {
  T n = norm(a);
  return n > T(0) ? a * (T(1) / n) : a;
}
template <class T> V3<T> L3(const T* p) { return {p[0], p[1], p[2]}; }
template <class T> void S3(T* p, V3<T> a) {
  p[0] = a.x;
  p[1] = a.y;
  p[2] = a.z;
}

template <class T>
struct Q {
  T w, x, y, z;
};
template <class T> Q<T> LQ(const T* p) { return {p[0], p[1], p[2], p[3]}; }
template <class T> void SQ(T* p, Q<T> q) {
  p[0] = q.w;
  p[1] = q.x;
  p[2] = q.y;
  p[3] = q.z;
}
// math.py:23-30
template <class T> Q<T> qmul(Q<T> u, Q<T> v) {
  return {u.w * v.w - u.x * v.x - u.y * v.y - u.z * v.z, u.w * v.x + u.x * v.w + u.y * v.z - u.z * v.y,
          u.w * v.y - u.x * v.z + u.y * v.w + u.z * v.x, u.w * v.z + u.x * v.y - u.y * v.x + u.z * v.w};
}
// math.py:33-38
template <class T> V3<T> qrot(V3<T> v, Q<T> q) {
  V3<T> u{q.x, q.y, q.z};
  V3<T> r = u * (T(2) * dot(u, v)) + v * (q.w * q.w - dot(u, u));
  return r + cross(u, v) * (T(2) * q.w);
}
// math.py:41-45
template <class T> Q<T> axis_angle(V3<T> a, T ang) {
  T s = std::sin(ang * T(0.5)), c = std::cos(ang * T(0.5));
  return {c, a.x * s, a.y * s, a.z * s};
}
// math.py:48-64, row-major 3x3
template <class T> void q2m(Q<T> q, T* r) {
  T a = q.w, b = q.x, c = q.y, d = q.z;
  r[0] = a * a + b * b - c * c - d * d;
  r[1] = T(2) * (b * c - a * d);
  r[2] = T(2) * (b * d + a * c);
  r[3] = T(2) * (b * c + a * d);
  r[4] = a * a - b * b + c * c - d * d;
  r[5] = T(2) * (c * d - a * b);
  r[6] = T(2) * (b * d - a * c);
  r[7] = T(2) * (c * d + a * b);
  r[8] = a * a - b * b - c * c + d * d;
}
template <class T> Q<T> qnorm(Q<T> q) {
  T n = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  if (n > T(0)) return {q.w / n, q.x / n, q.y / n, q.z / n};
  return q;
}
// math.py:115-128
template <class T> V3<T> quat_to_vel(Q<T> q) {
  V3<T> ax{q.x, q.y, q.z};
  T s = norm(ax);
  if (s == T(0)) return {T(0), T(0), T(0)};
  T speed = T(2) * std::atan2(s, q.w);
  const T pi = T(3.14159265358979323846);
  if (speed > pi) speed -= T(2) * pi;
  return ax * (speed / s);
}
// math.py:131-139
template <class T> V3<T> quat_sub(Q<T> qa, Q<T> qb) { return quat_to_vel(qmul(Q<T>{qb.w, -qb.x, -qb.y, -qb.z}, qa)); }
// math.py:142-152
template <class T> Q<T> quat_integrate(Q<T> q, V3<T> v, T dt) {
  T n = norm(v);
  V3<T> a = normalize(v);
  return qnorm(qmul(q, axis_angle(a, dt * n)));
}
// mju_mulInertVec (math.py:72-82)
template <class T> void inert_vec(const T* i, const T* v, T* r) {
  r[0] = i[0] * v[0] + i[3] * v[1] + i[4] * v[2] - i[8] * v[4] + i[7] * v[5];
  r[1] = i[3] * v[0] + i[1] * v[1] + i[5] * v[2] + i[8] * v[3] - i[6] * v[5];
  r[2] = i[4] * v[0] + i[5] * v[1] + i[2] * v[2] - i[7] * v[3] + i[6] * v[4];
  r[3] = i[8] * v[1] - i[7] * v[2] + i[9] * v[3];
  r[4] = i[6] * v[2] - i[8] * v[0] + i[9] * v[4];
  r[5] = i[7] * v[0] - i[6] * v[1] + i[9] * v[5];
}
// math.py:85-97
template <class T> void motion_cross(const T* u, const T* v, T* r) {
  V3<T> a = cross(L3(u), L3(v));
  V3<T> b = cross(L3(u + 3), L3(v)) + cross(L3(u), L3(v + 3));
  S3(r, a);
  S3(r + 3, b);
}
// math.py:100-112
template <class T> void motion_cross_force(const T* v, const T* f, T* r) {
  V3<T> a = cross(L3(v), L3(f)) + cross(L3(v + 3), L3(f + 3));
  V3<T> b = cross(L3(v), L3(f + 3));
  S3(r, a);
  S3(r + 3, b);
}
template <class T> T dot6(const T* a, const T* b) {
  T s = 0;
  for (int i = 0; i < 6; i++) s += a[i] * b[i];
  return s;
}
// math.py:155-180 (frame rows: normal, two tangents)
template <class T> void make_frame(V3<T> a, T* f) {
  a = normalize(a);
  V3<T> b = (T(-0.5) < a.y && a.y < T(0.5)) ? V3<T>{0, 1, 0} : V3<T>{0, 0, 1};
  b = b - a * dot(a, b);
  b = normalize(b);
  if (norm(a) == T(0)) b = {0, 0, 0};
  V3<T> c = cross(a, b);
  S3(f, a);
  S3(f + 3, b);
  S3(f + 6, c);
}

template <class T> T clampv(T x, T lo, T hi) { return std::min(std::max(x, lo), hi); }

// ===========================================================================
// smooth.py:32-145 kinematics
template <class T>
void kinematics(const Model<T>& m, const Data<T>& d, int w) {
  const int nb = m.nbody;
  T* xpos = d.xpos + (size_t)w * nb * 3;
  T* xquat = d.xquat + (size_t)w * nb * 4;
  T* xmat = d.xmat + (size_t)w * nb * 9;
  T* xipos = d.xipos + (size_t)w * nb * 3;
  T* ximat = d.ximat + (size_t)w * nb * 9;
  T* xanchor = d.xanchor + (size_t)w * m.njnt * 3;
  T* xaxis = d.xaxis + (size_t)w * m.njnt * 3;
  const T* qpos = d.qpos + (size_t)w * m.nq;
  S3(xpos, V3<T>{0, 0, 0});
  SQ(xquat, Q<T>{1, 0, 0, 0});
  S3(xipos, V3<T>{0, 0, 0});
  for (int k = 0; k < 9; k++) xmat[k] = ximat[k] = (k % 4 == 0) ? T(1) : T(0);
  // bodies are numbered parents-first, so index order is a valid tree order
  for (int b = 1; b < nb; b++) {
    int ja = m.body_jntadr[b];
    const int jn = m.body_jntnum[b];
    V3<T> pos;
    Q<T> quat;
    if (jn == 1 && m.jnt_type[ja] == mjhJNT_FREE) {
      const int qa = m.jnt_qposadr[ja];
      pos = L3(qpos + qa);
      quat = LQ(qpos + qa + 3);
      S3(xanchor + 3 * ja, pos);
      S3(xaxis + 3 * ja, L3(m.jnt_axis + 3 * ja));
    } else {
      const int p = m.body_parentid[b];
      const T* pm = xmat + 9 * p;
      V3<T> bp = L3(m.body_pos + 3 * b);
      pos = V3<T>{pm[0] * bp.x + pm[1] * bp.y + pm[2] * bp.z, pm[3] * bp.x + pm[4] * bp.y + pm[5] * bp.z,
                  pm[6] * bp.x + pm[7] * bp.y + pm[8] * bp.z} + L3(xpos + 3 * p);
      quat = qmul(LQ(xquat + 4 * p), LQ(m.body_quat + 4 * b));
      for (int j = ja; j < ja + jn; j++) {
        const int qa = m.jnt_qposadr[j];
        const V3<T> jp = L3(m.jnt_pos + 3 * j), jx = L3(m.jnt_axis + 3 * j);
        const V3<T> anchor = qrot(jp, quat) + pos;
        const V3<T> axis = qrot(jx, quat);
        switch (m.jnt_type[j]) {
          case mjhJNT_BALL:
            quat = qmul(quat, LQ(qpos + qa));
            pos = anchor - qrot(jp, quat);
            break;
          case mjhJNT_SLIDE:
            pos = pos + axis * (qpos[qa] - m.qpos0[qa]);
            break;
          case mjhJNT_HINGE:
            quat = qmul(quat, axis_angle(jx, qpos[qa] - m.qpos0[qa]));
            pos = anchor - qrot(jp, quat);
            break;
        }
        S3(xanchor + 3 * j, anchor);
        S3(xaxis + 3 * j, axis);
      }
    }
    S3(xpos + 3 * b, pos);
    quat = qnorm(quat);
    SQ(xquat + 4 * b, quat);
    q2m(quat, xmat + 9 * b);
    S3(xipos + 3 * b, pos + qrot(L3(m.body_ipos + 3 * b), quat));
    q2m(qmul(quat, LQ(m.body_iquat + 4 * b)), ximat + 9 * b);
  }
  for (int g = 0; g < m.ngeom; g++) {
    const int b = m.geom_bodyid[g];
    const Q<T> q = LQ(xquat + 4 * b);
    S3(d.geom_xpos + ((size_t)w * m.ngeom + g) * 3, L3(xpos + 3 * b) + qrot(L3(m.geom_pos + 3 * g), q));
    q2m(qmul(q, LQ(m.geom_quat + 4 * g)), d.geom_xmat + ((size_t)w * m.ngeom + g) * 9);
  }
  for (int s = 0; s < m.nsite; s++) {
    const int b = m.site_bodyid[s];
    const Q<T> q = LQ(xquat + 4 * b);
    S3(d.site_xpos + ((size_t)w * m.nsite + s) * 3, L3(xpos + 3 * b) + qrot(L3(m.site_pos + 3 * s), q));
    q2m(qmul(q, LQ(m.site_quat + 4 * s)), d.site_xmat + ((size_t)w * m.nsite + s) * 9);
  }
}

// smooth.py:148-245 com_pos
template <class T>
void com_pos(const Model<T>& m, const Data<T>& d, int w) {
  const int nb = m.nbody, nv = m.nv;
  T* sc = d.subtree_com + (size_t)w * nb * 3;
  const T* xipos = d.xipos + (size_t)w * nb * 3;
  for (int b = 0; b < nb; b++) S3(sc + 3 * b, L3(xipos + 3 * b) * m.body_mass[b]);
  // children before parents; body 0 is its own parent and adds itself
  for (int b = nb - 1; b >= 0; b--) {
    const int p = m.body_parentid[b];
    for (int k = 0; k < 3; k++) sc[3 * p + k] += sc[3 * b + k];
  }
  for (int b = 0; b < nb; b++)
    for (int k = 0; k < 3; k++) sc[3 * b + k] /= m.subtree_mass[b];
  const T* ximat = d.ximat + (size_t)w * nb * 9;
  for (int b = 0; b < nb; b++) {
    const T* R = ximat + 9 * b;
    const T* I = m.body_inertia + 3 * b;
    const T mass = m.body_mass[b];
    const V3<T> dif = L3(xipos + 3 * b) - L3(sc + 3 * m.body_rootid[b]);
    T tmp[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        T s = 0;
        for (int k = 0; k < 3; k++) s += R[3 * i + k] * I[k] * R[3 * j + k];
        tmp[3 * i + j] = s;
      }
    T* r = d.cinert + ((size_t)w * nb + b) * 10;
    r[0] = tmp[0] + mass * (dif.y * dif.y + dif.z * dif.z);
    r[1] = tmp[4] + mass * (dif.x * dif.x + dif.z * dif.z);
    r[2] = tmp[8] + mass * (dif.x * dif.x + dif.y * dif.y);
    r[3] = tmp[1] - mass * dif.x * dif.y;
    r[4] = tmp[2] - mass * dif.x * dif.z;
    r[5] = tmp[5] - mass * dif.y * dif.z;
    r[6] = mass * dif.x;
    r[7] = mass * dif.y;
    r[8] = mass * dif.z;
    r[9] = mass;
  }
  T* cdof = d.cdof + (size_t)w * nv * 6;
  for (int j = 0; j < m.njnt; j++) {
    const int b = m.jnt_bodyid[j];
    const int da = m.jnt_dofadr[j];
    const V3<T> axis = L3(d.xaxis + ((size_t)w * m.njnt + j) * 3);
    const T* X = d.xmat + ((size_t)w * nb + b) * 9;
    const V3<T> off = L3(sc + 3 * m.body_rootid[b]) - L3(d.xanchor + ((size_t)w * m.njnt + j) * 3);
    auto set = [&](int i, V3<T> ang, V3<T> lin) {
      S3(cdof + 6 * i, ang);
      S3(cdof + 6 * i + 3, lin);
    };
    const int jt = m.jnt_type[j];
    if (jt == mjhJNT_FREE || jt == mjhJNT_BALL) {
      int r0 = da;
      if (jt == mjhJNT_FREE) {
        set(da + 0, {0, 0, 0}, {1, 0, 0});
        set(da + 1, {0, 0, 0}, {0, 1, 0});
        set(da + 2, {0, 0, 0}, {0, 0, 1});
        r0 = da + 3;
      }
      for (int i = 0; i < 3; i++) {
        V3<T> c{X[i], X[3 + i], X[6 + i]};  // column i of xmat
        set(r0 + i, c, cross(c, off));
      }
    } else if (jt == mjhJNT_SLIDE) {
      set(da, {0, 0, 0}, axis);
    } else {
      set(da, axis, cross(axis, off));
    }
  }
}

// MuJoCo's sparse layout of a dense nv x nv matrix (smooth.py:281-294
// qM_sparse): row i holds M(i, j) for j = i, parent(i), parent(parent(i)),
// ... from dof_Madr[i]
template <class T>
void dense_to_sparse(const Model<T>& m, const T* M, T* S) {
  for (int i = 0; i < m.nv; i++) {
    int a = m.dof_Madr[i];
    for (int j = i; j >= 0; j = m.dof_parentid[j]) S[a++] = M[i * m.nv + j];
  }
}

// smooth.py:317-350 _factor_i_sparse: L'DL of the sparse matrix S into L
// (unit L below the diagonal, D on it; leaves first, as MuJoCo), and
// Dinv = 1 / D
template <class T>
void factor_sparse(const Model<T>& m, const T* S, T* L, T* Dinv) {
  std::memcpy(L, S, sizeof(T) * m.nM);
  for (int k = m.nv - 1; k >= 0; k--) {
    const int ak = m.dof_Madr[k];
    int p = 1;
    for (int i = m.dof_parentid[k]; i >= 0; i = m.dof_parentid[i], p++) {
      const T tmp = L[ak + p] / L[ak];
      const int ai = m.dof_Madr[i];
      const int len = m.dof_Madr[i + 1] - ai;  // i < k <= nv - 1
      for (int q = 0; q < len; q++) L[ai + q] -= L[ak + p + q] * tmp;
      L[ak + p] = tmp;
    }
  }
  for (int k = 0; k < m.nv; k++) Dinv[k] = T(1) / L[m.dof_Madr[k]];
}

// smooth.py:248-314 crb: dense qMd, and for sparse models qM in MuJoCo's
// sparse layout
template <class T>
void crb(const Model<T>& m, const Data<T>& d, int w) {
  const int nb = m.nbody, nv = m.nv;
  T* c = d.crb + (size_t)w * nb * 10;
  std::memcpy(c, d.cinert + (size_t)w * nb * 10, sizeof(T) * nb * 10);
  for (int b = nb - 1; b >= 1; b--) {
    const int p = m.body_parentid[b];
    if (p == 0) continue;
    for (int k = 0; k < 10; k++) c[10 * p + k] += c[10 * b + k];
  }
  T* M = d.qMd + (size_t)w * nv * nv;
  std::fill(M, M + nv * nv, T(0));
  const T* cdof = d.cdof + (size_t)w * nv * 6;
  for (int i = 0; i < nv; i++) {
    T buf[6];
    inert_vec(c + 10 * m.dof_bodyid[i], cdof + 6 * i, buf);
    M[i * nv + i] = m.dof_armature[i] + dot6(cdof + 6 * i, buf);
    for (int j = m.dof_parentid[i]; j >= 0; j = m.dof_parentid[j]) {
      const T v = dot6(cdof + 6 * j, buf);
      M[i * nv + j] += v;
      M[j * nv + i] += v;
    }
  }
  if (m.opt.is_sparse) dense_to_sparse(m, M, d.qM + (size_t)w * m.nM);
}

// dense Cholesky A = L L' of the n x n block at (off, off) of an ld-strided
// matrix, L written to the same block of out with zeros above the diagonal
template <class T>
void cholesky_block(const T* A, T* out, int off, int n, int ld) {
  std::vector<T> L(n * n, T(0));
  for (int j = 0; j < n; j++) {
    T s = A[(off + j) * ld + off + j];
    for (int k = 0; k < j; k++) s -= L[j * n + k] * L[j * n + k];
    const T djj = std::sqrt(s);
    L[j * n + j] = djj;
    for (int i = j + 1; i < n; i++) {
      T t = A[(off + i) * ld + off + j];
      for (int k = 0; k < j; k++) t -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = t / djj;
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) out[(off + i) * ld + off + j] = L[i * n + j];
}
// x = inv(L L') y on the block at off (x may alias y)
template <class T>
void cho_solve_block(const T* L, int off, int n, int ld, T* x, const T* y) {
  std::vector<T> z(n);
  for (int i = 0; i < n; i++) {
    T s = y[off + i];
    for (int k = 0; k < i; k++) s -= L[(off + i) * ld + off + k] * z[k];
    z[i] = s / L[(off + i) * ld + off + i];
  }
  for (int i = n - 1; i >= 0; i--) {
    T s = z[i];
    for (int k = i + 1; k < n; k++) s -= L[(off + k) * ld + off + i] * z[k];
    z[i] = s / L[(off + i) * ld + off + i];
  }
  for (int i = 0; i < n; i++) x[off + i] = z[i];
}

// per-tree dense blocks (io.py:150-162 qLD_tile*)
template <class T, class F>
void each_block(const Model<T>& m, F f) {
  for (int g = 0; g < m.nqLD_group; g++) {
    const int beg = m.qLD_tileadr[g];
    const int end = (g == m.nqLD_group - 1) ? m.nqLD_tile : m.qLD_tileadr[g + 1];
    for (int k = beg; k < end; k++) f(m.qLD_tile[k], m.qLD_tilesize[g]);
  }
}

// smooth.py:353-379
template <class T>
void factor_m(const Model<T>& m, const Data<T>& d, int w) {
  const int nv = m.nv;
  const T* M = d.qMd + (size_t)w * nv * nv;
  T* L = d.qLDd + (size_t)w * nv * nv;
  each_block(m, [&](int off, int n) { cholesky_block(M, L, off, n, nv); });
  if (m.opt.is_sparse)
    factor_sparse(m, d.qM + (size_t)w * m.nM, d.qLD + (size_t)w * m.nM, d.qLDiagInv + (size_t)w * nv);
}

// smooth.py:654-681
template <class T>
void solve_m(const Model<T>& m, const Data<T>& d, int w, T* x, const T* y) {
  const int nv = m.nv;
  const T* L = d.qLDd + (size_t)w * nv * nv;
  each_block(m, [&](int off, int n) { cho_solve_block(L, off, n, nv, x + (size_t)w * nv, y + (size_t)w * nv); });
}

// smooth.py:398-466 rne
template <class T>
void rne(const Model<T>& m, const Data<T>& d, int w) {
  const int nb = m.nbody, nv = m.nv;
  T* cacc = d.rne_cacc + (size_t)w * nb * 6;
  T* cfrc = d.rne_cfrc + (size_t)w * nb * 6;
  const T* qvel = d.qvel + (size_t)w * nv;
  const T* cdd = d.cdof_dot + (size_t)w * nv * 6;
  for (int k = 0; k < 6; k++) cacc[k] = 0;
  if (!(m.opt.disableflags & mjhDSBL_GRAVITY))
    for (int k = 0; k < 3; k++) cacc[3 + k] = -m.opt.gravity[k];
  for (int b = 1; b < nb; b++) {
    const int p = m.body_parentid[b];
    for (int k = 0; k < 6; k++) cacc[6 * b + k] = cacc[6 * p + k];
    for (int i = m.body_dofadr[b]; i < m.body_dofadr[b] + m.body_dofnum[b]; i++)
      for (int k = 0; k < 6; k++) cacc[6 * b + k] += cdd[6 * i + k] * qvel[i];
  }
  const T* ci = d.cinert + (size_t)w * nb * 10;
  const T* cvel = d.cvel + (size_t)w * nb * 6;
  for (int b = 0; b < nb; b++) {
    T a[6], iv[6], g[6];
    inert_vec(ci + 10 * b, cacc + 6 * b, a);
    inert_vec(ci + 10 * b, cvel + 6 * b, iv);
    motion_cross_force(cvel + 6 * b, iv, g);
    for (int k = 0; k < 6; k++) cfrc[6 * b + k] = a[k] + g[k];
  }
  for (int b = nb - 1; b >= 0; b--) {
    const int p = m.body_parentid[b];
    for (int k = 0; k < 6; k++) cfrc[6 * p + k] += cfrc[6 * b + k];
  }
  const T* cdof = d.cdof + (size_t)w * nv * 6;
  for (int i = 0; i < nv; i++) d.qfrc_bias[(size_t)w * nv + i] = dot6(cdof + 6 * i, cfrc + 6 * m.dof_bodyid[i]);
}

// smooth.py:469-536 transmission
template <class T>
void transmission(const Model<T>& m, const Data<T>& d, int w) {
  const int nv = m.nv;
  const T* qpos = d.qpos + (size_t)w * m.nq;
  for (int a = 0; a < m.nu; a++) {
    const int j = m.actuator_trnid[2 * a];
    const int qa = m.jnt_qposadr[j], va = m.jnt_dofadr[j], jt = m.jnt_type[j];
    const int trn = m.actuator_trntype[a];
    const T* gear = m.actuator_gear + 6 * a;
    T* mom = d.actuator_moment + ((size_t)w * m.nu + a) * nv;
    T* len = d.actuator_length + (size_t)w * m.nu + a;
    if (trn != mjhTRN_JOINT && trn != mjhTRN_JOINTINPARENT) continue;
    if (jt == mjhJNT_FREE) {
      *len = 0;
      if (trn == mjhTRN_JOINTINPARENT) {
        Q<T> q = LQ(qpos + qa + 3);
        V3<T> ga = qrot(V3<T>{gear[3], gear[4], gear[5]}, Q<T>{q.w, -q.x, -q.y, -q.z});
        mom[va + 0] = gear[0];
        mom[va + 1] = gear[1];
        mom[va + 2] = gear[2];
        mom[va + 3] = ga.x;
        mom[va + 4] = ga.y;
        mom[va + 5] = ga.z;
      } else {
        for (int i = 0; i < 6; i++) mom[va + i] = gear[i];
      }
    } else if (jt == mjhJNT_BALL) {
      Q<T> q = LQ(qpos + qa);
      V3<T> ga{gear[0], gear[1], gear[2]};
      if (trn == mjhTRN_JOINTINPARENT) ga = qrot(ga, Q<T>{q.w, -q.x, -q.y, -q.z});
      *len = dot(quat_to_vel(q), ga);
      mom[va + 0] = ga.x;
      mom[va + 1] = ga.y;
      mom[va + 2] = ga.z;
    } else {
      *len = qpos[qa] * gear[0];
      mom[va] = gear[0];
    }
  }
}

// smooth.py:539-606 com_vel
template <class T>
void com_vel(const Model<T>& m, const Data<T>& d, int w) {
  const int nb = m.nbody, nv = m.nv;
  T* cvel = d.cvel + (size_t)w * nb * 6;
  T* cdd = d.cdof_dot + (size_t)w * nv * 6;
  const T* cdof = d.cdof + (size_t)w * nv * 6;
  const T* qvel = d.qvel + (size_t)w * nv;
  for (int k = 0; k < 6; k++) cvel[k] = 0;
  for (int b = 1; b < nb; b++) {
    const int p = m.body_parentid[b];
    T v[6];
    for (int k = 0; k < 6; k++) v[k] = cvel[6 * p + k];
    int da = m.body_dofadr[b];
    auto add = [&](int i) {
      for (int k = 0; k < 6; k++) v[k] += cdof[6 * i + k] * qvel[i];
    };
    for (int j = m.body_jntadr[b]; j < m.body_jntadr[b] + m.body_jntnum[b]; j++) {
      const int jt = m.jnt_type[j];
      if (jt == mjhJNT_FREE) {
        add(da);
        add(da + 1);
        add(da + 2);
        for (int i = 3; i < 6; i++) motion_cross(v, cdof + 6 * (da + i), cdd + 6 * (da + i));
        add(da + 3);
        add(da + 4);
        add(da + 5);
        da += 6;
      } else if (jt == mjhJNT_BALL) {
        for (int i = 0; i < 3; i++) motion_cross(v, cdof + 6 * (da + i), cdd + 6 * (da + i));
        add(da);
        add(da + 1);
        add(da + 2);
        da += 3;
      } else {
        motion_cross(v, cdof + 6 * da, cdd + 6 * da);
        add(da);
        da += 1;
      }
    }
    for (int k = 0; k < 6; k++) cvel[6 * b + k] = v[k];
  }
}

// passive.py:27-108
template <class T>
void passive(const Model<T>& m, const Data<T>& d, int w) {
  const int nv = m.nv;
  T* pas = d.qfrc_passive + (size_t)w * nv;
  if (m.opt.disableflags & mjhDSBL_PASSIVE) {
    for (int i = 0; i < nv; i++) pas[i] = 0;
    return;
  }
  T* spr = d.qfrc_spring + (size_t)w * nv;
  T* dmp = d.qfrc_damper + (size_t)w * nv;
  const T* qpos = d.qpos + (size_t)w * m.nq;
  const T* qvel = d.qvel + (size_t)w * nv;
  for (int i = 0; i < nv; i++) spr[i] = 0;
  for (int j = 0; j < m.njnt; j++) {
    const T k = m.jnt_stiffness[j];
    if (k == T(0)) continue;
    const int da = m.jnt_dofadr[j], qa = m.jnt_qposadr[j], jt = m.jnt_type[j];
    if (jt == mjhJNT_FREE) {
      for (int i = 0; i < 3; i++) spr[da + i] = -k * (qpos[qa + i] - m.qpos_spring[qa + i]);
      V3<T> dif = quat_sub(LQ(qpos + qa + 3), LQ(m.qpos_spring + qa + 3));
      spr[da + 3] = -k * dif.x;
      spr[da + 4] = -k * dif.y;
      spr[da + 5] = -k * dif.z;
    } else if (jt == mjhJNT_BALL) {
      V3<T> dif = quat_sub(LQ(qpos + qa), LQ(m.qpos_spring + qa));
      spr[da + 0] = -k * dif.x;
      spr[da + 1] = -k * dif.y;
      spr[da + 2] = -k * dif.z;
    } else {
      spr[da] = -k * (qpos[qa] - m.qpos_spring[qa]);
    }
  }
  for (int i = 0; i < nv; i++) {
    dmp[i] = -m.dof_damping[i] * qvel[i];
    pas[i] = dmp[i] + spr[i];
  }
}

// actuator tile groups (io.py:164-204), last-group end as passed
template <class T, class F>
void each_act_tile(const Model<T>& m, int last_end, F f) {
  for (int g = 0; g < m.nact_group; g++) {
    const int beg = m.actuator_moment_tileadr[g];
    const int end = (g == m.nact_group - 1) ? last_end : m.actuator_moment_tileadr[g + 1];
    for (int k = beg; k < end; k++)
      f(m.actuator_moment_offset_nv[k], m.actuator_moment_offset_nu[k], m.actuator_moment_tilesize_nv[g],
        m.actuator_moment_tilesize_nu[g]);
  }
}

// forward.py:423-500 fwd_velocity
template <class T>
void fwd_velocity(const Model<T>& m, const Data<T>& d, int w) {
  const int nv = m.nv, nu = m.nu;
  const T* mom = d.actuator_moment + (size_t)w * nu * nv;
  const T* qvel = d.qvel + (size_t)w * nv;
  T* vel = d.actuator_velocity + (size_t)w * nu;
  // sparse models: moment * qvel over all actuators and dofs (forward.py:427-438)
  if (m.opt.is_sparse) {
    for (int a = 0; a < nu; a++) {
      T s = 0;
      for (int j = 0; j < nv; j++) s += mom[a * nv + j] * qvel[j];
      vel[a] = s;
    }
  } else
  each_act_tile(m, m.nact_group, [&](int onv, int onu, int tnv, int tnu) {
    if (tnu == 0 || tnv == 0) return;
    for (int a = 0; a < tnu; a++) {
      T s = 0;
      for (int j = 0; j < tnv; j++) s += mom[(onu + a) * nv + onv + j] * qvel[onv + j];
      vel[onu + a] = s;
    }
  });
  com_vel(m, d, w);
  passive(m, d, w);
  rne(m, d, w);
}

// forward.py:503-641 fwd_actuation
template <class T>
void fwd_actuation(const Model<T>& m, const Data<T>& d, int w) {
  const int nv = m.nv, nu = m.nu;
  T* qfrc = d.qfrc_actuator + (size_t)w * nv;
  if (nu == 0 || (m.opt.disableflags & mjhDSBL_ACTUATION)) {
    for (int i = 0; i < m.na; i++) d.act_dot[(size_t)w * m.na + i] = 0;
    for (int i = 0; i < nv; i++) qfrc[i] = 0;
    return;
  }
  T* force = d.actuator_force + (size_t)w * nu;
  for (int a = 0; a < nu; a++) {
    const T len = d.actuator_length[(size_t)w * nu + a], vel = d.actuator_velocity[(size_t)w * nu + a];
    const T* gp = m.actuator_gainprm + 10 * a;
    const T* bp = m.actuator_biasprm + 10 * a;
    const T gain = gp[0] + gp[1] * len + gp[2] * vel;
    const T bias = bp[0] + bp[1] * len + bp[2] * vel;
    T c = d.ctrl[(size_t)w * nu + a];
    if (m.actuator_ctrllimited[a] && !(m.opt.disableflags & mjhDSBL_CLAMPCTRL))
      c = clampv(c, m.actuator_ctrlrange[2 * a], m.actuator_ctrlrange[2 * a + 1]);
    force[a] = gain * c + bias;  // forcerange is not applied (forward.py:539-541)
  }
  const T* mom = d.actuator_moment + (size_t)w * nu * nv;
  // sparse models: moment' force over all actuators and dofs (forward.py:571-578)
  if (m.opt.is_sparse) {
    for (int j = 0; j < nv; j++) {
      T s = 0;
      for (int a = 0; a < nu; a++) s += mom[a * nv + j] * force[a];
      qfrc[j] = s;
    }
  } else
  each_act_tile(m, m.nact_group, [&](int onv, int onu, int tnv, int tnu) {
    if (tnu == 0 || tnv == 0) return;
    for (int j = 0; j < tnv; j++) {
      T s = 0;
      for (int a = 0; a < tnu; a++) s += mom[(onu + a) * nv + onv + j] * force[onu + a];
      qfrc[onv + j] = s;
    }
  });
  for (int i = 0; i < nv; i++) {
    const int j = m.dof_jntid[i];
    if (m.jnt_actfrclimited[j]) qfrc[i] = clampv(qfrc[i], m.jnt_actfrcrange[2 * j], m.jnt_actfrcrange[2 * j + 1]);
  }
}

// support.py:140-172
template <class T>
void xfrc_accumulate(const Model<T>& m, const Data<T>& d, int w, T* qfrc_all) {
  const int nv = m.nv, nb = m.nbody;
  const T* cdof = d.cdof + (size_t)w * nv * 6;
  for (int i = 0; i < nv; i++) {
    const V3<T> rot = L3(cdof + 6 * i), lin = L3(cdof + 6 * i + 3);
    T acc = 0;
    for (int b = m.dof_bodyid[i]; b < nb; b++) {
      // is b in the subtree of dof i's body?
      int p = b;
      while (p != 0 && p != m.dof_bodyid[i]) p = m.body_parentid[p];
      if (p == 0) continue;
      const V3<T> off = L3(d.xipos + ((size_t)w * nb + b) * 3) - L3(d.subtree_com + ((size_t)w * nb + m.body_rootid[b]) * 3);
      const T* f = d.xfrc_applied + ((size_t)w * nb + b) * 6;
      acc += dot(lin, L3(f)) + dot(rot, L3(f + 3)) + dot(cross(rot, off), L3(f));
    }
    qfrc_all[(size_t)w * nv + i] += acc;
  }
}

// forward.py:644-661
template <class T>
void fwd_acceleration(const Model<T>& m, const Data<T>& d, int w) {
  const int nv = m.nv;
  const size_t o = (size_t)w * nv;
  for (int i = 0; i < nv; i++)
    d.qfrc_smooth[o + i] = d.qfrc_passive[o + i] - d.qfrc_bias[o + i] + d.qfrc_actuator[o + i] + d.qfrc_applied[o + i];
  xfrc_accumulate(m, d, w, d.qfrc_smooth);
  solve_m(m, d, w, d.qacc_smooth, d.qfrc_smooth);
}

// support.py:35-93 dense mul_m over tree blocks
template <class T>
void mul_m(const Model<T>& m, const T* M, T* res, const T* vec) {
  const int nv = m.nv;
  each_block(m, [&](int off, int n) {
    for (int i = 0; i < n; i++) {
      T s = 0;
      for (int j = 0; j < n; j++) s += M[(off + i) * nv + off + j] * vec[off + j];
      res[off + i] = s;
    }
  });
}

// ===========================================================================
// collision

// collision_driver.py:27-49
template <class T>
bool geom_filter(const Model<T>& m, int g1, int g2, bool filterparent) {
  const int b1 = m.geom_bodyid[g1], b2 = m.geom_bodyid[g2];
  const int w1 = m.body_weldid[b1], w2 = m.body_weldid[b2];
  const int wp1 = m.body_weldid[m.body_parentid[w1]], wp2 = m.body_weldid[m.body_parentid[w2]];
  const bool self = w1 == w2;
  const bool pc = filterparent && w1 != 0 && w2 != 0 && (w1 == wp2 || w2 == wp1);
  const bool mask = (m.geom_contype[g1] & m.geom_conaffinity[g2]) || (m.geom_contype[g2] & m.geom_conaffinity[g1]);
  return mask && !self && !pc;
}

struct Pair {
  int g1, g2;
};

// collision_driver.py:458-512 candidate pairs of one world in (g1, g2) order
template <class T>
void nxn_pairs(const Model<T>& m, const Data<T>& d, int w, std::vector<Pair>& out) {
  const int ng = m.ngeom;
  const bool fp = !(m.opt.disableflags & mjhDSBL_FILTERPARENT);
  const T* gx = d.geom_xpos + (size_t)w * ng * 3;
  const T* gm = d.geom_xmat + (size_t)w * ng * 9;
  for (int g1 = 0; g1 < ng; g1++)
    for (int g2 = g1 + 1; g2 < ng; g2++) {
      const T s1 = m.geom_rbound[g1], s2 = m.geom_rbound[g2];
      const T bound = s1 + s2 + std::max(m.geom_margin[g1], m.geom_margin[g2]);
      const V3<T> dif = L3(gx + 3 * g2) - L3(gx + 3 * g1);
      bool ok;
      if (s1 != T(0) && s2 != T(0)) {
        ok = dot(dif, dif) <= bound * bound;
      } else if (s1 == T(0)) {
        ok = dot(dif, V3<T>{gm[9 * g1 + 2], gm[9 * g1 + 5], gm[9 * g1 + 8]}) <= bound;
      } else {
        ok = dot(dif * T(-1), V3<T>{gm[9 * g2 + 2], gm[9 * g2 + 5], gm[9 * g2 + 8]}) <= bound;
      }
      if (ok && geom_filter(m, g1, g2, fp)) out.push_back({g1, g2});
    }
}

// collision_driver.py:101-199 sphere / plane bounding volumes of SAP
template <class T>
struct BV {
  T x, y, z, w;
};
template <class T>
BV<T> encode_plane(V3<T> n, V3<T> p, T margin) {
  n = normalize(n);
  const T off = -dot(n, p + n * margin);
  const T s = std::fabs(off);
  if (std::fabs(off) < T(1e-6)) return {n.x, n.y, n.z, T(-1)};
  if (off < T(0)) return {s * n.x, s * n.y, s * n.z, T(-2)};
  return {s * n.x, s * n.y, s * n.z, T(-3)};
}
template <class T>
bool sap_overlap(BV<T> a, BV<T> b) {
  if (a.w < T(0) && b.w < T(0)) return false;
  if (a.w < T(0) || b.w < T(0)) {
    if (b.w < T(0)) std::swap(a, b);
    const T mag = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w);
    const V3<T> n = normalize(V3<T>{a.x, a.y, a.z});
    const T off = a.w == T(-1) ? T(0) : (a.w == T(-2) ? -mag : mag);
    return dot(V3<T>{b.x, b.y, b.z}, n) + off <= b.w;
  }
  const V3<T> dl{a.x - b.x, a.y - b.y, a.z - b.z};
  const T rs = a.w + b.w;
  return dot(dl, dl) <= rs * rs;
}

// collision_driver.py:350-455 sweep and prune of one world
template <class T>
void sap_pairs(const Model<T>& m, const Data<T>& d, int w, std::vector<Pair>& out) {
  const int ng = m.ngeom;
  const bool fp = !(m.opt.disableflags & mjhDSBL_FILTERPARENT);
  const T* gx = d.geom_xpos + (size_t)w * ng * 3;
  const T* gm = d.geom_xmat + (size_t)w * ng * 9;
  const V3<T> dir = normalize(V3<T>{T(0.5935), T(0.7790), T(0.1235)});
  std::vector<T> lower(ng), upper(ng);
  std::vector<int> idx(ng);
  for (int i = 0; i < ng; i++) {
    T r = m.geom_rbound[i];
    if (r == T(0)) r = T(1000000000.0);
    const T rad = r + m.geom_margin[i];
    const T c = dot(dir, L3(gx + 3 * i));
    lower[i] = c - rad;
    upper[i] = c + rad;
    idx[i] = i;
  }
  // stable sort by lower bound (ties keep geom order)
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return lower[a] < lower[b]; });
  std::vector<T> sl(ng);
  std::vector<BV<T>> bv(ng);
  for (int i = 0; i < ng; i++) {
    const int g = idx[i];
    sl[i] = lower[g];
    if (m.geom_rbound[g] == T(0))
      bv[i] = encode_plane(V3<T>{gm[9 * g + 2], gm[9 * g + 5], gm[9 * g + 8]}, L3(gx + 3 * g), m.geom_margin[g]);
    else
      bv[i] = {gx[3 * g], gx[3 * g + 1], gx[3 * g + 2], m.geom_rbound[g] + m.geom_margin[g]};
  }
  for (int i = 0; i < ng; i++) {
    // first sorted position after i whose lower bound exceeds upper(i)
    int lo = i + 1, hi = ng;
    const T end = upper[idx[i]];
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (sl[mid] > end) hi = mid;
      else lo = mid + 1;
    }
    const int limit = std::min(ng - 1, lo);
    for (int j = i + 1; j <= limit; j++)
      if (geom_filter(m, idx[i], idx[j], fp) && sap_overlap(bv[i], bv[j])) out.push_back({idx[i], idx[j]});
  }
}

// narrowphase (collision_primitive.py)
template <class T>
struct Con {
  T dist;
  V3<T> pos;
  T frame[9];
  int g1, g2;
};

template <class T>
void emit(std::vector<Con<T>>& out, T dist, V3<T> pos, const T* frame, T margin, int g1, int g2) {
  if (!((dist - margin) < T(0))) return;
  Con<T> c;
  c.dist = dist;
  c.pos = pos;
  std::memcpy(c.frame, frame, sizeof(T) * 9);
  c.g1 = g1;
  c.g2 = g2;
  out.push_back(c);
}

// math.py:191-242
template <class T>
V3<T> closest_segment_point(V3<T> a, V3<T> b, V3<T> pt) {
  const V3<T> ab = b - a;
  const T t = dot(pt - a, ab) / (dot(ab, ab) + T(1e-6));
  return a + ab * clampv(t, T(0), T(1));
}
template <class T>
void closest_segment_to_segment_points(V3<T> a0, V3<T> a1, V3<T> b0, V3<T> b1, V3<T>* ra, V3<T>* rb) {
  V3<T> da = a1 - a0, db = b1 - b0;
  const T la = norm(da), lb = norm(db);
  if (la != T(0)) da = da * (T(1) / la);
  if (lb != T(0)) db = db * (T(1) / lb);
  const T ha = la * T(0.5), hb = lb * T(0.5);
  const V3<T> am = a0 + da * ha, bm = b0 + db * hb;
  const V3<T> tr = am - bm;
  const T dd = dot(da, db), dat = dot(da, tr), dbt = dot(db, tr);
  const T denom = T(1) - dd * dd;
  const T ota = (-dat + dd * dbt) / (denom + T(1e-6));
  const T otb = dbt + ota * dd;
  const V3<T> ba = am + da * clampv(ota, -ha, ha);
  const V3<T> bb = bm + db * clampv(otb, -hb, hb);
  const V3<T> na = closest_segment_point(a0, a1, bb);
  const V3<T> nb = closest_segment_point(b0, b1, ba);
  const T d1 = dot(bb - na, bb - na), d2 = dot(ba - nb, ba - nb);
  if (d1 < d2) {
    *ra = na;
    *rb = bb;
  } else {
    *ra = ba;
    *rb = nb;
  }
}

template <class T>
void sphere_sphere(std::vector<Con<T>>& out, V3<T> p1, T r1, V3<T> p2, T r2, T margin, int g1, int g2) {
  const V3<T> dir = p2 - p1;
  T dist = norm(dir);
  const V3<T> n = dist == T(0) ? V3<T>{1, 0, 0} : dir * (T(1) / dist);
  dist = dist - (r1 + r2);
  T f[9];
  make_frame(n, f);
  emit(out, dist, p1 + n * (r1 + T(0.5) * dist), f, margin, g1, g2);
}

template <class T>
void narrowphase_pair(const Model<T>& m, const Data<T>& d, int w, int g1, int g2, std::vector<Con<T>>& out) {
  const int ng = m.ngeom;
  const int t1 = m.geom_type[g1], t2 = m.geom_type[g2];
  const T margin = std::max(m.geom_margin[g1], m.geom_margin[g2]);
  const V3<T> p1 = L3(d.geom_xpos + ((size_t)w * ng + g1) * 3), p2 = L3(d.geom_xpos + ((size_t)w * ng + g2) * 3);
  const T* R1 = d.geom_xmat + ((size_t)w * ng + g1) * 9;
  const T* R2 = d.geom_xmat + ((size_t)w * ng + g2) * 9;
  const T* s1 = m.geom_size + 3 * g1;
  const T* s2 = m.geom_size + 3 * g2;
  auto col = [](const T* R, int c) { return V3<T>{R[c], R[3 + c], R[6 + c]}; };
  auto plane_sphere = [&](V3<T> n, V3<T> sp, T r, T* dist, V3<T>* pos) {
    *dist = dot(sp - p1, n) - r;
    *pos = sp - n * (r + T(0.5) * (*dist));
  };
  if (t1 == mjhGEOM_PLANE && t2 == mjhGEOM_SPHERE) {
    const V3<T> n = col(R1, 2);
    T dist, f[9];
    V3<T> pos;
    plane_sphere(n, p2, s2[0], &dist, &pos);
    make_frame(n, f);
    emit(out, dist, pos, f, margin, g1, g2);
  } else if (t1 == mjhGEOM_SPHERE && t2 == mjhGEOM_SPHERE) {
    sphere_sphere(out, p1, s1[0], p2, s2[0], margin, g1, g2);
  } else if (t1 == mjhGEOM_PLANE && t2 == mjhGEOM_CAPSULE) {
    const V3<T> n = col(R1, 2), axis = col(R2, 2);
    V3<T> b = axis - n * dot(n, axis);
    const T bn = norm(b);
    if (bn != T(0)) b = b * (T(1) / bn);
    if (bn < T(0.5)) b = (T(-0.5) < n.y && n.y < T(0.5)) ? V3<T>{0, 1, 0} : V3<T>{0, 0, 1};
    const V3<T> c = cross(n, b);
    T f[9] = {n.x, n.y, n.z, b.x, b.y, b.z, c.x, c.y, c.z};
    const V3<T> seg = axis * s2[1];
    T dist;
    V3<T> pos;
    plane_sphere(n, p2 + seg, s2[0], &dist, &pos);
    emit(out, dist, pos, f, margin, g1, g2);
    plane_sphere(n, p2 - seg, s2[0], &dist, &pos);
    emit(out, dist, pos, f, margin, g1, g2);
  } else if (t1 == mjhGEOM_PLANE && t2 == mjhGEOM_BOX) {
    const V3<T> n = col(R1, 2);
    const T dist = dot(p2 - p1, n);
    int count = 0;
    for (int i = 0; i < 8; i++) {
      const V3<T> cl{(i & 1) ? s2[0] : -s2[0], (i & 2) ? s2[1] : -s2[1], (i & 4) ? s2[2] : -s2[2]};
      const V3<T> corner{R2[0] * cl.x + R2[1] * cl.y + R2[2] * cl.z, R2[3] * cl.x + R2[4] * cl.y + R2[5] * cl.z,
                         R2[6] * cl.x + R2[7] * cl.y + R2[8] * cl.z};
      const T ld = dot(n, corner);
      if (dist + ld > margin || ld > T(0)) continue;
      const T cd = dist + ld;
      T f[9];
      make_frame(n, f);
      emit(out, cd, corner + p2 + (n * cd) * (T(1) / T(-2)), f, margin, g1, g2);
      if (++count >= 4) break;
    }
  } else if (t1 == mjhGEOM_CAPSULE && t2 == mjhGEOM_CAPSULE) {
    const V3<T> seg1 = col(R1, 2) * s1[1], seg2 = col(R2, 2) * s2[1];
    V3<T> a, b;
    closest_segment_to_segment_points(p1 - seg1, p1 + seg1, p2 - seg2, p2 + seg2, &a, &b);
    sphere_sphere(out, a, s1[0], b, s2[0], margin, g1, g2);
  }
}

// collision_driver.py:304-347 for one written contact
template <class T>
void contact_params(const Model<T>& m, const Data<T>& d, int i) {
  const int g1 = d.contact_geom[2 * i], g2 = d.contact_geom[2 * i + 1];
  const T margin = std::max(m.geom_margin[g1], m.geom_margin[g2]);
  const T gap = std::max(m.geom_gap[g1], m.geom_gap[g2]);
  const T s1 = m.geom_solmix[g1], s2 = m.geom_solmix[g2];
  const T mv = T(1e-15);
  T mix = s1 / (s1 + s2);
  if (s1 < mv && s2 < mv) mix = T(0.5);
  else if (s1 < mv) mix = T(0);
  else if (s2 < mv) mix = T(1);
  const int p1 = m.geom_priority[g1], p2 = m.geom_priority[g2];
  if (p1 != p2) mix = p1 > p2 ? T(1) : T(0);
  d.contact_dim[i] = p1 == p2 ? std::max(m.geom_condim[g1], m.geom_condim[g2]) : (p1 > p2 ? m.geom_condim[g1] : m.geom_condim[g2]);
  const T* r1 = m.geom_solref + 2 * g1;
  const T* r2 = m.geom_solref + 2 * g2;
  for (int k = 0; k < 2; k++)
    d.contact_solref[2 * i + k] =
        (r1[0] > T(0) && r2[0] > T(0)) ? mix * r1[k] + (T(1) - mix) * r2[k] : std::min(r1[k], r2[k]);
  d.contact_includemargin[i] = margin - gap;
  T f[3];
  for (int k = 0; k < 3; k++) f[k] = std::max(m.geom_friction[3 * g1 + k], m.geom_friction[3 * g2 + k]);
  T* fr = d.contact_friction + 5 * i;
  fr[0] = f[0];
  fr[1] = f[0];
  fr[2] = f[1];
  fr[3] = f[2];
  fr[4] = f[2];
  for (int k = 0; k < 5; k++) d.contact_solimp[5 * i + k] = mix * m.geom_solimp[5 * g1 + k] + (T(1) - mix) * m.geom_solimp[5 * g2 + k];
}

// reserve n entries of a global pool for one world; returns the base index
inline int reserve(int32_t* counter, int n) {
  if (n <= 0) return 0;
  return std::atomic_ref<int32_t>(*counter).fetch_add(n);
}

// broadphase of one world into the pair pool (mode 0 auto, 1 nxn, 2 sap)
template <class T>
void broadphase(const Model<T>& m, const Data<T>& d, int w, int mode) {
  std::vector<Pair> pairs;
  const bool sap = mode == 2 || (mode == 0 && m.ngeom > 100);
  if (sap) sap_pairs(m, d, w, pairs);
  else nxn_pairs(m, d, w, pairs);
  const int n = (int)pairs.size();
  const int base = reserve(d.ncollision, n);
  const int nw = n <= 0 ? 0 : std::max(0, std::min(n, d.nconmax - base));
  d.world_pair[2 * w] = n ? base : 0;
  d.world_pair[2 * w + 1] = nw;
  for (int k = 0; k < nw; k++) {
    int g1 = pairs[k].g1, g2 = pairs[k].g2;
    if (m.geom_type[g1] > m.geom_type[g2]) std::swap(g1, g2);  // collision_driver.py:62-65
    d.collision_pair[2 * (base + k)] = g1;
    d.collision_pair[2 * (base + k) + 1] = g2;
    d.collision_worldid[base + k] = w;
  }
}

template <class T>
void narrowphase(const Model<T>& m, const Data<T>& d, int w) {
  std::vector<Con<T>> cons;
  const int padr = d.world_pair[2 * w], np = d.world_pair[2 * w + 1];
  for (int p = 0; p < np; p++) narrowphase_pair(m, d, w, d.collision_pair[2 * (padr + p)], d.collision_pair[2 * (padr + p) + 1], cons);
  const int n = (int)cons.size();
  const int base = reserve(d.ncon, n);
  const int nw = n <= 0 ? 0 : std::max(0, std::min(n, d.nconmax - base));
  d.world_con[2 * w] = n ? base : 0;
  d.world_con[2 * w + 1] = nw;
  for (int k = 0; k < nw; k++) {
    const int i = base + k;
    d.contact_dist[i] = cons[k].dist;
    S3(d.contact_pos + 3 * i, cons[k].pos);
    std::memcpy(d.contact_frame + 9 * i, cons[k].frame, sizeof(T) * 9);
    d.contact_geom[2 * i] = cons[k].g1;
    d.contact_geom[2 * i + 1] = cons[k].g2;
    d.contact_worldid[i] = w;
    contact_params(m, d, i);
  }
}

// ===========================================================================
// constraint.py:22-73
template <class T>
void update_efc_row(const Model<T>& m, const Data<T>& d, int e, T pos, T invweight, const T* solref,
                    const T* solimp, T margin, bool refsafe, T Jqvel) {
  T tc = solref[0];
  const T dr = solref[1];
  T dmin = solimp[0], dmax = solimp[1], width = solimp[2], mid = solimp[3], power = solimp[4];
  if (refsafe) tc = std::max(tc, T(2) * m.opt.timestep);
  dmin = clampv(dmin, T(0.0001), T(0.9999));
  dmax = clampv(dmax, T(0.0001), T(0.9999));
  width = std::max(T(1e-15), width);
  mid = clampv(mid, T(0.0001), T(0.9999));
  power = std::max(T(1), power);
  T k = T(1) / (dmax * dmax * tc * tc * dr * dr);
  T b = T(2) / (dmax * tc);
  if (solref[0] <= T(0)) k = -solref[0] / (dmax * dmax);
  if (solref[1] <= T(0)) b = -solref[1] / dmax;
  const T x = std::fabs(pos) / width;
  const T ya = (T(1) / std::pow(mid, power - T(1))) * std::pow(x, power);
  const T yb = T(1) - (T(1) / std::pow(T(1) - mid, power - T(1))) * std::pow(T(1) - x, power);
  const T y = x < mid ? ya : yb;
  T imp = dmin + y * (dmax - dmin);
  imp = clampv(imp, dmin, dmax);
  if (x > T(1)) imp = dmax;
  d.efc_D[e] = T(1) / std::max(invweight * (T(1) - imp) / imp, T(1e-15));
  d.efc_aref[e] = -k * imp * pos - b * Jqvel;
  d.efc_pos[e] = pos + margin;
  d.efc_margin[e] = margin;
}

// constraint.py:76-101: translational Jacobian of `point` on `body` for dof i
template <class T>
V3<T> jac(const Model<T>& m, const Data<T>& d, int w, V3<T> point, int body, int i) {
  const int db = m.dof_bodyid[i];
  bool in_tree = db == 0;
  for (int p = body; p != 0; p = m.body_parentid[p])
    if (p == db) {
      in_tree = true;
      break;
    }
  if (!in_tree) return {0, 0, 0};
  const T* cd = d.cdof + ((size_t)w * m.nv + i) * 6;
  const V3<T> off = point - L3(d.subtree_com + ((size_t)w * m.nbody + m.body_rootid[body]) * 3);
  return L3(cd + 3) + cross(L3(cd), off);
}

// constraint.py:104-240 for one world
template <class T>
void make_constraint(const Model<T>& m, const Data<T>& d, int w) {
  if (m.opt.disableflags & mjhDSBL_CONSTRAINT) return;
  const int nv = m.nv;
  const bool refsafe = !(m.opt.disableflags & mjhDSBL_REFSAFE);
  const bool do_lim = !(m.opt.disableflags & mjhDSBL_LIMIT) && m.nlimited > 0;
  const bool do_con = !(m.opt.disableflags & mjhDSBL_CONTACT) && m.opt.cone == mjhCONE_PYRAMIDAL;
  const T* qpos = d.qpos + (size_t)w * m.nq;
  const T* qvel = d.qvel + (size_t)w * nv;
  struct Lim {
    int j;
    T pos, J;
  };
  std::vector<Lim> lims;
  if (do_lim)
    for (int k = 0; k < m.nlimited; k++) {
      const int j = m.jnt_limited_slide_hinge_adr[k];
      const T q = qpos[m.jnt_qposadr[j]];
      const T dmin = q - m.jnt_range[2 * j], dmax = m.jnt_range[2 * j + 1] - q;
      const T pos = std::min(dmin, dmax) - m.jnt_margin[j];
      if (pos < T(0)) lims.push_back({j, pos, (dmin < dmax) ? T(1) : T(-1)});
    }
  std::vector<int> cons;
  if (do_con) {
    const int ca = d.world_con[2 * w], nc = d.world_con[2 * w + 1];
    for (int c = ca; c < ca + nc; c++)
      if (d.contact_dim[c] == 3 && d.contact_dist[c] - d.contact_includemargin[c] < T(0)) cons.push_back(c);
  }
  const int total = (int)lims.size() + 4 * (int)cons.size();
  const int base = reserve(d.nefc, total);
  const int nw = total <= 0 ? 0 : std::max(0, std::min(total, d.njmax - base));
  d.world_efc[2 * w] = total ? base : 0;
  d.world_efc[2 * w + 1] = nw;
  int r = base;
  for (const Lim& l : lims) {
    if (r - base >= nw) break;
    T* J = d.efc_J + (size_t)r * nv;
    std::fill(J, J + nv, T(0));
    const int da = m.jnt_dofadr[l.j];
    J[da] = l.J;
    d.efc_worldid[r] = w;
    update_efc_row(m, d, r, l.pos, m.dof_invweight0[da], m.jnt_solref + 2 * l.j, m.jnt_solimp + 5 * l.j,
                   m.jnt_margin[l.j], refsafe, l.J * qvel[da]);
    r++;
  }
  for (int c : cons) {
    if (r - base + 4 > nw) break;
    const int b1 = m.geom_bodyid[d.contact_geom[2 * c]], b2 = m.geom_bodyid[d.contact_geom[2 * c + 1]];
    const T* fr = d.contact_friction + 5 * c;
    const T f0 = fr[0];
    T iw = m.body_invweight0[2 * b1] + m.body_invweight0[2 * b2];
    iw = iw + f0 * f0 * iw;
    iw = iw * T(2) * f0 * f0 / m.opt.impratio;
    const T* F = d.contact_frame + 9 * c;
    const V3<T> cp = L3(d.contact_pos + 3 * c);
    const T pos = d.contact_dist[c] - d.contact_includemargin[c];
    for (int e = 0; e < 4; e++) {
      const int row = e / 2 + 1;           // tangent row of the frame
      const T sgn = (e % 2 == 0) ? T(1) : T(-1);
      const T mu = fr[row - 1];
      T* J = d.efc_J + (size_t)(r + e) * nv;
      T Jqvel = 0;
      for (int i = 0; i < nv; i++) {
        const V3<T> jd = jac(m, d, w, cp, b2, i) - jac(m, d, w, cp, b1, i);
        const T d0 = F[0] * jd.x + F[1] * jd.y + F[2] * jd.z;
        const T dt = F[3 * row] * jd.x + F[3 * row + 1] * jd.y + F[3 * row + 2] * jd.z;
        J[i] = d0 + sgn * dt * mu;
        Jqvel += J[i] * qvel[i];
      }
      d.efc_worldid[r + e] = w;
      update_efc_row(m, d, r + e, pos, iw, d.contact_solref + 2 * c, d.contact_solimp + 5 * c,
                     d.contact_includemargin[c], refsafe, Jqvel);
    }
    r += 4;
  }
}

// ===========================================================================
// solver.py:26-1000 for one world
template <class T>
struct P3 {
  T a, b, c;
};
template <class T> P3<T> eval_pt(const T* q, T alpha) {
  return {alpha * alpha * q[2] + alpha * q[1] + q[0], T(2) * alpha * q[2] + q[1], T(2) * q[2]};
}
template <class T> T safe_div(T x, T y) { return x / (y != T(0) ? y : T(1e-15)); }
template <class T> bool in_bracket(P3<T> x, P3<T> y) { return (x.b < y.b && y.b < T(0)) || (x.b > y.b && y.b > T(0)); }

template <class T>
struct Solver {
  const Model<T>& m;
  const Data<T>& d;
  int w, nv, r0, nr;
  T *qacc, *Ma, *grad, *Mgrad, *search, *mv, *qc;
  const T *qs, *qacc_smooth, *M;

  T* J(int r) const { return d.efc_J + (size_t)(r0 + r) * nv; }

  // solver.py:75-173
  void update_constraint() {
    d.efc_prev_cost[w] = d.efc_cost[w];
    T cost = 0, gauss = 0;
    for (int r = 0; r < nr; r++) {
      const int e = r0 + r;
      const T jar = d.efc_Jaref[e];
      const int act = jar < T(0);
      d.efc_active[e] = act;
      d.efc_force[e] = act ? T(-1) * d.efc_D[e] * jar : T(0);
      if (act) cost += T(0.5) * d.efc_D[e] * jar * jar;
    }
    for (int i = 0; i < nv; i++) {
      T s = 0;
      for (int r = 0; r < nr; r++) s += J(r)[i] * d.efc_force[r0 + r];
      qc[i] = s;
    }
    for (int i = 0; i < nv; i++) gauss += T(0.5) * (Ma[i] - qs[i]) * (qacc[i] - qacc_smooth[i]);
    d.efc_gauss[w] = gauss;
    d.efc_cost[w] = cost + gauss;
  }

  // solver.py:176-312
  void update_gradient() {
    T gd = 0;
    for (int i = 0; i < nv; i++) {
      grad[i] = Ma[i] - qs[i] - qc[i];
      gd += grad[i] * grad[i];
    }
    d.efc_grad_dot[w] = gd;
    if (m.opt.solver == mjhSOL_CG) {
      solve_m(m, d, w, d.efc_Mgrad, d.efc_grad);
    } else {
      T* H = d.efc_h + (size_t)w * nv * nv;
      for (int i = 0; i < nv; i++)
        for (int j = 0; j <= i; j++) {
          T h = M[i * nv + j];
          for (int r = 0; r < nr; r++) {
            const int e = r0 + r;
            if (d.efc_D[e] == T(0) || !d.efc_active[e]) continue;
            h += J(r)[i] * J(r)[j] * d.efc_D[e];
          }
          H[i * nv + j] = h;
        }
      std::vector<T> L(nv * nv);
      cholesky_block(H, L.data(), 0, nv, nv);
      cho_solve_block(L.data(), 0, nv, nv, Mgrad, grad);
    }
  }

  P3<T> rows_at(T alpha) {
    P3<T> s{0, 0, 0};
    for (int r = 0; r < nr; r++) {
      const int e = r0 + r;
      if (d.efc_Jaref[e] + alpha * d.efc_jv[e] < T(0)) {
        const P3<T> p = eval_pt(d.efc_quad + 3 * e, alpha);
        s.a += p.a;
        s.b += p.b;
        s.c += p.c;
      }
    }
    return s;
  }

  // solver.py:339-664
  void linesearch_iterative() {
    const T* qg = d.efc_quad_gauss + 3 * w;
    const T gtol = m.opt.tolerance * m.opt.ls_tolerance * std::sqrt(d.efc_search_dot[w]) * m.meaninertia * T(std::max(1, nv));
    d.efc_gtol[w] = gtol;
    P3<T> p0{qg[0], qg[1], T(2) * qg[2]};
    for (int r = 0; r < nr; r++) {
      const int e = r0 + r;
      if (d.efc_Jaref[e] < T(0)) {
        const T* q = d.efc_quad + 3 * e;
        p0.a += q[0];
        p0.b += q[1];
        p0.c += T(2) * q[2];
      }
    }
    T lo_alpha = -safe_div(p0.b, p0.c);
    P3<T> lo = eval_pt(qg, lo_alpha);
    {
      P3<T> s = rows_at(lo_alpha);
      lo = {lo.a + s.a, lo.b + s.b, lo.c + s.c};
    }
    const bool lo_less = lo.b < p0.b;
    P3<T> hi = lo_less ? p0 : lo;
    T hi_alpha = lo_less ? T(0) : lo_alpha;
    if (!lo_less) {
      lo = p0;
      lo_alpha = 0;
    }
    bool done = false;
    T alpha = d.efc_alpha[w];
    for (int it = 0; it < m.opt.ls_iterations && !done; it++) {
      const T la = lo_alpha - safe_div(lo.b, lo.c);
      const T ha = hi_alpha - safe_div(hi.b, hi.c);
      const T ma = T(0.5) * (lo_alpha + hi_alpha);
      P3<T> ln = eval_pt(qg, la), hn = eval_pt(qg, ha), mn = eval_pt(qg, ma);
      P3<T> s1 = rows_at(la), s2 = rows_at(ha), s3 = rows_at(ma);
      ln = {ln.a + s1.a, ln.b + s1.b, ln.c + s1.c};
      hn = {hn.a + s2.a, hn.b + s2.b, hn.c + s2.c};
      mn = {mn.a + s3.a, mn.b + s3.b, mn.c + s3.c};
      bool sl = false, sh = false;
      if (in_bracket(lo, ln)) { lo = ln; lo_alpha = la; sl = true; }
      if (in_bracket(lo, mn)) { lo = mn; lo_alpha = ma; sl = true; }
      if (in_bracket(lo, hn)) { lo = hn; lo_alpha = ha; sl = true; }
      if (in_bracket(hi, hn)) { hi = hn; hi_alpha = ha; sh = true; }
      if (in_bracket(hi, mn)) { hi = mn; hi_alpha = ma; sh = true; }
      if (in_bracket(hi, ln)) { hi = ln; hi_alpha = la; sh = true; }
      done = (!sl && !sh) || (lo.b < T(0) && lo.b > -gtol) || (hi.b > T(0) && hi.b < gtol);
      const bool improved = lo.a < p0.a || hi.a < p0.a;
      alpha = 0;
      if (improved) alpha = lo.a < hi.a ? lo_alpha : hi_alpha;
    }
    d.efc_alpha[w] = alpha;
  }

  // solver.py:667-734
  void linesearch_parallel() {
    const T* qg = d.efc_quad_gauss + 3 * w;
    int best = 0;
    T bestc = 0;
    for (int k = 0; k < m.nlsp; k++) {
      const T a = m.alpha_candidate[k];
      T t[3] = {qg[0], qg[1], qg[2]};
      for (int r = 0; r < nr; r++) {
        const int e = r0 + r;
        if (d.efc_Jaref[e] + a * d.efc_jv[e] < T(0))
          for (int c = 0; c < 3; c++) t[c] += d.efc_quad[3 * e + c];
      }
      const T cost = a * a * t[2] + a * t[1] + t[0];
      d.efc_cost_candidate[w * m.nlsp + k] = cost;
      if (k == 0 || cost < bestc) {
        best = k;
        bestc = cost;
      }
    }
    d.efc_alpha[w] = m.alpha_candidate[best];
  }

  // solver.py:737-869
  void linesearch() {
    mul_m(m, M, mv, search);
    for (int r = 0; r < nr; r++) {
      T s = 0;
      for (int i = 0; i < nv; i++) s += J(r)[i] * search[i];
      d.efc_jv[r0 + r] = s;
    }
    T g0 = 0, g1 = 0, g2 = 0;
    for (int i = 0; i < nv; i++) {
      g0 += d.efc_gauss[w] / T(nv);
      g1 += search[i] * (Ma[i] - qs[i]);
      g2 += T(0.5) * search[i] * mv[i];
    }
    d.efc_quad_gauss[3 * w] = g0;
    d.efc_quad_gauss[3 * w + 1] = g1;
    d.efc_quad_gauss[3 * w + 2] = g2;
    for (int r = 0; r < nr; r++) {
      const int e = r0 + r;
      const T jar = d.efc_Jaref[e], jv = d.efc_jv[e], D = d.efc_D[e];
      d.efc_quad[3 * e] = T(0.5) * jar * jar * D;
      d.efc_quad[3 * e + 1] = jv * jar * D;
      d.efc_quad[3 * e + 2] = T(0.5) * jv * jv * D;
    }
    if (m.opt.ls_parallel) linesearch_parallel();
    else linesearch_iterative();
    const T alpha = d.efc_alpha[w];
    for (int i = 0; i < nv; i++) {
      qacc[i] += alpha * search[i];
      Ma[i] += alpha * mv[i];
    }
    for (int r = 0; r < nr; r++) d.efc_Jaref[r0 + r] += alpha * d.efc_jv[r0 + r];
  }

  void search_update(T beta) {
    T sd = 0;
    for (int i = 0; i < nv; i++) {
      T s = T(-1) * Mgrad[i];
      if (m.opt.solver == mjhSOL_CG) s += beta * search[i];
      search[i] = s;
      sd += s * s;
    }
    d.efc_search_dot[w] = sd;
  }
};

// solver.py:872-1000
template <class T>
void solve(const Model<T>& m, const Data<T>& d, int w) {
  const int nv = m.nv;
  const size_t o = (size_t)w * nv;
  Solver<T> s{m, d, w, nv, d.world_efc[2 * w], d.world_efc[2 * w + 1], d.qacc + o, d.efc_Ma + o, d.efc_grad + o,
              d.efc_Mgrad + o, d.efc_search + o, d.efc_mv + o, d.qfrc_constraint + o, d.qfrc_smooth + o,
              d.qacc_smooth + o, d.qMd + o * nv};
  for (int i = 0; i < nv; i++) s.qacc[i] = d.qacc_warmstart[o + i];
  d.efc_cost[w] = std::numeric_limits<T>::infinity();
  d.efc_solver_niter[w] = 0;
  d.efc_done[w] = 0;
  d.efc_search_dot[w] = 0;
  for (int r = 0; r < s.nr; r++) {
    const int e = s.r0 + r;
    T a = 0;
    for (int i = 0; i < nv; i++) a += s.J(r)[i] * s.qacc[i] - d.efc_aref[e] / T(nv);
    d.efc_Jaref[e] = a;
  }
  mul_m(m, s.M, s.Ma, s.qacc);
  s.update_constraint();
  s.update_gradient();
  s.search_update(T(0));
  const int iters = m.opt.iterations;
  for (int it = 0; it < iters; it++) {
    if (iters > 1 && d.efc_done[w]) break;
    s.linesearch();
    if (m.opt.solver == mjhSOL_CG)
      for (int i = 0; i < nv; i++) {
        d.efc_prev_grad[o + i] = s.grad[i];
        d.efc_prev_Mgrad[o + i] = s.Mgrad[i];
      }
    s.update_constraint();
    s.update_gradient();
    T beta = 0;
    if (m.opt.solver == mjhSOL_CG) {
      T num = 0, den = 0;
      for (int i = 0; i < nv; i++) {
        num += s.grad[i] * (s.Mgrad[i] - d.efc_prev_Mgrad[o + i]);
        den += d.efc_prev_grad[o + i] * d.efc_prev_Mgrad[o + i];
      }
      beta = std::max(T(0), num / std::max(T(1e-15), den));
      d.efc_beta[w] = beta;
    }
    s.search_update(beta);
    const T scale = m.meaninertia * T(std::max(1, nv));
    const T improvement = (d.efc_prev_cost[w] - d.efc_cost[w]) / scale;
    const T gradient = std::sqrt(d.efc_grad_dot[w]) / scale;
    d.efc_done[w] = (improvement < m.opt.tolerance) || (gradient < m.opt.tolerance);
  }
  for (int i = 0; i < nv; i++) d.qacc_warmstart[o + i] = s.qacc[i];
}

// ===========================================================================
// forward.py:43-162 _advance (time is advanced by the caller)
template <class T>
void advance(const Model<T>& m, const Data<T>& d, int w, const T* qacc_all) {
  const int nv = m.nv;
  const T dt = m.opt.timestep;
  for (int a = 0; a < m.nu && m.na; a++) {
    const int adr = m.actuator_actadr[a];
    if (adr == -1) continue;
    T& x = d.act[(size_t)w * m.na + adr];
    const T xd = d.act_dot[(size_t)w * m.na + adr];
    if (m.actuator_dyntype[a] == mjhDYN_FILTEREXACT) {
      const T tau = std::max(m.actuator_dynprm[10 * a], T(1e-15));
      x = x + xd * tau * (T(1) - std::exp(-dt / tau));
    } else {
      x = x + xd * dt;
    }
  }
  T* qvel = d.qvel + (size_t)w * nv;
  for (int i = 0; i < nv; i++) qvel[i] = qvel[i] + qacc_all[(size_t)w * nv + i] * dt;
  T* qpos = d.qpos + (size_t)w * m.nq;
  for (int j = 0; j < m.njnt; j++) {
    const int qa = m.jnt_qposadr[j], da = m.jnt_dofadr[j];
    switch (m.jnt_type[j]) {
      case mjhJNT_FREE: {
        S3(qpos + qa, L3(qpos + qa) + L3(qvel + da) * dt);
        SQ(qpos + qa + 3, quat_integrate(LQ(qpos + qa + 3), L3(qvel + da + 3), dt));
        break;
      }
      case mjhJNT_BALL:
        SQ(qpos + qa, quat_integrate(LQ(qpos + qa), L3(qvel + da), dt));
        break;
      default:
        qpos[qa] = qpos[qa] + dt * qvel[da];
    }
  }
}

// forward.py:165-242
template <class T>
void euler(const Model<T>& m, const Data<T>& d, int w) {
  if (m.opt.disableflags & mjhDSBL_EULERDAMP) {
    advance(m, d, w, d.qacc);
    return;
  }
  const int nv = m.nv;
  const size_t o = (size_t)w * nv;
  std::vector<T> A(nv * nv), L(nv * nv, T(0)), f(nv);
  for (int k = 0; k < nv * nv; k++) A[k] = d.qMd[o * nv + k];
  for (int i = 0; i < nv; i++) {
    A[i * nv + i] += m.dof_damping[i] * m.opt.timestep;
    f[i] = d.qfrc_smooth[o + i] + d.qfrc_constraint[o + i];
  }
  each_block(m, [&](int off, int n) {
    cholesky_block(A.data(), L.data(), off, n, nv);
    cho_solve_block(L.data(), off, n, nv, d.qacc_integration + o, f.data());
  });
  if (m.opt.is_sparse) {
    // forward.py:171-193 eulerdamp_sparse: the integration matrix and its
    // L'DL factor in the sparse layout
    T* Mi = d.qM_integration + (size_t)w * m.nM;
    dense_to_sparse(m, A.data(), Mi);
    factor_sparse(m, Mi, d.qLD_integration + (size_t)w * m.nM, d.qLDiagInv_integration + o);
    for (int i = 0; i < nv; i++) d.qfrc_integration[o + i] = f[i];
  }
  advance(m, d, w, d.qacc_integration);
}

// forward.py:245-405 (implicitfast, dense)
template <class T>
void implicit(const Model<T>& m, const Data<T>& d, int w) {
  const int nv = m.nv, nu = m.nu;
  const bool pas = !(m.opt.disableflags & mjhDSBL_PASSIVE);
  const bool act = !(m.opt.disableflags & mjhDSBL_ACTUATION) && m.actuator_affine_bias_gain;
  if (!pas && !act) {
    advance(m, d, w, d.qacc);
    return;
  }
  const size_t o = (size_t)w * nv;
  T* avi = d.act_vel_integration + (size_t)w * nu;
  if (act)
    for (int a = 0; a < nu; a++) {
      T bv = 0, gv = 0;
      if (m.actuator_biastype[a] == mjhBIAS_AFFINE) bv = m.actuator_biasprm[10 * a + 2];
      if (m.actuator_gaintype[a] == mjhGAIN_AFFINE) gv = m.actuator_gainprm[10 * a + 2];
      const T c = m.actuator_dyntype[a] != mjhDYN_NONE ? d.act[(size_t)w * m.na + a] : d.ctrl[(size_t)w * nu + a];
      avi[a] = bv + gv * c;
    }
  T* Mi = d.qM_integration + o * nv;
  const T* M = d.qMd + o * nv;
  const T* mom = d.actuator_moment + (size_t)w * nu * nv;
  each_act_tile(m, m.nqLD_tile, [&](int onv, int onu, int tnv, int tnu) {
    if (tnv == 0) return;
    for (int r = 0; r < tnv; r++)
      for (int c = 0; c < tnv; c++) {
        T q = 0;
        if (act && tnu != 0)
          for (int a = 0; a < tnu; a++) q += mom[(onu + a) * nv + onv + r] * avi[onu + a] * mom[(onu + a) * nv + onv + c];
        if (pas && r == c) q += -m.dof_damping[onv + r];
        Mi[(onv + r) * nv + onv + c] = M[(onv + r) * nv + onv + c] - q * m.opt.timestep;
      }
    for (int j = 0; j < tnv; j++) d.qfrc_integration[o + onv + j] = d.qfrc_smooth[o + onv + j] + d.qfrc_constraint[o + onv + j];
  });
  std::vector<T> L(nv * nv, T(0));
  each_block(m, [&](int off, int n) {
    cholesky_block(Mi, L.data(), off, n, nv);
    cho_solve_block(L.data(), off, n, nv, d.qacc_integration + o, d.qfrc_integration + o);
  });
  advance(m, d, w, d.qacc_integration);
}

// ---- per-world stage groups ----
template <class T>
void fwd_position(const Model<T>& m, const Data<T>& d, int w) {
  kinematics(m, d, w);
  com_pos(m, d, w);
  crb(m, d, w);
  factor_m(m, d, w);
  broadphase(m, d, w, 0);
  narrowphase(m, d, w);
  make_constraint(m, d, w);
  transmission(m, d, w);
}

template <class T>
void forward(const Model<T>& m, const Data<T>& d, int w) {
  fwd_position(m, d, w);
  fwd_velocity(m, d, w);
  fwd_actuation(m, d, w);
  fwd_acceleration(m, d, w);
  if (d.njmax == 0) {
    for (int i = 0; i < m.nv; i++) d.qacc[(size_t)w * m.nv + i] = d.qacc_smooth[(size_t)w * m.nv + i];
  } else {
    solve(m, d, w);
  }
}

template <class T>
void step(const Model<T>& m, const Data<T>& d, int w) {
  forward(m, d, w);
  if (m.opt.integrator == mjhINT_EULER) euler(m, d, w);
  else implicit(m, d, w);
}

template <class T>
void reset_counters(const Model<T>& m, const Data<T>& d, int which) {
  if (which & 1) {
    d.ncollision[0] = 0;
    d.ncon[0] = 0;
  }
  if ((which & 2) && !(m.opt.disableflags & mjhDSBL_CONSTRAINT)) d.nefc[0] = 0;
}

// runs f on every world; worlds are independent except for the pool
// reservations, which are atomic
template <class T, class F>
void each_world(const Data<T>& d, int nthreads, F f) {
#ifdef _OPENMP
  if (nthreads > 1) {
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
    for (int w = 0; w < d.nworld; w++) f(w);
    return;
  }
#endif
  for (int w = 0; w < d.nworld; w++) f(w);
}

}  // namespace oracle

// ---------------------------------------------------------------------------
// C entry points: oracle_<stage>_<f32|f64>(model, data[, extra], nthreads)
#define ORACLE_STAGE(NAME, BODY)                                                          \
  extern "C" int oracle_##NAME##_f32(const oracle::Model<float>* m, const oracle::Data<float>* d, int nt) { \
    using T = float;                                                                       \
    const auto& M = *m;                                                                    \
    const auto& D = *d;                                                                    \
    (void)sizeof(T);                                                                       \
    BODY;                                                                                  \
    return 0;                                                                              \
  }                                                                                        \
  extern "C" int oracle_##NAME##_f64(const oracle::Model<double>* m, const oracle::Data<double>* d, int nt) { \
    using T = double;                                                                      \
    const auto& M = *m;                                                                    \
    const auto& D = *d;                                                                    \
    (void)sizeof(T);                                                                       \
    BODY;                                                                                  \
    return 0;                                                                              \
  }

using namespace oracle;

ORACLE_STAGE(kinematics, each_world(D, nt, [&](int w) { kinematics<T>(M, D, w); }))
ORACLE_STAGE(com_pos, each_world(D, nt, [&](int w) { com_pos<T>(M, D, w); }))
ORACLE_STAGE(crb, each_world(D, nt, [&](int w) { crb<T>(M, D, w); }))
ORACLE_STAGE(factor_m, each_world(D, nt, [&](int w) { factor_m<T>(M, D, w); }))
ORACLE_STAGE(rne, each_world(D, nt, [&](int w) { rne<T>(M, D, w); }))
ORACLE_STAGE(transmission, each_world(D, nt, [&](int w) { transmission<T>(M, D, w); }))
ORACLE_STAGE(com_vel, each_world(D, nt, [&](int w) { com_vel<T>(M, D, w); }))
ORACLE_STAGE(passive, each_world(D, nt, [&](int w) { passive<T>(M, D, w); }))
ORACLE_STAGE(nxn_broadphase, reset_counters(M, D, 1); each_world(D, nt, [&](int w) { broadphase<T>(M, D, w, 1); }))
ORACLE_STAGE(sap_broadphase, reset_counters(M, D, 1); each_world(D, nt, [&](int w) { broadphase<T>(M, D, w, 2); }))
ORACLE_STAGE(primitive_narrowphase, each_world(D, nt, [&](int w) { narrowphase<T>(M, D, w); }))
ORACLE_STAGE(collision, reset_counters(M, D, 1); each_world(D, nt, [&](int w) {
               broadphase<T>(M, D, w, 0);
               narrowphase<T>(M, D, w);
             }))
ORACLE_STAGE(make_constraint, reset_counters(M, D, 2); each_world(D, nt, [&](int w) { make_constraint<T>(M, D, w); }))
ORACLE_STAGE(fwd_position, reset_counters(M, D, 3); each_world(D, nt, [&](int w) { fwd_position<T>(M, D, w); }))
ORACLE_STAGE(fwd_velocity, each_world(D, nt, [&](int w) { fwd_velocity<T>(M, D, w); }))
ORACLE_STAGE(fwd_actuation, each_world(D, nt, [&](int w) { fwd_actuation<T>(M, D, w); }))
ORACLE_STAGE(fwd_acceleration, each_world(D, nt, [&](int w) { fwd_acceleration<T>(M, D, w); }))
ORACLE_STAGE(solve, each_world(D, nt, [&](int w) { solve<T>(M, D, w); }))
ORACLE_STAGE(euler, each_world(D, nt, [&](int w) { euler<T>(M, D, w); }))
ORACLE_STAGE(implicit, each_world(D, nt, [&](int w) { implicit<T>(M, D, w); }))
ORACLE_STAGE(forward, reset_counters(M, D, 3); each_world(D, nt, [&](int w) { forward<T>(M, D, w); }))
ORACLE_STAGE(step, reset_counters(M, D, 3); each_world(D, nt, [&](int w) { step<T>(M, D, w); }))

// closest points between segments (math.py:209-242), for the known-answer tests
extern "C" void oracle_closest_segments_f64(const double* a0, const double* a1, const double* b0, const double* b1,
                                            double* ra, double* rb) {
  V3<double> x, y;
  closest_segment_to_segment_points(L3(a0), L3(a1), L3(b0), L3(b1), &x, &y);
  S3(ra, x);
  S3(rb, y);
}
extern "C" void oracle_closest_segments_f32(const float* a0, const float* a1, const float* b0, const float* b1,
                                            float* ra, float* rb) {
  V3<float> x, y;
  closest_segment_to_segment_points(L3(a0), L3(a1), L3(b0), L3(b1), &x, &y);
  S3(ra, x);
  S3(rb, y);
}
