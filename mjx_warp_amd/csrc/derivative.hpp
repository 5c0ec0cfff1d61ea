// Transition derivatives of one step by finite differences, and one step of
// the adjoint recursion over them.
//
// The reference has no gradients (its kernels overwrite arrays in place and
// use integer atomics; collision_convex.py:22 turns backward code generation
// off).  The interface follows MuJoCo's own transition derivatives,
// mjd_transitionFD (engine_derivative_fd.c): the state is
// x = (dq [nv], qvel [nv], act [na]) with dq the tangent-space position
// (a local-frame rotation vector for free / ball joints, as mj_integratePos
// and mj_differentiatePos use it), the input is u = ctrl [nu], and
//   A[i][j] = d x'_i / d x_j   [nx, nx],   B[i][k] = d x'_i / d u_k   [nx, nu].
//
// Perturbed copies of each world live in a second Data ("fd") of
//   K = 1 + S * (nx + nu)  worlds per nominal world, S = 2 (centered) or 1:
// fd world w*K holds the nominal state and w*K + 1 + c*S + s column c's
// perturbation on side s (centered: s = 0 is +eps, s = 1 is -eps; forward:
// +eps, or -eps for a control that cannot move up).  One ordinary step of the
// fd Data then gives every perturbed successor; the differences become the
// columns of A and B.  A control column respects ctrlrange the way
// mjd_transitionFD does: a side that would leave the range is not used
// (one-sided difference), and a control that can move neither way gets a
// zero column.
#pragma once

#include "common.hpp"
#include "forward.hpp"

namespace mjh {

MJH_DEV int fd_nx(const mjhModel& m) { return 2 * m.nv + m.na; }
MJH_DEV int fd_worlds_per(const mjhModel& m, int centered) { return 1 + (centered ? 2 : 1) * (fd_nx(m) + m.nu); }

// Whether control a may move up / down by eps inside its range.
MJH_DEV void fd_ctrl_sides(const mjhModel& m, int a, float u, float eps, bool* up, bool* down) {
  const bool limited = m.actuator_ctrllimited[a] && !(m.opt.disableflags & mjhDSBL_CLAMPCTRL);
  if (!limited) {
    *up = *down = true;
    return;
  }
  const float lo = m.actuator_ctrlrange[2 * a], hi = m.actuator_ctrlrange[2 * a + 1];
  *up = u + eps <= hi;
  *down = u - eps >= lo;
}

// Copies nominal world w's inputs into fd world f and applies f's
// perturbation (one fd world per team).
template <int TEAM>
MJH_DEV void fd_perturb(const mjhModel& m, const mjhData& d, const mjhData& fd, float eps, int centered,
                        const Team<TEAM>& t) {
  if (!t.valid) return;
  const int nv = m.nv, nx = fd_nx(m), S = centered ? 2 : 1;
  const int K = fd_worlds_per(m, centered);
  const int f = t.world, w = f / K, k = f - w * K;
  auto copy = [&](float* dst, const float* src, int n) {
    for (int i = t.lane; i < n; i += TEAM) dst[(size_t)f * n + i] = src[(size_t)w * n + i];
  };
  copy(fd.qpos, d.qpos, m.nq);
  copy(fd.qvel, d.qvel, nv);
  copy(fd.act, d.act, m.na);
  copy(fd.ctrl, d.ctrl, m.nu);
  copy(fd.qacc_warmstart, d.qacc_warmstart, nv);
  copy(fd.qfrc_applied, d.qfrc_applied, nv);
  copy(fd.xfrc_applied, d.xfrc_applied, 6 * m.nbody);
  copy(fd.mocap_pos, d.mocap_pos, 3 * m.nmocap);
  copy(fd.mocap_quat, d.mocap_quat, 4 * m.nmocap);
  sync();
  if (k == 0 || t.lane != 0) return;
  const int c = (k - 1) / S, s = (k - 1) - c * S;
  float h = s ? -eps : eps;
  if (c < nv) {
    float* qpos = fd.qpos + (size_t)f * m.nq;
    const int j = m.dof_jntid[c], jt = m.jnt_type[j];
    const int qa = m.jnt_qposadr[j], r = c - m.jnt_dofadr[j];
    if (jt == mjhJNT_FREE && r < 3) {
      qpos[qa + r] += h;
    } else if (jt == mjhJNT_FREE || jt == mjhJNT_BALL) {
      const int q0 = jt == mjhJNT_FREE ? qa + 3 : qa;
      const int axis = jt == mjhJNT_FREE ? r - 3 : r;
      const V3 e = V3{axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f};
      st4(qpos + q0, quat_integrate(ld4(qpos + q0), e, h));
    } else {
      qpos[qa] += h;
    }
  } else if (c < 2 * nv) {
    fd.qvel[(size_t)f * nv + (c - nv)] += h;
  } else if (c < nx) {
    fd.act[(size_t)f * m.na + (c - 2 * nv)] += h;
  } else {
    const int a = c - nx;
    float* u = fd.ctrl + (size_t)f * m.nu + a;
    if (!centered) {
      bool up, down;
      fd_ctrl_sides(m, a, *u, eps, &up, &down);
      if (!up) h = -eps;
    }
    *u += h;
  }
}

// Unit quaternion of the rotation vector v.
MJH_DEV Q4 quat_exp(V3 v) { return axis_angle_to_quat(normalize(v), length(v)); }

// Right Jacobian of the rotation exponential at phi applied to b:
// exp(phi + b) = exp(phi) exp(J_r(phi) b + O(b^2)),
// J_r = I - (1 - cos th) / th^2 [phi]x + (th - sin th) / th^3 [phi]x^2.
MJH_DEV V3 right_jacobian(V3 phi, V3 b) {
  const float th2 = dot(phi, phi), th = sqrtf(th2);
  float c1, c2;
  if (th < 1e-2f) {
    c1 = 0.5f - th2 / 24.0f;
    c2 = 1.0f / 6.0f - th2 / 120.0f;
  } else {
    c1 = (1.0f - cosf(th)) / th2;
    c2 = (th - sinf(th)) / (th2 * th);
  }
  const V3 pb = cross(phi, b);
  return b - c1 * pb + c2 * cross(phi, pb);
}

// Column c of [A | B] for nominal world w (one team per (w, c)), from the
// stepped fd Data.  The successor difference x'[fp] (-) x'[fm] is taken
// through the step's last stage, qvel' = qvel + dt a and qpos' = qpos (+)
// dt qvel' (a: the acceleration the integrator advanced with), from the
// input difference of column c and dt (a[fp] - a[fm]).  That equals the
// difference of the outputs, without differencing two FP32 positions of
// order 1 whose perturbation-induced change is a few ulps: slide, hinge and
// translation exactly; a rotation through
//   q'[fm]^-1 q'[fp] = exp(-p) exp(dh e) exp(p + b)
//                    = exp(R(p)' dh e) exp(J_r(p) b + O(b^2)),
// p = dt qvel'[fm], b = dt (qvel'[fp] - qvel'[fm]).
template <int TEAM>
MJH_DEV void fd_difference(const mjhModel& m, const mjhData& fd, int nworld, float eps, int centered, float* A,
                           float* B, const Team<TEAM>& t) {
  if (!t.valid) return;
  const int nv = m.nv, nx = fd_nx(m), nc = nx + m.nu, S = centered ? 2 : 1;
  const int K = fd_worlds_per(m, centered);
  const int g = t.world, w = g / nc, c = g - w * nc;
  (void)nworld;
  const int f0 = w * K, f1 = f0 + 1 + c * S;
  // worlds differenced and the perturbation of column c each holds
  int fp = f1, fm = centered ? f1 + 1 : f0;
  float hp = eps, hm = centered ? -eps : 0.0f;
  if (c >= nx) {
    const int a = c - nx;
    bool up, down;
    // the nominal world's control is unperturbed
    fd_ctrl_sides(m, a, fd.ctrl[(size_t)f0 * m.nu + a], eps, &up, &down);
    if (centered) {
      if (up && !down) {
        fm = f0;
        hm = 0.0f;
      } else if (down && !up) {
        fp = f0;
        hp = 0.0f;
      } else if (!up && !down) {
        hp = hm = 0.0f;
      }
    } else if (!up) {
      // f1 holds the -eps perturbation
      fp = f0;
      fm = f1;
      hp = 0.0f;
      hm = down ? -eps : 0.0f;
    }
  }
  const float dh = hp - hm;
  const float scale = dh != 0.0f ? 1.0f / dh : 0.0f;
  float* col;
  size_t stride;
  if (c < nx) {
    col = A + (size_t)w * nx * nx + c;
    stride = nx;
  } else {
    col = B + (size_t)w * nx * m.nu + (c - nx);
    stride = m.nu;
  }
  const float dt = m.opt.timestep;
  const float* acc = advances_with_qacc(m) ? fd.qacc : fd.qacc_integration;
  const float* ap = acc + (size_t)fp * nv;
  const float* am = acc + (size_t)fm * nv;
  // qvel'[fp] - qvel'[fm] of dof i
  auto dvel = [&](int i) { return (c == nv + i ? dh : 0.0f) + dt * (ap[i] - am[i]); };
  for (int j = t.lane; j < m.njnt; j += TEAM) {
    const int jt = m.jnt_type[j], da = m.jnt_dofadr[j];
    if (jt == mjhJNT_FREE || jt == mjhJNT_BALL) {
      int d0 = da;
      if (jt == mjhJNT_FREE) {
        for (int r = 0; r < 3; r++) col[(da + r) * stride] = ((c == da + r ? dh : 0.0f) + dt * dvel(da + r)) * scale;
        d0 += 3;
      }
      const V3 p = dt * ld3(fd.qvel + (size_t)fm * nv + d0);
      const V3 b = dt * V3{dvel(d0), dvel(d0 + 1), dvel(d0 + 2)};
      const int r = c - d0;
      const V3 in = (r >= 0 && r < 3) ? V3{r == 0 ? dh : 0.0f, r == 1 ? dh : 0.0f, r == 2 ? dh : 0.0f}
                                      : V3{0.0f, 0.0f, 0.0f};
      const V3 din = rot_vec_quat(in, quat_inv(quat_exp(p)));
      const V3 rot = quat_to_vel(mul_quat(quat_exp(din), quat_exp(right_jacobian(p, b))));
      col[(d0 + 0) * stride] = rot.x * scale;
      col[(d0 + 1) * stride] = rot.y * scale;
      col[(d0 + 2) * stride] = rot.z * scale;
    } else {
      col[da * stride] = ((c == da ? dh : 0.0f) + dt * dvel(da)) * scale;
    }
  }
  for (int i = t.lane; i < nv; i += TEAM) col[(nv + i) * stride] = dvel(i) * scale;
  const float* xp = fd.act + (size_t)fp * m.na;
  const float* xm = fd.act + (size_t)fm * m.na;
  for (int i = t.lane; i < m.na; i += TEAM) col[(2 * nv + i) * stride] = (xp[i] - xm[i]) * scale;
}

// d(loss)/d(dq) of position dof i from the gradient with respect to the raw
// qpos coordinates at qpos: a rotation dof of a free / ball joint moves the
// quaternion q by 0.5 q (0, e_r) per unit.
MJH_DEV float qpos_grad_to_tangent(const mjhModel& m, const float* qpos, const float* gq, int i) {
  const int j = m.dof_jntid[i], jt = m.jnt_type[j];
  const int qa = m.jnt_qposadr[j], r = i - m.jnt_dofadr[j];
  if (jt == mjhJNT_FREE && r < 3) return gq[qa + r];
  if (jt == mjhJNT_FREE || jt == mjhJNT_BALL) {
    const int q0 = jt == mjhJNT_FREE ? qa + 3 : qa;
    const int axis = jt == mjhJNT_FREE ? r - 3 : r;
    const Q4 e = Q4{0.0f, axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f};
    const Q4 dq = mul_quat(ld4(qpos + q0), e);
    const Q4 g = ld4(gq + q0);
    return 0.5f * (dq.w * g.w + dq.x * g.x + dq.y * g.y + dq.z * g.z);
  }
  return gq[qa];
}

// One step of the adjoint recursion for nominal world w (one team per
// world):  lam = A' lam_next + g_x,  grad_ctrl = B' lam_next, where g_x is
// the loss gradient at this state (g_qpos raw, turned into the tangent
// space at d.qpos; g_qvel; g_act).  Null gradients and lam_next are zero.
template <int TEAM>
MJH_DEV void adjoint_step(const mjhModel& m, const mjhData& d, const float* A, const float* B, const float* g_qpos,
                          const float* g_qvel, const float* g_act, const float* lam_next, float* lam,
                          float* grad_ctrl, const Team<TEAM>& t) {
  if (!t.valid) return;
  const int w = t.world, nv = m.nv, nx = fd_nx(m), nu = m.nu;
  const float* Aw = A + (size_t)w * nx * nx;
  const float* Bw = B + (size_t)w * nx * nu;
  const float* ln = lam_next ? lam_next + (size_t)w * nx : nullptr;
  for (int i = t.lane; i < nx; i += TEAM) {
    float s = 0.0f;
    if (ln)
      for (int j = 0; j < nx; j++) s += Aw[(size_t)j * nx + i] * ln[j];
    float g = 0.0f;
    if (i < nv) {
      if (g_qpos) g = qpos_grad_to_tangent(m, d.qpos + (size_t)w * m.nq, g_qpos + (size_t)w * m.nq, i);
    } else if (i < 2 * nv) {
      if (g_qvel) g = g_qvel[(size_t)w * nv + (i - nv)];
    } else if (g_act) {
      g = g_act[(size_t)w * m.na + (i - 2 * nv)];
    }
    lam[(size_t)w * nx + i] = s + g;
  }
  if (!grad_ctrl) return;
  for (int k = t.lane; k < nu; k += TEAM) {
    float s = 0.0f;
    if (ln)
      for (int j = 0; j < nx; j++) s += Bw[(size_t)j * nu + k] * ln[j];
    grad_ctrl[(size_t)w * nu + k] = s;
  }
}

// Splits lam [nworld, nx] into gradients with respect to d's raw state:
// qpos (a rotation's tangent gradient g becomes 2 q (0, g), orthogonal to
// q), qvel and act.  One team per world.
template <int TEAM>
MJH_DEV void adjoint_state_grad(const mjhModel& m, const mjhData& d, const float* lam, float* g_qpos, float* g_qvel,
                                float* g_act, const Team<TEAM>& t) {
  if (!t.valid) return;
  const int w = t.world, nv = m.nv, nx = fd_nx(m);
  const float* l = lam + (size_t)w * nx;
  const float* qpos = d.qpos + (size_t)w * m.nq;
  float* gq = g_qpos + (size_t)w * m.nq;
  for (int j = t.lane; j < m.njnt; j += TEAM) {
    const int jt = m.jnt_type[j], qa = m.jnt_qposadr[j], da = m.jnt_dofadr[j];
    if (jt == mjhJNT_FREE || jt == mjhJNT_BALL) {
      int q0 = qa, d0 = da;
      if (jt == mjhJNT_FREE) {
        gq[qa] = l[da];
        gq[qa + 1] = l[da + 1];
        gq[qa + 2] = l[da + 2];
        q0 += 3;
        d0 += 3;
      }
      const Q4 g = mul_quat(ld4(qpos + q0), Q4{0.0f, l[d0], l[d0 + 1], l[d0 + 2]});
      st4(gq + q0, Q4{2.0f * g.w, 2.0f * g.x, 2.0f * g.y, 2.0f * g.z});
    } else {
      gq[qa] = l[da];
    }
  }
  for (int i = t.lane; i < nv; i += TEAM) g_qvel[(size_t)w * nv + i] = l[nv + i];
  if (g_act)
    for (int i = t.lane; i < m.na; i += TEAM) g_act[(size_t)w * m.na + i] = l[2 * nv + i];
}

}  // namespace mjh
