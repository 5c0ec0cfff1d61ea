// Analytic reverse mode of one step: the vector-Jacobian product
//   (lam_x, lam_u) = (dx'/dx)' lam',  (dx'/du)' lam'
// of the step x' = step(x, u), x = (dq, qvel) in the tangent space of
// mjd_transitionFD (engine_derivative_fd.c), at the state a step was just
// taken from.  The reference has no gradients (collision_convex.py:22 turns
// Warp's backward code generation off); this is the hand-written adjoint of
// the forward path the library runs (forward.py:682-695).
//
// The step is differentiated as
//   qpos' = qpos (+) dt qvel',  qvel' = qvel + dt a,
//   M_hat a = qfrc_smooth + J' f,          M_hat = M + dt diag(damping)
//                                           (implicitfast / Euler damping)
//   H a_s = qfrc_smooth + J_A' D_A aref_A, H = M + J_A' D_A J_A,
//   f_A = -D_A (J_A a_s - aref_A),          f = 0 off the active set A
// i.e. implicit differentiation of the converged constraint optimum
// (solver.py:176-312 assembles the same Hessian) with the active set held:
// one extra Cholesky solve with M_hat and one with H per step.
//
// Configuration derivatives use the screw of each dof: moving dof k by e
// moves every body and joint axis after it on its chain by e cdof_k, so a
// motion vector X attached to them changes by cdof_k x X, an inertia I by
// cdof_k x* I - I cdof_k x (the spatial algebra of rne).  With the
// reference point of the com-frame quantities held fixed (the generalized
// forces are invariant to it), the derivative of mu' rne(q, qvel, acc)
// along dof k is a sum over the bodies after k of
//   df_b . V_b + f_b . dV_b,   f_b = I_b cacc_b + cvel_b x* I_b cvel_b,
//   V_b = sum of mu_j cdof_j over b's chain,
// where each body quantity's change is the screw applied to its part after
// k (its value minus the prefix accumulated up to k), plus the change of
// cdof_dot of the moved joints.  Velocity derivatives of the bias force
// follow the same pattern with d cvel_b = cdof_k.
//
// Constraint rows: aref and D depend on the row's distance through the
// impedance (constraint.py:22-73); the distance of a limit row moves with
// its Jacobian row, a contact's with the normal row (the mean of its first
// two pyramid rows).  The contact Jacobian changes with the configuration
// through the joint axes (the screw term), the contact point and the
// contact frame, differentiated per pair type as the narrowphase builds
// them: plane pairs (the sphere centre / capsule end / box corner is a
// point of the body, the contact point sits half the distance off it, the
// plane-capsule frame's first tangent follows the capsule axis), sphere /
// capsule pairs (closest points of the core segments sliding as their
// optimality conditions require, n = dir / |dir|, frame = make_frame(n)).
// Other pair types move the contact point with the mean of the two bodies'
// points and hold the frame.
#pragma once

#include "common.hpp"
#include "constraint.hpp"
#include "derivative.hpp"
#include "forward.hpp"
#include "smooth.hpp"

namespace mjh {

// LDS floats of adjoint_vjp: per body 6 six-vectors and the com-frame
// inertia, per dof 6 prefixes, cdof and cdof_dot, 7 nv-vectors (one of them
// integer), the nv x nv matrix and the register factor's two 32-float
// broadcast buffers, 8 per row, and two per (dof, body) pair.
__host__ __device__ inline int lds_adjoint_floats(int nbody, int nv, int rows) {
  return 46 * nbody + 48 * nv + 7 * nv + nv * nv + 64 + 8 * rows + 2 * nv * nbody;
}

// k, b, impedance and its derivative with respect to the row's distance,
// as update_efc_row computes them.
MJH_DEV void efc_coefs(const mjhModel& m, float pos, const float* solref, const float* solimp, bool refsafe,
                       float* k_out, float* b_out, float* imp_out, float* dimp_out) {
  float timeconst = solref[0];
  const float dampratio = solref[1];
  float dmin = solimp[0], dmax = solimp[1], width = solimp[2], mid = solimp[3], power = solimp[4];
  if (refsafe) timeconst = fmaxf(timeconst, 2.0f * m.opt.timestep);
  dmin = fminf(fmaxf(dmin, MJH_MINIMP), MJH_MAXIMP);
  dmax = fminf(fmaxf(dmax, MJH_MINIMP), MJH_MAXIMP);
  width = fmaxf(MJH_MINVAL, width);
  mid = fminf(fmaxf(mid, MJH_MINIMP), MJH_MAXIMP);
  power = fmaxf(1.0f, power);
  float k = 1.0f / (dmax * dmax * timeconst * timeconst * dampratio * dampratio);
  float b = 2.0f / (dmax * timeconst);
  if (solref[0] <= 0.0f) k = -solref[0] / (dmax * dmax);
  if (solref[1] <= 0.0f) b = -solref[1] / dmax;
  const float x = fabsf(pos) / width;
  float y, dy;
  if (x < mid) {
    const float s = 1.0f / powf(mid, power - 1.0f);
    y = s * powf(x, power);
    dy = s * power * powf(x, power - 1.0f);
  } else {
    const float s = 1.0f / powf(1.0f - mid, power - 1.0f);
    y = 1.0f - s * powf(1.0f - x, power);
    dy = s * power * powf(1.0f - x, power - 1.0f);
  }
  float imp = dmin + y * (dmax - dmin);
  float dimp = (dmax - dmin) * dy * (pos > 0.0f ? 1.0f : (pos < 0.0f ? -1.0f : 0.0f)) / width;
  if (imp < dmin || imp > dmax) dimp = 0.0f;
  imp = fminf(fmaxf(imp, dmin), dmax);
  if (x > 1.0f) {
    imp = dmax;
    dimp = 0.0f;
  }
  *k_out = k;
  *b_out = b;
  *imp_out = imp;
  *dimp_out = dimp;
}

MJH_DEV V3 s6_rot(const S6& a) { return V3{a.v[0], a.v[1], a.v[2]}; }
MJH_DEV V3 s6_lin(const S6& a) { return V3{a.v[3], a.v[4], a.v[5]}; }
MJH_DEV S6 s6_make(V3 r, V3 l) { return S6{{r.x, r.y, r.z, l.x, l.y, l.z}}; }
MJH_DEV S6 s6_axpy(float a, const S6& x, const S6& y) {
  S6 r;
#pragma unroll
  for (int i = 0; i < 6; i++) r.v[i] = a * x.v[i] + y.v[i];
  return r;
}
MJH_DEV S6 s6_sub(const S6& x, const S6& y) { return s6_axpy(-1.0f, y, x); }
MJH_DEV S6 s6_add(const S6& x, const S6& y) { return s6_axpy(1.0f, y, x); }
// d(I x) for an inertia moved by the screw s: s x* (I x) - I (s x x)
MJH_DEV S6 inert_moved(const float* I, const S6& s, const S6& x) {
  return s6_sub(motion_cross_force(s, inert_vec(I, x)), inert_vec(I, motion_cross(s, x)));
}

// closest_segment_point and its tangent (dA, dB, dPT: tangents of a, b, pt)
MJH_DEV V3 csp_tangent(V3 a, V3 b, V3 pt, V3 dA, V3 dB, V3 dPT, V3* dout) {
  const V3 ab = b - a, dab = dB - dA;
  const float den = dot(ab, ab) + 1e-6f;
  const float tt = dot(pt - a, ab) / den;
  const float dtt = (dot(dPT - dA, ab) + dot(pt - a, dab)) / den - tt * 2.0f * dot(ab, dab) / den;
  const float tc = fminf(fmaxf(tt, 0.0f), 1.0f);
  const float dtc = (tt > 0.0f && tt < 1.0f) ? dtt : 0.0f;
  *dout = dA + dtc * ab + tc * dab;
  return a + tc * ab;
}
// closest_segment_to_segment_points (collision.hpp) and the tangent of its
// two points for tangents dA0.. of the segment ends: the same operations,
// clamps and choice, differentiated
MJH_DEV void segseg_tangent(V3 a0, V3 a1, V3 b0, V3 b1, V3 dA0, V3 dA1, V3 dB0, V3 dB1, V3* ra, V3* rb, V3* dra,
                            V3* drb) {
  V3 da = a1 - a0, db = b1 - b0, dda = dA1 - dA0, ddb = dB1 - dB0;
  const float la = length(da), lb = length(db);
  if (la != 0.0f) {
    da = da * (1.0f / la);
    dda = (dda - da * dot(da, dda)) * (1.0f / la);
  }
  if (lb != 0.0f) {
    db = db * (1.0f / lb);
    ddb = (ddb - db * dot(db, ddb)) * (1.0f / lb);
  }
  const float dla = la != 0.0f ? dot(da, dA1 - dA0) : 0.0f, dlb = lb != 0.0f ? dot(db, dB1 - dB0) : 0.0f;
  const float ha = la * 0.5f, hb = lb * 0.5f, dha = 0.5f * dla, dhb = 0.5f * dlb;
  const V3 am = a0 + da * ha, bm = b0 + db * hb;
  const V3 dam = dA0 + dda * ha + da * dha, dbm = dB0 + ddb * hb + db * dhb;
  const V3 tr = am - bm, dtr = dam - dbm;
  const float dadb = dot(da, db), datr = dot(da, tr), dbtr = dot(db, tr);
  const float ddadb = dot(dda, db) + dot(da, ddb), ddatr = dot(dda, tr) + dot(da, dtr);
  const float ddbtr = dot(ddb, tr) + dot(db, dtr);
  const float denom = 1.0f - dadb * dadb, ddenom = -2.0f * dadb * ddadb;
  const float num = -datr + dadb * dbtr, dnum = -ddatr + ddadb * dbtr + dadb * ddbtr;
  const float den = denom + 1e-6f;
  const float ota = num / den, dota = (dnum * den - num * ddenom) / (den * den);
  const float otb = dbtr + ota * dadb, dotb = ddbtr + dota * dadb + ota * ddadb;
  const float ta = fminf(fmaxf(ota, -ha), ha), tb = fminf(fmaxf(otb, -hb), hb);
  const float dta = ota < -ha ? -dha : (ota > ha ? dha : dota);
  const float dtb = otb < -hb ? -dhb : (otb > hb ? dhb : dotb);
  const V3 ba = am + da * ta, bb = bm + db * tb;
  const V3 dba = dam + dda * ta + da * dta, dbb = dbm + ddb * tb + db * dtb;
  V3 dna, dnb;
  const V3 na = csp_tangent(a0, a1, bb, dA0, dA1, dbb, &dna);
  const V3 nb = csp_tangent(b0, b1, ba, dB0, dB1, dba, &dnb);
  const float d1 = dot(bb - na, bb - na), d2 = dot(ba - nb, ba - nb);
  if (d1 < d2) {
    *ra = na;
    *rb = bb;
    *dra = dna;
    *drb = dbb;
  } else {
    *ra = ba;
    *rb = nb;
    *dra = dba;
    *drb = dnb;
  }
}

// One world per workgroup of 64 lanes.  d holds the step just taken from
// (qpos_t, qvel_t): the forward quantities at that state and qvel' in
// d.qvel.  lam_next [nworld, nx] is lam'; lam [nworld, nx] gets
// (dx'/dx)' lam' (+ gadd [nworld, nx] if given), grad_ctrl [nworld, nu]
// (dx'/du)' lam'.  rcap: rows per world the LDS holds; a world with more
// active rows gets NaN gradients.
template <int TEAM>
MJH_DEV void adjoint_vjp(const mjhModel& m, const mjhData& d, const float* qpos_t_all, const float* qvel_t_all,
                         const float* lam_next, const float* gadd, float* lam, float* grad_ctrl, int rcap,
                         float* lds, const Team<TEAM>& t) {
  static_assert(TEAM == 64, "one world per wavefront");
  if (!t.valid) return;
  const int w = t.world, nv = m.nv, nb = m.nbody, nu = m.nu, nx = 2 * nv + m.na;
  const float dt = m.opt.timestep;
  const int dis = m.opt.disableflags;
  const bool passive_on = !(dis & mjhDSBL_PASSIVE);
  const float* qpos_t = qpos_t_all + (size_t)w * m.nq;
  const float* vt = qvel_t_all + (size_t)w * nv;
  const float* v1 = d.qvel + (size_t)w * nv;
  const float* ln = lam_next + (size_t)w * nx;
  const float* cdof_g = WPTR(d.cdof, w, nv * 6);
  const float* cdot_g = WPTR(d.cdof_dot, w, nv * 6);
  const float* cinert_g = WPTR(d.cinert, w, nb * 10);
  const float* sc = WPTR(d.subtree_com, w, nb * 3);
  const float* qM = WPTR(d.qMd, w, nv * nv);
  const float* a_s = WPTR(d.qacc, w, nv);
  const float* a_i = advances_with_qacc(m) ? a_s : WPTR(d.qacc_integration, w, nv);
  const bool damp = !advances_with_qacc(m) && (m.opt.integrator == mjhINT_EULER || passive_on);

  float* cv = lds;
  float* caA = cv + 6 * nb;
  float* caS = caA + 6 * nb;
  float* Vm = caS + 6 * nb;
  float* Vn = Vm + 6 * nb;
  float* Va = Vn + 6 * nb;
  float* pv = Va + 6 * nb;
  float* paA = pv + 6 * nv;
  float* paS = paA + 6 * nv;
  float* pm = paS + 6 * nv;
  float* pn = pm + 6 * nv;
  float* pa = pn + 6 * nv;
  float* lq = pa + 6 * nv;
  float* lv = lq + nv;
  float* mu = lv + nv;
  float* nuv = mu + nv;
  float* wv = nuv + nv;
  float* tmp = wv + nv;
  int* blk = (int*)(tmp + nv);
  float* A = (float*)(blk + nv);
  float* rlf = A + nv * nv + 64;
  float* rg = rlf + rcap;
  float* rcq = rg + rcap;
  float* rcv = rcq + rcap;
  float* ral = rcv + rcap;
  float* rbe = ral + rcap;
  float* rga = rbe + rcap;
  int* rsrc = (int*)(rga + rcap);
  // the motion axes, their derivatives and the com-frame inertias, read
  // many times by the per-dof loops below
  float* cdof = (float*)(rsrc + rcap);
  float* cdot = cdof + 6 * nv;
  float* cinert = cdot + 6 * nv;
  for (int i = t.lane; i < 6 * nv; i += TEAM) {
    cdof[i] = cdof_g[i];
    cdot[i] = cdot_g[i];
  }
  for (int i = t.lane; i < 10 * nb; i += TEAM) cinert[i] = cinert_g[i];

  // ---- integration: qpos' = qpos (+) dt qvel', qvel' = qvel + dt a
  for (int i = t.lane; i < nv; i += TEAM) {
    lq[i] = 0.0f;
    lv[i] = ln[nv + i];
    int beg, end;
    tree_block(m, i, &beg, &end);
    blk[i] = beg;
    tmp[i] = end;
  }
  sync();
  for (int j = t.lane; j < m.njnt; j += TEAM) {
    const int jt = m.jnt_type[j], da = m.jnt_dofadr[j];
    if (jt == mjhJNT_FREE) {
      for (int r = 0; r < 3; r++) {
        lq[da + r] += ln[da + r];
        lv[da + r] += dt * ln[da + r];
      }
      // rotation: dq' = R(th)' dq + dt J_r(th) dw', th = dt w'
      const V3 th = dt * ld3(v1 + da + 3);
      const V3 l = ld3(ln + da + 3);
      const V3 rl = rot_vec_quat(l, quat_exp(th));
      const V3 jl = right_jacobian(-1.0f * th, l);
      lq[da + 3] += rl.x;
      lq[da + 4] += rl.y;
      lq[da + 5] += rl.z;
      lv[da + 3] += dt * jl.x;
      lv[da + 4] += dt * jl.y;
      lv[da + 5] += dt * jl.z;
    } else {
      lq[da] += ln[da];
      lv[da] += dt * ln[da];
    }
  }
  sync();

  // ---- mu = M_hat^-1 (dt lam_qvel'), M_hat block diagonal by tree.
  // nv <= 32: a row per lane in registers (RegChol; both 32-lane halves of
  // the wave factor the same matrix, each with its own broadcast buffer);
  // larger: the LDS factor
  const bool reg = nv <= kRegNvMax;
  const int rl = t.lane & 31;
  float* cbuf = A + nv * nv + (t.lane & 32);
  int rbeg = rl, rend = rl + 1;
  if (reg && rl < nv) tree_block(m, rl, &rbeg, &rend);
  if (reg) {
    const float dg = (rl < nv && damp) ? dt * m.dof_damping[rl] : 0.0f;
    const float b = rl < nv ? dt * lv[rl] : 0.0f;
    const float x = with_nvc(nv, [&](auto nvc) {
      RegChol<decltype(nvc)::value> c{rl, cbuf};
      c.load_sym(qM, nv, rbeg, rend, dg);
      c.factor();
      return c.solve(b);
    });
    if (t.lane < nv) mu[t.lane] = x;
    sync();
  } else {
    for (int p = t.lane; p < nv * nv; p += TEAM) {
      const int r = p / nv, c = p % nv;
      float v = (c >= blk[r] && c < (int)tmp[r]) ? qM[p] : 0.0f;
      if (r == c && damp) v += dt * m.dof_damping[r];
      A[p] = v;
    }
    for (int i = t.lane; i < nv; i += TEAM) mu[i] = dt * lv[i];
    sync();
    chol_lds(A, nv, nv, m, t);
    cho_solve_lds(A, nv, nv, mu, t);
    sync();
  }

  // ---- constraint rows of this world
  int base = 0, n = 0;
  if (!(dis & mjhDSBL_CONSTRAINT)) {
    base = d.world_efc[2 * w];
    n = d.world_efc[2 * w + 1];
  }
  if (n > rcap) {
    for (int i = t.lane; i < nx; i += TEAM) lam[(size_t)w * nx + i] = __int_as_float(0x7fc00000);
    for (int i = t.lane; i < nu; i += TEAM) grad_ctrl[(size_t)w * nu + i] = __int_as_float(0x7fc00000);
    return;
  }
  const float* J = d.efc_J + (size_t)base * nv;
  const float* fr = d.efc_force + base;
  const float* Dr = d.efc_D + base;
  for (int i = t.lane; i < n; i += TEAM) {
    float s = 0.0f;
    if (fr[i] > 0.0f)
      for (int j = 0; j < nv; j++) s += J[(size_t)i * nv + j] * mu[j];
    rlf[i] = s;
  }
  sync();
  // H = M + J_A' D_A J_A, rhs -J_A' D_A lam_f
  for (int j = t.lane; j < nv; j += TEAM) {
    float s = 0.0f;
    for (int i = 0; i < n; i++)
      if (fr[i] > 0.0f) s -= J[(size_t)i * nv + j] * Dr[i] * rlf[i];
    nuv[j] = s;
  }
  sync();
  if (n == 0) {
    for (int i = t.lane; i < nv; i += TEAM) nuv[i] = 0.0f;
  } else if (reg) {
    // the lane's row of H: its block of M, then D_i J_i[row] J_i per active row
    const float b = rl < nv ? nuv[rl] : 0.0f;
    const float x = with_nvc(nv, [&](auto nvc) {
      constexpr int NVC = decltype(nvc)::value;
      RegChol<NVC> c{rl, cbuf};
      c.load_sym(qM, nv, rbeg, rend, 0.0f);
      for (int i = 0; i < n; i++) {
        if (!(fr[i] > 0.0f)) continue;
        const float* Ji = J + (size_t)i * nv;
        const float g = rl < nv ? Dr[i] * Ji[rl] : 0.0f;
#pragma unroll
        for (int j = 0; j < NVC; j++)
          if (j < nv) c.h[j] = fmaf(g, Ji[j], c.h[j]);
      }
      c.factor();
      return c.solve(b);
    });
    sync();
    if (t.lane < nv) nuv[t.lane] = x;
  } else {
    for (int p = t.lane; p < nv * nv; p += TEAM) {
      const int r = p / nv, c = p % nv;
      if (c > r) continue;
      float v = (c >= blk[r] && c < (int)tmp[r]) ? qM[p] : 0.0f;
      for (int i = 0; i < n; i++)
        if (fr[i] > 0.0f) v += Dr[i] * J[(size_t)i * nv + r] * J[(size_t)i * nv + c];
      A[p] = v;
    }
    sync();
    chol_lds(A, nv, nv, m, t);
    cho_solve_lds(A, nv, nv, nuv, t);
  }
  sync();
  for (int i = t.lane; i < n; i += TEAM) {
    float s = 0.0f;
    if (fr[i] > 0.0f) {
      s = rlf[i];
      for (int j = 0; j < nv; j++) s += J[(size_t)i * nv + j] * nuv[j];
    }
    rg[i] = s;
  }
  // each row's source, in make_constraint's order: active limits in joint
  // order, then four rows per active contact
  if (t.lane == 0 && n > 0) {
    int r = 0;
    if (!(dis & mjhDSBL_LIMIT))
      for (int k = 0; k < m.nlimited && r < n; k++) {
        float pos, Jl;
        const int j = m.jnt_limited_slide_hinge_adr[k];
        if (limit_active(m, qpos_t, j, &pos, &Jl)) rsrc[r++] = -(j + 1);
      }
    if (!(dis & mjhDSBL_CONTACT) && m.opt.cone == mjhCONE_PYRAMIDAL) {
      const int cadr = d.world_con[2 * w], ncon = d.world_con[2 * w + 1];
      for (int c = cadr; c < cadr + ncon && r + 4 <= n; c++)
        if (d.contact_dim[c] == 3 && d.contact_dist[c] - d.contact_includemargin[c] < 0.0f)
          for (int e = 0; e < 4; e++) rsrc[r++] = 4 * c + e;
    }
    for (; r < n; r++) rsrc[r] = 0x7fffffff;  // a row of no known source: no distance terms
  }
  sync();
  const bool refsafe = !(dis & mjhDSBL_REFSAFE);
  for (int i = t.lane; i < n; i += TEAM) {
    float cq = 0.0f, cvr = 0.0f, al = 0.0f, be = 0.0f, ga = 0.0f, cd = 0.0f;
    const int src = rsrc[i];
    if (fr[i] > 0.0f && src != 0x7fffffff) {
      const float* solref;
      const float* solimp;
      if (src < 0) {
        const int j = -src - 1;
        solref = m.jnt_solref + 2 * j;
        solimp = m.jnt_solimp + 5 * j;
      } else {
        solref = d.contact_solref + 2 * (src >> 2);
        solimp = d.contact_solimp + 5 * (src >> 2);
      }
      const float pos = d.efc_pos[base + i] - d.efc_margin[base + i];
      float k, b, imp, dimp;
      efc_coefs(m, pos, solref, solimp, refsafe, &k, &b, &imp, &dimp);
      const float D = Dr[i], g = rg[i], f = fr[i];
      const float laref = D * g;
      const float lD = (f / D) * g;
      cq = laref * (-k * imp - k * pos * dimp);
      if (dimp != 0.0f && imp > 0.0f && imp < 1.0f) cq += lD * D / (imp * (1.0f - imp)) * dimp;
      cvr = -b * laref;
      if (src >= 0) {
        al = f;
        be = -D * g;
        ga = -D * g * b;
        // sphere / capsule pairs: the distance's derivative comes from the
        // narrowphase's own closest points (contact loop below), not the
        // normal row
        const int c = src >> 2;
        const int t1 = m.geom_type[d.contact_geom[2 * c]], t2 = m.geom_type[d.contact_geom[2 * c + 1]];
        if ((t1 == mjhGEOM_SPHERE || t1 == mjhGEOM_CAPSULE) && (t2 == mjhGEOM_SPHERE || t2 == mjhGEOM_CAPSULE)) {
          cd = cq;
          cq = 0.0f;
        }
      }
    }
    rlf[i] = cd;  // (the row multipliers are no longer needed)
    rcq[i] = cq;
    rcv[i] = cvr;
    ral[i] = al;
    rbe[i] = be;
    rga[i] = ga;
  }
  sync();
  // distance and aref terms: d pos / dq is the limit's Jacobian row, or the
  // contact's normal row (the mean of its first two pyramid rows)
  for (int j = t.lane; j < nv; j += TEAM) {
    float sq = 0.0f, sv = 0.0f;
    for (int i = 0; i < n; i++) {
      if (rcq[i] == 0.0f && rcv[i] == 0.0f) continue;
      const float Jij = J[(size_t)i * nv + j];
      sv += rcv[i] * Jij;
      const int src = rsrc[i];
      if (src < 0) {
        sq += rcq[i] * Jij;
      } else {
        const int r0 = i - (src & 3);
        sq += rcq[i] * 0.5f * (J[(size_t)r0 * nv + j] + J[(size_t)(r0 + 1) * nv + j]);
      }
    }
    lq[j] += sq;
    lv[j] += sv;
    wv[j] = mu[j] + nuv[j];
  }
  sync();

  // ---- smooth forces: actuators, damping, springs
  if (nu && !(dis & mjhDSBL_ACTUATION)) {
    const float* moment = WPTR(d.actuator_moment, w, nu * nv);
    const float* ctrl = WPTR(d.ctrl, w, nu);
    const float* len = WPTR(d.actuator_length, w, nu);
    const float* vel = WPTR(d.actuator_velocity, w, nu);
    const bool noclamp = dis & mjhDSBL_CLAMPCTRL;
    // d qfrc_actuator / d ctrl_a = gain_a (inside ctrlrange) times the
    // moment entries the step applies: all of them for sparse models, the
    // dense tiles' (the reference's tile loop) otherwise
    auto act = [&](int a, int onv, int tnv) {
      bool inr = true;
      if (m.actuator_ctrllimited[a] && !noclamp) {
        const float c = ctrl[a];
        inr = c > m.actuator_ctrlrange[2 * a] && c < m.actuator_ctrlrange[2 * a + 1];
      }
      float s = 0.0f;
      if (inr) {
        const float* g = m.actuator_gainprm + 10 * a;
        const float gain = g[0] + g[1] * len[a] + g[2] * vel[a];
        for (int j = onv; j < onv + tnv; j++) s += moment[(size_t)a * nv + j] * wv[j];
        s *= gain;
      }
      grad_ctrl[(size_t)w * nu + a] = s;
    };
    if (m.opt.is_sparse) {
      for (int a = t.lane; a < nu; a += TEAM) act(a, 0, nv);
    } else {
      for (int a = t.lane; a < nu; a += TEAM) grad_ctrl[(size_t)w * nu + a] = 0.0f;
      sync();
      for_each_act_tile(m, m.nact_group, [&](int onv, int onu, int tnv, int tnu) {
        if (tnu == 0 || tnv == 0) return;
        for (int a = t.lane; a < tnu; a += TEAM) act(onu + a, onv, tnv);
      });
    }
  } else {
    for (int a = t.lane; a < nu; a += TEAM) grad_ctrl[(size_t)w * nu + a] = 0.0f;
  }
  if (passive_on)
    for (int j = t.lane; j < nv; j += TEAM) {
      lv[j] -= m.dof_damping[j] * wv[j];
      const int jn = m.dof_jntid[j];
      const int jt = m.jnt_type[jn];
      if (jt == mjhJNT_HINGE || jt == mjhJNT_SLIDE) lq[j] -= m.jnt_stiffness[jn] * wv[j];
    }

  // ---- body velocities, accelerations (at a and a_s), V(mu), V(nu),
  // V(a_s) and their per-dof prefixes (lanes over the six components)
  if (t.lane < 6) {
    const int c = t.lane;
    const bool nograv = dis & mjhDSBL_GRAVITY;
    const float gl = c == 3 ? m.opt.gravity[0] : (c == 4 ? m.opt.gravity[1] : m.opt.gravity[2]);
    const float g0 = (c < 3 || nograv) ? 0.0f : -gl;
    cv[c] = 0.0f;
    caA[c] = g0;
    caS[c] = g0;
    Vm[c] = 0.0f;
    Vn[c] = 0.0f;
    Va[c] = 0.0f;
    for (int b = 1; b < nb; b++) {
      const int p = m.body_parentid[b];
      float xv = cv[6 * p + c], xa = caA[6 * p + c], xs = caS[6 * p + c];
      float xm = Vm[6 * p + c], xn = Vn[6 * p + c], xx = Va[6 * p + c];
      const int ja = m.body_jntadr[b], jn = m.body_jntnum[b];
      for (int j = ja; j < ja + jn; j++) {
        const int jt = m.jnt_type[j], da = m.jnt_dofadr[j];
        auto add = [&](int i, bool with_dot) {
          const float cd = cdof[6 * i + c], dd = with_dot ? cdot[6 * i + c] * vt[i] : 0.0f;
          xv += cd * vt[i];
          xa += cd * a_i[i] + dd;
          xs += cd * a_s[i] + dd;
          xm += cd * mu[i];
          xn += cd * nuv[i];
          xx += cd * a_s[i];
        };
        auto prefix = [&](int i) {
          pv[6 * i + c] = xv;
          paA[6 * i + c] = xa;
          paS[6 * i + c] = xs;
          pm[6 * i + c] = xm;
          pn[6 * i + c] = xn;
          pa[6 * i + c] = xx;
        };
        if (jt == mjhJNT_FREE) {
          for (int i = 0; i < 3; i++) add(da + i, false);
          for (int i = 0; i < 6; i++) prefix(da + i);
          for (int i = 3; i < 6; i++) add(da + i, true);
        } else {
          add(da, true);
          prefix(da);
        }
      }
      cv[6 * b + c] = xv;
      caA[6 * b + c] = xa;
      caS[6 * b + c] = xs;
      Vm[6 * b + c] = xm;
      Vn[6 * b + c] = xn;
      Va[6 * b + c] = xx;
    }
  }
  sync();

  // ---- configuration and velocity derivatives of rne, a lane per (dof k,
  // body b) pair into LDS, then summed per dof in body order (the same sum
  // every run)
  float* pq = cinert + 10 * nb;
  float* pw = pq + nv * nb;
  for (int pr = t.lane; pr < nv * nb; pr += TEAM) {
    const int k = pr / nb, b = pr - k * nb;
    const int kb = m.dof_bodyid[k];
    pq[pr] = 0.0f;
    pw[pr] = 0.0f;
    if (b < kb || !m.body_dof_tree[b * nv + k]) continue;
    const S6 S = ld6(cdof + 6 * k);
    const int kj = m.dof_jntid[k];
    const bool ktrans = m.jnt_type[kj] == mjhJNT_FREE && k - m.jnt_dofadr[kj] < 3;
    const S6 Uv = ld6(pv + 6 * k);
    const S6 SUv = motion_cross(S, Uv);
    // velocity derivative: the prefix through k, and k's own cdof_dot
    const S6 P = (m.jnt_type[kj] == mjhJNT_FREE && !ktrans) ? ld6(cv + 6 * kb) : Uv;
    const S6 cdk = ktrans ? S6{{0.f, 0.f, 0.f, 0.f, 0.f, 0.f}} : ld6(cdot + 6 * k);
    float accq = 0.0f;
    const float* I = cinert + 10 * b;
    const S6 cvb = ld6(cv + 6 * b);
    const S6 Iv = inert_vec(I, cvb);
    const S6 dvv = motion_cross(S, s6_sub(cvb, Uv));
    const S6 dIcv = inert_moved(I, S, cvb);
    const S6 gyro = s6_add(motion_cross_force(dvv, Iv), motion_cross_force(cvb, s6_add(dIcv, inert_vec(I, dvv))));
    const S6 cvI = motion_cross_force(cvb, Iv);
    // (a, mu) and (a_s, nu)
#pragma unroll
    for (int which = 0; which < 2; which++) {
      const S6 ca = ld6((which == 0 ? caA : caS) + 6 * b);
      const S6 Uac = ld6((which == 0 ? paA : paS) + 6 * k);
      const S6 V = ld6((which == 0 ? Vm : Vn) + 6 * b);
      const S6 UV = ld6((which == 0 ? pm : pn) + 6 * k);
      const S6 dca = s6_sub(motion_cross(S, s6_sub(ca, Uac)), motion_cross(SUv, s6_sub(cvb, Uv)));
      const S6 fb = s6_add(inert_vec(I, ca), cvI);
      const S6 df = s6_add(s6_add(inert_moved(I, S, ca), inert_vec(I, dca)), gyro);
      const S6 dV = motion_cross(S, s6_sub(V, UV));
      accq += dot6(df, V) + dot6(fb, dV);
    }
    // bias force along qvel_k, left vector mu + nu
    const S6 Vw = s6_add(ld6(Vm + 6 * b), ld6(Vn + 6 * b));
    const S6 dcv = s6_add(motion_cross(S, s6_sub(cvb, P)), cdk);
    const S6 dfv = s6_add(s6_add(inert_vec(I, dcv), motion_cross_force(S, Iv)),
                          motion_cross_force(cvb, inert_vec(I, S)));
    pq[pr] = accq;
    pw[pr] = dot6(dfv, Vw);
  }
  sync();
  for (int k = t.lane; k < nv; k += TEAM) {
    float sq = 0.0f, sw = 0.0f;
    for (int b = 0; b < nb; b++) {
      sq += pq[k * nb + b];
      sw += pw[k * nb + b];
    }
    tmp[k] = sq;
    wv[k] = sw;
  }
  sync();
  // ---- contact rows moving with the configuration, lane per dof k
  for (int k = t.lane; k < nv; k += TEAM) {
    const S6 S = ld6(cdof + 6 * k);
    const S6 Uv = ld6(pv + 6 * k), Um = ld6(pm + 6 * k), Un = ld6(pn + 6 * k), Ua = ld6(pa + 6 * k);
    // contact Jacobian rows moving with the configuration
    float accc = 0.0f;
    for (int i = 0; i < n; i++) {
      const int src = rsrc[i];
      if (src < 0 || src == 0x7fffffff || fr[i] <= 0.0f) continue;
      const int c = src >> 2, e = src & 3;
      const float al = ral[i], be = rbe[i], ga = rga[i];
      const int g1 = d.contact_geom[2 * c], g2 = d.contact_geom[2 * c + 1];
      const int b1 = m.geom_bodyid[g1], b2 = m.geom_bodyid[g2];
      const V3 p = ld3(d.contact_pos + 3 * c);
      const M33 fm = ld9(d.contact_frame + 9 * c);
      const V3 nrm = V3{fm.m[0], fm.m[1], fm.m[2]};
      const V3 t1 = V3{fm.m[3], fm.m[4], fm.m[5]};
      const V3 t2 = V3{fm.m[6], fm.m[7], fm.m[8]};
      const float* fri = d.contact_friction + 5 * c;
      const float sgn = (e & 1) ? -1.0f : 1.0f;
      const V3 dir = e < 2 ? nrm + (sgn * fri[0]) * t1 : nrm + (sgn * fri[1]) * t2;
      V3 om = V3{0.f, 0.f, 0.f}, vp = V3{0.f, 0.f, 0.f};
      V3 jp = V3{0.f, 0.f, 0.f};
#pragma unroll
      for (int side = 0; side < 2; side++) {
        const int b = side ? b2 : b1;
        if (b == 0) continue;
        const float sg = side ? 1.0f : -1.0f;
        const V3 cr = ld3(sc + 3 * m.body_rootid[b]);
        // V_b(x) for x = al (mu + nu) + be a_s + ga qvel
        const S6 Vx = s6_axpy(al, s6_add(ld6(Vm + 6 * b), ld6(Vn + 6 * b)),
                              s6_axpy(be, ld6(Va + 6 * b), s6_axpy(ga, ld6(cv + 6 * b), S6{{0.f, 0.f, 0.f, 0.f, 0.f, 0.f}})));
        om = om + sg * s6_rot(Vx);
        vp = vp + sg * (s6_lin(Vx) + cross(s6_rot(Vx), p - cr));
        if (!m.body_dof_tree[b * nv + k]) continue;
        // joint axes after k: -(S x* F) . (V_b(x) - prefix_k(x))
        const S6 F = s6_make(sg * cross(p - cr, dir), sg * dir);
        const S6 Ux = s6_axpy(al, s6_add(Um, Un), s6_axpy(be, Ua, s6_axpy(ga, Uv, S6{{0.f, 0.f, 0.f, 0.f, 0.f, 0.f}})));
        accc -= dot6(motion_cross_force(S, F), s6_sub(Vx, Ux));
        jp = jp + 0.5f * (s6_lin(S) + cross(s6_rot(S), p - cr));
      }
      const int ty1 = m.geom_type[g1], ty2 = m.geom_type[g2];
      const bool plane = ty1 == mjhGEOM_PLANE && (ty2 == mjhGEOM_SPHERE || ty2 == mjhGEOM_CAPSULE || ty2 == mjhGEOM_BOX);
      const bool round = (ty1 == mjhGEOM_SPHERE || ty1 == mjhGEOM_CAPSULE) &&
                         (ty2 == mjhGEOM_SPHERE || ty2 == mjhGEOM_CAPSULE);
      V3 dp = jp;
      if (plane) {
        // p = centre - (r + dist / 2) n with the centre (sphere centre,
        // capsule end, box corner) a point of body 2
        dp = V3{0.f, 0.f, 0.f};
        if (b2 != 0 && m.body_dof_tree[b2 * nv + k]) {
          const float r = ty2 == mjhGEOM_BOX ? 0.0f : m.geom_size[3 * g2];
          const V3 cen = p + nrm * (r + 0.5f * d.contact_dist[c]);
          const V3 jc = s6_lin(S) + cross(s6_rot(S), cen - ld3(sc + 3 * m.body_rootid[b2]));
          dp = jc - (0.5f * dot(nrm, jc)) * nrm;
        }
        // plane-capsule frame: the first tangent follows the capsule axis
        if (ty2 == mjhGEOM_CAPSULE && b2 != 0 && m.body_dof_tree[b2 * nv + k]) {
          const float* gx = d.geom_xmat + ((size_t)w * m.ngeom + g2) * 9;
          const V3 ax = V3{gx[2], gx[5], gx[8]};
          const V3 bv = ax - nrm * dot(nrm, ax);
          const float bn = length(bv);
          if (bn >= 0.5f) {
            const V3 bu = bv * (1.0f / bn);
            const V3 u = cross(s6_rot(S), ax);
            const V3 up = u - nrm * dot(nrm, u);
            const V3 db = (up - bu * dot(bu, up)) * (1.0f / bn);
            const V3 dd = e < 2 ? (sgn * fri[0]) * db : (sgn * fri[1]) * cross(nrm, db);
            accc += dot(dd, vp);
          }
        }
      } else if (round) {
        // sphere / capsule pairs: sphere_sphere on the closest points of the
        // core segments (closest_segment_to_segment_points differentiated
        // as computed, clamps and choice included); n = dir / |dir|,
        // p = P1 + (r1 + dist / 2) n, frame = make_frame(n)
        const float* gxp = d.geom_xpos + (size_t)w * m.ngeom * 3;
        const float* gxm = d.geom_xmat + (size_t)w * m.ngeom * 9;
        const float r1 = m.geom_size[3 * g1];
        const float dist = d.contact_dist[c];
        const float L = dist + r1 + m.geom_size[3 * g2];
        if (L > 1e-6f) {
          auto matv = [&](int b, V3 P) {
            if (b == 0 || !m.body_dof_tree[b * nv + k]) return V3{0.f, 0.f, 0.f};
            return s6_lin(S) + cross(s6_rot(S), P - ld3(sc + 3 * m.body_rootid[b]));
          };
          const V3 c1 = ld3(gxp + 3 * g1), c2 = ld3(gxp + 3 * g2);
          V3 P1, P2, dP1, dP2;
          if (ty1 == mjhGEOM_CAPSULE && ty2 == mjhGEOM_CAPSULE) {
            const V3 s1 = V3{gxm[9 * g1 + 2], gxm[9 * g1 + 5], gxm[9 * g1 + 8]} * m.geom_size[3 * g1 + 1];
            const V3 s2 = V3{gxm[9 * g2 + 2], gxm[9 * g2 + 5], gxm[9 * g2 + 8]} * m.geom_size[3 * g2 + 1];
            segseg_tangent(c1 - s1, c1 + s1, c2 - s2, c2 + s2, matv(b1, c1 - s1), matv(b1, c1 + s1),
                           matv(b2, c2 - s2), matv(b2, c2 + s2), &P1, &P2, &dP1, &dP2);
          } else {
            dP1 = matv(b1, c1);
            dP2 = matv(b2, c2);
          }
          const V3 ddir = dP2 - dP1;
          const V3 dn = (ddir - nrm * dot(nrm, ddir)) * (1.0f / L);
          const float ddist = dot(nrm, ddir);
          dp = dP1 + dn * (r1 + 0.5f * dist) + nrm * (0.5f * ddist);
          // the row's distance terms (aref, D) through this distance
          accc += rlf[i] * ddist;
          const V3 ev = (-0.5f < nrm.y && nrm.y < 0.5f) ? V3{0.f, 1.f, 0.f} : V3{0.f, 0.f, 1.f};
          const V3 braw = ev - nrm * dot(nrm, ev);
          const float bl = length(braw);
          const V3 tt1 = braw * (1.0f / bl);
          const V3 dbraw = -1.0f * (dn * dot(nrm, ev) + nrm * dot(dn, ev));
          const V3 dt1 = (dbraw - tt1 * dot(tt1, dbraw)) * (1.0f / bl);
          const V3 dt2 = cross(dn, tt1) + cross(nrm, dt1);
          const V3 dd = dn + (e < 2 ? (sgn * fri[0]) * dt1 : (sgn * fri[1]) * dt2);
          accc += dot(dd, vp);
        }
      }
      // contact point: d . (omega_rel x dp)
      accc += dot(dp, cross(dir, om));
    }
    lq[k] += accc - tmp[k];
    lv[k] -= wv[k];
  }
  sync();
  float* lo = lam + (size_t)w * nx;
  const float* ga = gadd ? gadd + (size_t)w * nx : nullptr;
  for (int i = t.lane; i < nv; i += TEAM) {
    lo[i] = lq[i] + (ga ? ga[i] : 0.0f);
    lo[nv + i] = lv[i] + (ga ? ga[nv + i] : 0.0f);
  }
}

}  // namespace mjh
