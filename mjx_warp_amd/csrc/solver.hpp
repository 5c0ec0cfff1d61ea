// Primal CG / Newton constraint solver of one world (solver.py:26-1000).
//
// A team owns its world's rows (the contiguous efc block recorded in
// world_efc) and dofs.  Per-world scalars (cost, gauss, bracket points ...)
// are reduced across the team with shuffles so every lane holds them; the
// reference's global atomics become these deterministic team reductions.
// The Newton Hessian H = M + J' D J (lower triangle, solver.py:176-312) is
// assembled and factored in LDS and is not written back to efc_h.
#pragma once

#include "forward.hpp"

namespace mjh {

struct V3f {
  float a, b, c;
};
MJH_DEV V3f eval_pt(float q0, float q1, float q2, float alpha) {
  return V3f{alpha * alpha * q2 + alpha * q1 + q0, 2.0f * alpha * q2 + q1, 2.0f * q2};
}
MJH_DEV float safe_div(float x, float y) { return x / (y != 0.0f ? y : MJH_MINVAL); }
MJH_DEV bool in_bracket(V3f x, V3f y) { return (x.b < y.b && y.b < 0.0f) || (x.b > y.b && y.b > 0.0f); }

template <int TEAM>
struct SolverWorld {
  const mjhModel& m;
  const mjhData& d;
  const Team<TEAM>& t;
  int w, nv, radr, nr;
  float* qacc;
  float* Ma;
  float* grad;
  float* Mgrad;
  float* search;
  float* mv;
  const float* qs;
  const float* qacc_smooth;
  float* qc;
  const float* qM;
  float* lds;
  bool wide;  // H in efc_h instead of LDS

  MJH_DEV float* J(int r) const { return d.efc_J + (size_t)(radr + r) * nv; }

  // solver.py:75-173 _update_constraint
  MJH_DEV void update_constraint() {
    float c = 0.0f;
    for (int r = t.lane; r < nr; r += TEAM) {
      const int e = radr + r;
      const float jar = d.efc_Jaref[e], D = d.efc_D[e];
      const int active = jar < 0.0f;
      d.efc_active[e] = active;
      if (active) {
        d.efc_force[e] = -1.0f * D * jar;
        c += 0.5f * D * jar * jar;
      } else {
        d.efc_force[e] = 0.0f;
      }
    }
    c = team_sum<TEAM>(c);
    sync();
    float g = 0.0f;
    for (int i = t.lane; i < nv; i += TEAM) {
      float s = 0.0f;
      for (int r = 0; r < nr; r++) s += J(r)[i] * d.efc_force[radr + r];
      qc[i] = s;
      g += 0.5f * (Ma[i] - qs[i]) * (qacc[i] - qacc_smooth[i]);
    }
    g = team_sum<TEAM>(g);
    if (t.lane == 0) {
      d.efc_prev_cost[w] = d.efc_cost[w];
      d.efc_gauss[w] = g;
      d.efc_cost[w] = c + g;
    }
    sync();
  }

  // solver.py:176-312 _update_gradient
  MJH_DEV void update_gradient() {
    float gd = 0.0f;
    for (int i = t.lane; i < nv; i += TEAM) {
      const float g = Ma[i] - qs[i] - qc[i];
      grad[i] = g;
      gd += g * g;
    }
    gd = team_sum<TEAM>(gd);
    if (t.lane == 0) d.efc_grad_dot[w] = gd;
    sync();
    if (m.opt.solver == mjhSOL_CG) {
      solve_blocks(m, d.qLDd + (size_t)w * nv * nv, Mgrad, grad, lds, t);
    } else {
      float* H = wide ? d.efc_h + (size_t)w * nv * nv : lds;
      float* x = wide ? lds : lds + nv * nv;
      float* pbuf = wide ? lds + nv : nullptr;
      const int ntri = nv * (nv + 1) / 2;
      for (int p = t.lane; p < ntri; p += TEAM) {
        const int i = m.dof_tri_row[p], j = m.dof_tri_col[p];
        float h = qM[i * nv + j];
        for (int r = 0; r < nr; r++) {
          const int e = radr + r;
          const float D = d.efc_D[e];
          if (D == 0.0f || d.efc_active[e] == 0) continue;
          const float* Jr = J(r);
          h += Jr[i] * Jr[j] * D;
        }
        H[i * nv + j] = h;
      }
      for (int i = t.lane; i < nv; i += TEAM) x[i] = grad[i];
      sync();
      if (pbuf) chol_blocked(H, nv, nv, pbuf, m, t);
      else chol_lds(H, nv, nv, m, t);
      cho_solve_lds(H, nv, nv, x, t);
      for (int i = t.lane; i < nv; i += TEAM) Mgrad[i] = x[i];
    }
    sync();
  }

  // sum over rows active at alpha (Jaref + alpha jv < 0) of eval_pt(quad, alpha)
  MJH_DEV V3f row_points(float alpha) {
    float a = 0.f, b = 0.f, c = 0.f;
    for (int r = t.lane; r < nr; r += TEAM) {
      const int e = radr + r;
      if (d.efc_Jaref[e] + alpha * d.efc_jv[e] < 0.0f) {
        const float* q = d.efc_quad + 3 * e;
        const V3f p = eval_pt(q[0], q[1], q[2], alpha);
        a += p.a;
        b += p.b;
        c += p.c;
      }
    }
    return V3f{team_sum<TEAM>(a), team_sum<TEAM>(b), team_sum<TEAM>(c)};
  }

  // solver.py:339-664 _linesearch_iterative
  MJH_DEV void linesearch_iterative() {
    const float* qg = d.efc_quad_gauss + 3 * w;
    const float qg0 = qg[0], qg1 = qg[1], qg2 = qg[2];
    const float snorm = sqrtf(d.efc_search_dot[w]);
    const float scale = m.meaninertia * (float)max(1, nv);
    const float gtol = m.opt.tolerance * m.opt.ls_tolerance * snorm * scale;
    // p0: cost at alpha = 0 over rows active at 0 (Jaref < 0)
    float pa = 0.f, pb = 0.f, pc = 0.f;
    for (int r = t.lane; r < nr; r += TEAM) {
      const int e = radr + r;
      if (d.efc_Jaref[e] < 0.0f) {
        const float* q = d.efc_quad + 3 * e;
        pa += q[0];
        pb += q[1];
        pc += 2.0f * q[2];
      }
    }
    const V3f p0 = V3f{qg0 + team_sum<TEAM>(pa), qg1 + team_sum<TEAM>(pb), 2.0f * qg2 + team_sum<TEAM>(pc)};
    // lo at the Newton step from alpha = 0
    float lo_alpha = -safe_div(p0.b, p0.c);
    V3f lo = eval_pt(qg0, qg1, qg2, lo_alpha);
    {
      const V3f s = row_points(lo_alpha);
      lo = V3f{lo.a + s.a, lo.b + s.b, lo.c + s.c};
    }
    // bracket
    const bool lo_less = lo.b < p0.b;
    V3f hi = lo_less ? p0 : lo;
    float hi_alpha = lo_less ? 0.0f : lo_alpha;
    lo = lo_less ? lo : p0;
    lo_alpha = lo_less ? lo_alpha : 0.0f;
    bool done = false;
    float alpha = d.efc_alpha[w];
    V3f lon = {}, hin = {}, mid = {};
    float lon_alpha = 0.f, hin_alpha = 0.f, mid_alpha = 0.f;
    for (int it = 0; it < m.opt.ls_iterations; it++) {
      if (done) break;
      lon_alpha = lo_alpha - safe_div(lo.b, lo.c);
      hin_alpha = hi_alpha - safe_div(hi.b, hi.c);
      mid_alpha = 0.5f * (lo_alpha + hi_alpha);
      lon = eval_pt(qg0, qg1, qg2, lon_alpha);
      hin = eval_pt(qg0, qg1, qg2, hin_alpha);
      mid = eval_pt(qg0, qg1, qg2, mid_alpha);
      float s[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      for (int r = t.lane; r < nr; r += TEAM) {
        const int e = radr + r;
        const float* q = d.efc_quad + 3 * e;
        const float jar = d.efc_Jaref[e], jv = d.efc_jv[e];
        if (jar + lon_alpha * jv < 0.0f) {
          const V3f p = eval_pt(q[0], q[1], q[2], lon_alpha);
          s[0] += p.a; s[1] += p.b; s[2] += p.c;
        }
        if (jar + hin_alpha * jv < 0.0f) {
          const V3f p = eval_pt(q[0], q[1], q[2], hin_alpha);
          s[3] += p.a; s[4] += p.b; s[5] += p.c;
        }
        if (jar + mid_alpha * jv < 0.0f) {
          const V3f p = eval_pt(q[0], q[1], q[2], mid_alpha);
          s[6] += p.a; s[7] += p.b; s[8] += p.c;
        }
      }
#pragma unroll
      for (int k = 0; k < 9; k++) s[k] = team_sum<TEAM>(s[k]);
      lon = V3f{lon.a + s[0], lon.b + s[1], lon.c + s[2]};
      hin = V3f{hin.a + s[3], hin.b + s[4], hin.c + s[5]};
      mid = V3f{mid.a + s[6], mid.b + s[7], mid.c + s[8]};
      // swap lo (solver.py:573-585)
      const bool sl1 = in_bracket(lo, lon);
      if (sl1) { lo = lon; lo_alpha = lon_alpha; }
      const bool sl2 = in_bracket(lo, mid);
      if (sl2) { lo = mid; lo_alpha = mid_alpha; }
      const bool sl3 = in_bracket(lo, hin);
      if (sl3) { lo = hin; lo_alpha = hin_alpha; }
      const bool swap_lo = sl1 || sl2 || sl3;
      // swap hi (solver.py:587-599)
      const bool sh1 = in_bracket(hi, hin);
      if (sh1) { hi = hin; hi_alpha = hin_alpha; }
      const bool sh2 = in_bracket(hi, mid);
      if (sh2) { hi = mid; hi_alpha = mid_alpha; }
      const bool sh3 = in_bracket(hi, lon);
      if (sh3) { hi = lon; hi_alpha = lon_alpha; }
      const bool swap_hi = sh1 || sh2 || sh3;
      done = (!swap_lo && !swap_hi) || (lo.b < 0.0f && lo.b > -gtol) || (hi.b > 0.0f && hi.b < gtol);
      const bool improved = lo.a < p0.a || hi.a < p0.a;
      const bool lo_better = lo.a < hi.a;
      alpha = 0.0f;
      if (improved && lo_better) alpha = lo_alpha;
      if (improved && !lo_better) alpha = hi_alpha;
    }
    if (t.lane == 0) {
      d.efc_gtol[w] = gtol;
      d.efc_ls_done[w] = done;
      st3(d.efc_p0 + 3 * w, V3{p0.a, p0.b, p0.c});
      st3(d.efc_lo + 3 * w, V3{lo.a, lo.b, lo.c});
      d.efc_lo_alpha[w] = lo_alpha;
      st3(d.efc_hi + 3 * w, V3{hi.a, hi.b, hi.c});
      d.efc_hi_alpha[w] = hi_alpha;
      st3(d.efc_lo_next + 3 * w, V3{lon.a, lon.b, lon.c});
      d.efc_lo_next_alpha[w] = lon_alpha;
      st3(d.efc_hi_next + 3 * w, V3{hin.a, hin.b, hin.c});
      d.efc_hi_next_alpha[w] = hin_alpha;
      st3(d.efc_mid + 3 * w, V3{mid.a, mid.b, mid.c});
      d.efc_mid_alpha[w] = mid_alpha;
      d.efc_alpha[w] = alpha;
    }
    sync();
  }

  // solver.py:667-734 _linesearch_parallel
  MJH_DEV void linesearch_parallel() {
    const float* qg = d.efc_quad_gauss + 3 * w;
    int best = 0;
    float best_cost = 0.0f;
    for (int k = 0; k < m.nlsp; k++) {
      const float a = m.alpha_candidate[k];
      float s0 = 0.f, s1 = 0.f, s2 = 0.f;
      for (int r = t.lane; r < nr; r += TEAM) {
        const int e = radr + r;
        if (d.efc_Jaref[e] + a * d.efc_jv[e] < 0.0f) {
          const float* q = d.efc_quad + 3 * e;
          s0 += q[0];
          s1 += q[1];
          s2 += q[2];
        }
      }
      const float t0 = qg[0] + team_sum<TEAM>(s0);
      const float t1 = qg[1] + team_sum<TEAM>(s1);
      const float t2 = qg[2] + team_sum<TEAM>(s2);
      const float cst = a * a * t2 + a * t1 + t0;
      if (t.lane == 0) {
        d.efc_cost_candidate[w * m.nlsp + k] = cst;
        float* qt = d.efc_quad_total_candidate + 3 * (w * m.nlsp + k);
        qt[0] = t0;
        qt[1] = t1;
        qt[2] = t2;
      }
      if (k == 0 || cst < best_cost) {
        best = k;
        best_cost = cst;
      }
    }
    if (t.lane == 0) d.efc_alpha[w] = m.alpha_candidate[best];
    sync();
  }

  // solver.py:737-869 _linesearch
  MJH_DEV void linesearch() {
    mul_m(m, qM, mv, search, t);
    for (int r = t.lane; r < nr; r += TEAM) {
      const float* Jr = J(r);
      float s = 0.0f;
      for (int i = 0; i < nv; i++) s += Jr[i] * search[i];
      d.efc_jv[radr + r] = s;
    }
    sync();
    float g1 = 0.0f, g2 = 0.0f;
    const float gauss = d.efc_gauss[w];
    float g0 = 0.0f;
    for (int i = t.lane; i < nv; i += TEAM) {
      g0 += gauss / (float)nv;
      g1 += search[i] * (Ma[i] - qs[i]);
      g2 += 0.5f * search[i] * mv[i];
    }
    g0 = team_sum<TEAM>(g0);
    g1 = team_sum<TEAM>(g1);
    g2 = team_sum<TEAM>(g2);
    if (t.lane == 0) st3(d.efc_quad_gauss + 3 * w, V3{g0, g1, g2});
    for (int r = t.lane; r < nr; r += TEAM) {
      const int e = radr + r;
      const float jar = d.efc_Jaref[e], jv = d.efc_jv[e], D = d.efc_D[e];
      float* q = d.efc_quad + 3 * e;
      q[0] = 0.5f * jar * jar * D;
      q[1] = jv * jar * D;
      q[2] = 0.5f * jv * jv * D;
    }
    sync();
    if (m.opt.ls_parallel) linesearch_parallel();
    else linesearch_iterative();
    const float alpha = d.efc_alpha[w];
    for (int i = t.lane; i < nv; i += TEAM) {
      qacc[i] += alpha * search[i];
      Ma[i] += alpha * mv[i];
    }
    for (int r = t.lane; r < nr; r += TEAM) d.efc_Jaref[radr + r] += alpha * d.efc_jv[radr + r];
    sync();
  }

  MJH_DEV void search_update(float beta, bool with_previous) {
    // with_previous is false for the first direction (solver.py:50-55),
    // which does not read the previous search vector
    float sd = 0.0f;
    for (int i = t.lane; i < nv; i += TEAM) {
      float s = -1.0f * Mgrad[i];
      if (with_previous && m.opt.solver == mjhSOL_CG) s += beta * search[i];
      search[i] = s;
      sd += s * s;
    }
    sd = team_sum<TEAM>(sd);
    if (t.lane == 0) d.efc_search_dot[w] = sd;
    sync();
  }
};

// solver.py:872-1000 solve.  lds: the largest tree block's n^2 + n floats
// and nv*nv + nv (H and the right-hand side), or when wide (H in efc_h) nv
// and chol_blocked's panel buffer.
template <int TEAM>
MJH_DEV void solve(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t, bool wide = false) {
  if (!t.valid) return;
  const int w = t.world, nv = m.nv;
  SolverWorld<TEAM> s{m, d, t, w, nv, d.world_efc[2 * w], d.world_efc[2 * w + 1],
                      d.qacc + (size_t)w * nv, d.efc_Ma + (size_t)w * nv, d.efc_grad + (size_t)w * nv,
                      d.efc_Mgrad + (size_t)w * nv, d.efc_search + (size_t)w * nv, d.efc_mv + (size_t)w * nv,
                      d.qfrc_smooth + (size_t)w * nv, d.qacc_smooth + (size_t)w * nv,
                      d.qfrc_constraint + (size_t)w * nv, d.qMd + (size_t)w * nv * nv, lds, wide};
  const float* ws = d.qacc_warmstart + (size_t)w * nv;
  for (int i = t.lane; i < nv; i += TEAM) s.qacc[i] = ws[i];
  sync();
  // _create_context (solver.py:26-72)
  if (t.lane == 0) {
    d.efc_cost[w] = __int_as_float(0x7f800000);
    d.efc_solver_niter[w] = 0;
    d.efc_done[w] = 0;
    d.efc_search_dot[w] = 0.0f;
  }
  for (int r = t.lane; r < s.nr; r += TEAM) {
    const int e = s.radr + r;
    const float* Jr = s.J(r);
    const float arn = d.efc_aref[e] / (float)nv;
    float a = 0.0f;
    for (int i = 0; i < nv; i++) a += Jr[i] * s.qacc[i] - arn;
    d.efc_Jaref[e] = a;
  }
  mul_m(m, s.qM, s.Ma, s.qacc, t);
  sync();
  s.update_constraint();
  s.update_gradient();
  s.search_update(0.0f, false);
  const int iters = m.opt.iterations;
  bool done = false;
  for (int it = 0; it < iters; it++) {
    // worlds that converged skip further iterations when iterations > 1
    if (iters > 1 && done) break;
    s.linesearch();
    if (m.opt.solver == mjhSOL_CG) {
      for (int i = t.lane; i < nv; i += TEAM) {
        d.efc_prev_grad[w * nv + i] = s.grad[i];
        d.efc_prev_Mgrad[w * nv + i] = s.Mgrad[i];
      }
      sync();
    }
    s.update_constraint();
    s.update_gradient();
    float beta = 0.0f;
    if (m.opt.solver == mjhSOL_CG) {
      float num = 0.0f, den = 0.0f;
      for (int i = t.lane; i < nv; i += TEAM) {
        const float pm = d.efc_prev_Mgrad[w * nv + i];
        num += s.grad[i] * (s.Mgrad[i] - pm);
        den += d.efc_prev_grad[w * nv + i] * pm;
      }
      num = team_sum<TEAM>(num);
      den = team_sum<TEAM>(den);
      beta = fmaxf(0.0f, num / fmaxf(MJH_MINVAL, den));
      if (t.lane == 0) {
        d.efc_beta_num[w] = num;
        d.efc_beta_den[w] = den;
        d.efc_beta[w] = beta;
      }
    }
    s.search_update(beta, true);
    const float scale = m.meaninertia * (float)max(1, nv);
    const float improvement = (d.efc_prev_cost[w] - d.efc_cost[w]) / scale;
    const float gradient = sqrtf(d.efc_grad_dot[w]) / scale;
    done = (improvement < m.opt.tolerance) || (gradient < m.opt.tolerance);
    if (t.lane == 0) d.efc_done[w] = done;
  }
  sync();
  float* wso = d.qacc_warmstart + (size_t)w * nv;
  for (int i = t.lane; i < nv; i += TEAM) wso[i] = s.qacc[i];
  sync();
}

}  // namespace mjh
