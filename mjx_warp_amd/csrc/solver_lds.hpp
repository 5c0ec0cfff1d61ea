// LDS-resident constraint solver of one world (solver.py:26-1000).
//
// Same iteration as solver.hpp, but every mutable quantity lives in LDS
// for the whole solve: the mass matrix (or, for CG, its factor), the Newton
// Hessian and its factor, the dof vectors (qacc, Ma, grad, Mgrad, search,
// mv, ...) and the per-row scalars (Jaref, jv, D, quadratic coefficients,
// force, active).  The constraint Jacobian rows are copied into LDS when
// they fit (rows <= rj) and read from global memory otherwise (they are
// read-only during the solve).  Global memory is written once, at the end:
// qacc, qacc_warmstart, qfrc_constraint, efc_force / efc_Jaref / efc_active
// of the world's rows and the per-world solver scalars.  Barriers therefore
// never wait on outstanding global stores.
//
// A world with more rows than the LDS row capacity `rmax` is not solved
// here; the caller records it in world_flags and the global-memory solver
// (solver.hpp) handles it in a tail launch.
#pragma once

#include "solver.hpp"

namespace mjh {

constexpr int kFlagDeferred = 1;

// LDS floats per world: nv^2 (H; when wide H is in efc_h and this is the
// blocked Cholesky's panel buffer) + 14 nv (vectors) + 8 rmax (rows) + rj nv
// (staged Jacobian rows).  The mass
// matrix is read from global memory, where it is read-only during the solve.
__host__ __device__ inline int lds_solver_floats(int nv, int rmax, int rj, bool wide = false) {
  return (wide ? chol_blocked_floats(nv) : nv * nv) + 14 * nv + 8 * rmax + rj * nv;
}

typedef float floatx16 __attribute__((ext_vector_type(16)));

// Lower triangle of H = M + J' diag(dr) J (solver.py:250-276 _JTDAJ plus the
// mass matrix) on the matrix cores, for one world per 64-lane wavefront:
// H is cut into 32 x 32 tiles; tile (ti, tj), tj <= ti, accumulates
// v_mfma_f32_32x32x2f32 over the rows two at a time.  Lane l supplies row
// k + l / 32 of the world at column l % 32 of the tile: dr J[r][ti*32 + c]
// as the A operand, J[r][tj*32 + c] as the B operand; result entry v of lane
// l is tile row 8 (v / 4) + 4 (l / 32) + v % 4, tile column l % 32.  Every
// lane of the wavefront must call it (the matrix instruction uses all 64).
MJH_DEV void hessian_mfma(const float* M, const float* J, const float* D, const float* act, int nr, int nv,
                          float* H, int lane) {
  const int c = lane & 31, half = lane >> 5;
  const int nt = (nv + 31) >> 5;
  for (int ti = 0; ti < nt; ti++)
    for (int tj = 0; tj <= ti; tj++) {
      floatx16 acc = {};
      const int ci = ti * 32 + c, cj = tj * 32 + c;
      // the operands of eight row pairs are loaded before their matrix
      // instructions, so that their loads are in flight together
      for (int k0 = 0; k0 < nr; k0 += 16) {
        float a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int r = k0 + 2 * u + half;
          a[u] = 0.0f;
          b[u] = 0.0f;
          if (r < nr) {
            const float dr = D[r] * act[r];
            if (ci < nv) a[u] = dr * J[r * nv + ci];
            if (cj < nv) b[u] = J[r * nv + cj];
          }
        }
#pragma unroll
        for (int u = 0; u < 8; u++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[u], b[u], acc, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int i = ti * 32 + 8 * (v >> 2) + 4 * half + (v & 3);
        if (i < nv && cj < nv && cj <= i) H[i * nv + cj] = M[i * nv + cj] + acc[v];
      }
    }
}

template <int TEAM>
struct LdsSolver {
  const mjhModel& m;
  const mjhData& d;
  const Team<TEAM>& t;
  int w, nv, radr, nr;
  const float* M;  // qM of the world (global, read-only)
  float *H, *qacc, *Ma, *grad, *Mgrad, *search, *mv, *qs, *qas, *qc, *pg, *pmg, *x;
  float *Jaref, *jv, *D, *q0, *q1, *q2, *force, *act;
  const float* J;  // row-major [nr, nv], LDS or global
  float* pbuf = nullptr;  // wide: chol_blocked's panel buffer
  // Newton: the factor in H is current.  H = M + J' D(active) J depends on
  // the iterate only through the active set, so while no row changes
  // activity the previous factor is the one a rebuild would produce, bit
  // for bit.  act_bits: bit s set when row lane + TEAM s was active at the
  // last factorisation; a world with more than 32 rows per lane refactors
  // every iteration.
  bool fact_valid = false;
  unsigned act_bits = 0;

  MJH_DEV void mulM(float* res, const float* vec) const {
    for_each_tree_block(m, [&](int dof, int n) {
      for (int r = t.lane; r < n; r += TEAM) {
        float s = 0.0f;
        const float* row = M + (dof + r) * nv + dof;
        for (int c = 0; c < n; c++) s += row[c] * vec[dof + c];
        res[dof + r] = s;
      }
    });
  }

  MJH_DEV float cost_rows() {
    float c = 0.0f;
    for (int r = t.lane; r < nr; r += TEAM) {
      const float jar = Jaref[r];
      const bool a = jar < 0.0f;
      act[r] = a ? 1.0f : 0.0f;
      force[r] = a ? -1.0f * D[r] * jar : 0.0f;
      if (a) c += 0.5f * D[r] * jar * jar;
    }
    return team_sum<TEAM>(c);
  }

  // solver.py:75-173; returns (cost, gauss)
  MJH_DEV void update_constraint(float* cost, float* gauss) {
    const float c = cost_rows();
    sync();
    float g = 0.0f;
    for (int i = t.lane; i < nv; i += TEAM) {
      float s = 0.0f;
      for (int r = 0; r < nr; r++) s += J[r * nv + i] * force[r];
      qc[i] = s;
      g += 0.5f * (Ma[i] - qs[i]) * (qacc[i] - qas[i]);
    }
    g = team_sum<TEAM>(g);
    *gauss = g;
    *cost = c + g;
    sync();
  }

  // solver.py:176-312; returns grad_dot
  MJH_DEV float update_gradient() {
    float gd = 0.0f;
    for (int i = t.lane; i < nv; i += TEAM) {
      const float g = Ma[i] - qs[i] - qc[i];
      grad[i] = g;
      x[i] = g;
      gd += g * g;
    }
    gd = team_sum<TEAM>(gd);
    bool refactor = true;
    if (m.opt.solver == mjhSOL_NEWTON) {
      unsigned bits = 0;
      bool many = false;
      for (int r = t.lane, k = 0; r < nr; r += TEAM, k++) {
        if (k >= 32) many = true;
        else if (act[r] != 0.0f) bits |= 1u << k;
      }
      const float changed = team_sum<TEAM>(bits != act_bits || many ? 1.0f : 0.0f);
      refactor = !fact_valid || changed != 0.0f;
      act_bits = bits;
      fact_valid = true;
    }
    if (m.opt.solver == mjhSOL_NEWTON && !refactor) {
      sync();
      cho_solve_lds(H, nv, nv, x, t);
    } else if (m.opt.solver == mjhSOL_NEWTON && TEAM == 64) {
      hessian_mfma(M, J, D, act, nr, nv, H, t.lane);
      sync();
      if (pbuf) chol_blocked(H, nv, nv, pbuf, m, t);
      else chol_lds(H, nv, nv, m, t);
      cho_solve_lds(H, nv, nv, x, t);
    } else if (m.opt.solver == mjhSOL_NEWTON) {
      const int ntri = nv * (nv + 1) / 2;
      for (int p = t.lane; p < ntri; p += TEAM) {
        const int i = m.dof_tri_row[p], j = m.dof_tri_col[p];
        float h = M[i * nv + j];
        for (int r = 0; r < nr; r++) {
          const float dr = D[r] * act[r];
          if (dr == 0.0f) continue;
          h += J[r * nv + i] * J[r * nv + j] * dr;
        }
        H[i * nv + j] = h;
      }
      sync();
      chol_lds(H, nv, nv, m, t);
      cho_solve_lds(H, nv, nv, x, t);
    } else {
      // H holds the tree-block factor of M (qLD) for CG
      sync();
      for_each_tree_block(m, [&](int dof, int n) { cho_solve_lds(H + dof * nv + dof, n, nv, x + dof, t); });
    }
    for (int i = t.lane; i < nv; i += TEAM) Mgrad[i] = x[i];
    sync();
    return gd;
  }

  MJH_DEV V3f row_sum(float alpha) const {
    float a = 0.f, b = 0.f, c = 0.f;
    for (int r = t.lane; r < nr; r += TEAM) {
      if (Jaref[r] + alpha * jv[r] < 0.0f) {
        const V3f p = eval_pt(q0[r], q1[r], q2[r], alpha);
        a += p.a;
        b += p.b;
        c += p.c;
      }
    }
    return V3f{team_sum<TEAM>(a), team_sum<TEAM>(b), team_sum<TEAM>(c)};
  }

  // solver.py:339-664, returns alpha (prev_alpha is used when the loop does
  // not run)
  MJH_DEV float linesearch_iterative(float g0, float g1, float g2, float search_dot, float prev_alpha) {
    const float gtol = m.opt.tolerance * m.opt.ls_tolerance * sqrtf(search_dot) * m.meaninertia * (float)max(1, nv);
    float pa = 0.f, pb = 0.f, pc = 0.f;
    for (int r = t.lane; r < nr; r += TEAM)
      if (Jaref[r] < 0.0f) {
        pa += q0[r];
        pb += q1[r];
        pc += 2.0f * q2[r];
      }
    const V3f p0 = V3f{g0 + team_sum<TEAM>(pa), g1 + team_sum<TEAM>(pb), 2.0f * g2 + team_sum<TEAM>(pc)};
    float lo_alpha = -safe_div(p0.b, p0.c);
    V3f lo = eval_pt(g0, g1, g2, lo_alpha);
    {
      const V3f s = row_sum(lo_alpha);
      lo = V3f{lo.a + s.a, lo.b + s.b, lo.c + s.c};
    }
    const bool lo_less = lo.b < p0.b;
    V3f hi = lo_less ? p0 : lo;
    float hi_alpha = lo_less ? 0.0f : lo_alpha;
    if (!lo_less) {
      lo = p0;
      lo_alpha = 0.0f;
    }
    float alpha = prev_alpha;
    bool done = false;
    for (int it = 0; it < m.opt.ls_iterations && !done; it++) {
      const float la = lo_alpha - safe_div(lo.b, lo.c);
      const float ha = hi_alpha - safe_div(hi.b, hi.c);
      const float ma = 0.5f * (lo_alpha + hi_alpha);
      float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f, s5 = 0.f, s6 = 0.f, s7 = 0.f, s8 = 0.f;
      for (int r = t.lane; r < nr; r += TEAM) {
        const float jar = Jaref[r], v = jv[r];
        const float c0 = q0[r], c1 = q1[r], c2 = q2[r];
        if (jar + la * v < 0.0f) {
          const V3f p = eval_pt(c0, c1, c2, la);
          s0 += p.a; s1 += p.b; s2 += p.c;
        }
        if (jar + ha * v < 0.0f) {
          const V3f p = eval_pt(c0, c1, c2, ha);
          s3 += p.a; s4 += p.b; s5 += p.c;
        }
        if (jar + ma * v < 0.0f) {
          const V3f p = eval_pt(c0, c1, c2, ma);
          s6 += p.a; s7 += p.b; s8 += p.c;
        }
      }
      V3f ln = eval_pt(g0, g1, g2, la), hn = eval_pt(g0, g1, g2, ha), mn = eval_pt(g0, g1, g2, ma);
      ln = V3f{ln.a + team_sum<TEAM>(s0), ln.b + team_sum<TEAM>(s1), ln.c + team_sum<TEAM>(s2)};
      hn = V3f{hn.a + team_sum<TEAM>(s3), hn.b + team_sum<TEAM>(s4), hn.c + team_sum<TEAM>(s5)};
      mn = V3f{mn.a + team_sum<TEAM>(s6), mn.b + team_sum<TEAM>(s7), mn.c + team_sum<TEAM>(s8)};
      bool sl = false, sh = false;
      if (in_bracket(lo, ln)) { lo = ln; lo_alpha = la; sl = true; }
      if (in_bracket(lo, mn)) { lo = mn; lo_alpha = ma; sl = true; }
      if (in_bracket(lo, hn)) { lo = hn; lo_alpha = ha; sl = true; }
      if (in_bracket(hi, hn)) { hi = hn; hi_alpha = ha; sh = true; }
      if (in_bracket(hi, mn)) { hi = mn; hi_alpha = ma; sh = true; }
      if (in_bracket(hi, ln)) { hi = ln; hi_alpha = la; sh = true; }
      done = (!sl && !sh) || (lo.b < 0.0f && lo.b > -gtol) || (hi.b > 0.0f && hi.b < gtol);
      const bool improved = lo.a < p0.a || hi.a < p0.a;
      alpha = 0.0f;
      if (improved) alpha = lo.a < hi.a ? lo_alpha : hi_alpha;
    }
    return alpha;
  }

  // solver.py:667-734
  MJH_DEV float linesearch_parallel(float g0, float g1, float g2) {
    int best = 0;
    float best_cost = 0.0f;
    for (int k = 0; k < m.nlsp; k++) {
      const float a = m.alpha_candidate[k];
      float s0 = 0.f, s1 = 0.f, s2 = 0.f;
      for (int r = t.lane; r < nr; r += TEAM)
        if (Jaref[r] + a * jv[r] < 0.0f) {
          s0 += q0[r];
          s1 += q1[r];
          s2 += q2[r];
        }
      const float t0 = g0 + team_sum<TEAM>(s0), t1 = g1 + team_sum<TEAM>(s1), t2 = g2 + team_sum<TEAM>(s2);
      const float c = a * a * t2 + a * t1 + t0;
      if (t.lane == 0) {
        d.efc_cost_candidate[w * m.nlsp + k] = c;
        float* qt = d.efc_quad_total_candidate + 3 * (w * m.nlsp + k);
        qt[0] = t0;
        qt[1] = t1;
        qt[2] = t2;
      }
      if (k == 0 || c < best_cost) {
        best = k;
        best_cost = c;
      }
    }
    return m.alpha_candidate[best];
  }

  // solver.py:737-869; returns alpha
  MJH_DEV float linesearch(float gauss, float search_dot, float prev_alpha) {
    mulM(mv, search);
    for (int r = t.lane; r < nr; r += TEAM) {
      float s = 0.0f;
      const float* Jr = J + r * nv;
      for (int i = 0; i < nv; i++) s += Jr[i] * search[i];
      jv[r] = s;
      const float jar = Jaref[r], dd = D[r];
      q0[r] = 0.5f * jar * jar * dd;
      q1[r] = s * jar * dd;
      q2[r] = 0.5f * s * s * dd;
    }
    sync();
    float g0 = 0.f, g1 = 0.f, g2 = 0.f;
    for (int i = t.lane; i < nv; i += TEAM) {
      g0 += gauss / (float)nv;
      g1 += search[i] * (Ma[i] - qs[i]);
      g2 += 0.5f * search[i] * mv[i];
    }
    g0 = team_sum<TEAM>(g0);
    g1 = team_sum<TEAM>(g1);
    g2 = team_sum<TEAM>(g2);
    const float alpha = m.opt.ls_parallel ? linesearch_parallel(g0, g1, g2)
                                          : linesearch_iterative(g0, g1, g2, search_dot, prev_alpha);
    for (int i = t.lane; i < nv; i += TEAM) {
      qacc[i] += alpha * search[i];
      Ma[i] += alpha * mv[i];
    }
    for (int r = t.lane; r < nr; r += TEAM) Jaref[r] += alpha * jv[r];
    sync();
    return alpha;
  }

  MJH_DEV float search_update(float beta, bool with_previous) {
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
    sync();
    return sd;
  }
};

// Returns false (and leaves the world untouched) when the world's rows do
// not fit the LDS row capacity.
template <int TEAM>
MJH_DEV bool solve_lds(const mjhModel& m, const mjhData& d, float* lds, int rmax, int rj, bool wide,
                       const Team<TEAM>& t) {
  if (!t.valid) return true;
  const int w = t.world, nv = m.nv;
  const int radr = d.world_efc[2 * w], nr = d.world_efc[2 * w + 1];
  if (nr > rmax) return false;
  LdsSolver<TEAM> s{m, d, t, w, nv, radr, nr};
  float* p = lds;
  s.M = d.qMd + (size_t)w * nv * nv;
  // wide: H in the world's efc_h (CG: its tree-block factor in qLDd)
  if (wide) {
    s.H = m.opt.solver == mjhSOL_CG ? d.qLDd + (size_t)w * nv * nv : d.efc_h + (size_t)w * nv * nv;
    s.pbuf = p;
    p += chol_blocked_floats(nv);
  } else {
    s.H = p;
    p += nv * nv;
  }
  s.qacc = p; p += nv;
  s.Ma = p; p += nv;
  s.grad = p; p += nv;
  s.Mgrad = p; p += nv;
  s.search = p; p += nv;
  s.mv = p; p += nv;
  s.qs = p; p += nv;
  s.qas = p; p += nv;
  s.qc = p; p += nv;
  s.pg = p; p += nv;
  s.pmg = p; p += nv;
  s.x = p; p += 3 * nv;
  s.Jaref = p; p += rmax;
  s.jv = p; p += rmax;
  s.D = p; p += rmax;
  s.q0 = p; p += rmax;
  s.q1 = p; p += rmax;
  s.q2 = p; p += rmax;
  s.force = p; p += rmax;
  s.act = p; p += rmax;
  const float* Jg = d.efc_J + (size_t)radr * nv;
  if (nr <= rj) {
    float* Js = p;
    for (int k = t.lane; k < nr * nv; k += TEAM) Js[k] = Jg[k];
    s.J = Js;
  } else {
    s.J = Jg;
  }
  // stage the world's inputs
  const size_t o = (size_t)w * nv;
  if (m.opt.solver == mjhSOL_CG && !wide) {
    const float* qLD = d.qLDd + o * nv;
    for (int k = t.lane; k < nv * nv; k += TEAM) s.H[k] = qLD[k];
  }
  for (int i = t.lane; i < nv; i += TEAM) {
    s.qacc[i] = d.qacc_warmstart[o + i];  // warm start (solver.py:973)
    s.qs[i] = d.qfrc_smooth[o + i];
    s.qas[i] = d.qacc_smooth[o + i];
  }
  for (int r = t.lane; r < nr; r += TEAM) s.D[r] = d.efc_D[radr + r];
  sync();
  // _create_context (solver.py:26-72)
  for (int r = t.lane; r < nr; r += TEAM) {
    const float* Jr = s.J + r * nv;
    const float arn = d.efc_aref[radr + r] / (float)nv;
    float a = 0.0f;
    for (int i = 0; i < nv; i++) a += Jr[i] * s.qacc[i] - arn;
    s.Jaref[r] = a;
  }
  s.mulM(s.Ma, s.qacc);
  sync();
  float cost = __int_as_float(0x7f800000), prev_cost, gauss;
  prev_cost = cost;
  s.update_constraint(&cost, &gauss);
  float grad_dot = s.update_gradient();
  float search_dot = s.search_update(0.0f, false);
  const int iters = m.opt.iterations;
  bool done = false;
  float alpha = d.efc_alpha[w], beta = d.efc_beta[w], beta_num = d.efc_beta_num[w], beta_den = d.efc_beta_den[w];
  for (int it = 0; it < iters; it++) {
    if (iters > 1 && done) break;
    alpha = s.linesearch(gauss, search_dot, alpha);
    if (m.opt.solver == mjhSOL_CG) {
      for (int i = t.lane; i < nv; i += TEAM) {
        s.pg[i] = s.grad[i];
        s.pmg[i] = s.Mgrad[i];
      }
      sync();
    }
    prev_cost = cost;
    s.update_constraint(&cost, &gauss);
    grad_dot = s.update_gradient();
    if (m.opt.solver == mjhSOL_CG) {
      float num = 0.0f, den = 0.0f;
      for (int i = t.lane; i < nv; i += TEAM) {
        num += s.grad[i] * (s.Mgrad[i] - s.pmg[i]);
        den += s.pg[i] * s.pmg[i];
      }
      beta_num = team_sum<TEAM>(num);
      beta_den = team_sum<TEAM>(den);
      beta = fmaxf(0.0f, beta_num / fmaxf(MJH_MINVAL, beta_den));
    }
    search_dot = s.search_update(beta, true);
    const float scale = m.meaninertia * (float)max(1, nv);
    done = ((prev_cost - cost) / scale < m.opt.tolerance) || (sqrtf(grad_dot) / scale < m.opt.tolerance);
  }
  // write back
  for (int i = t.lane; i < nv; i += TEAM) {
    d.qacc[o + i] = s.qacc[i];
    d.qacc_warmstart[o + i] = s.qacc[i];
    d.qfrc_constraint[o + i] = s.qc[i];
  }
  for (int r = t.lane; r < nr; r += TEAM) {
    d.efc_force[radr + r] = s.force[r];
    d.efc_Jaref[radr + r] = s.Jaref[r];
    d.efc_active[radr + r] = s.act[r] != 0.0f;
  }
  if (t.lane == 0) {
    d.efc_cost[w] = cost;
    d.efc_prev_cost[w] = prev_cost;
    d.efc_gauss[w] = gauss;
    d.efc_grad_dot[w] = grad_dot;
    d.efc_search_dot[w] = search_dot;
    d.efc_alpha[w] = alpha;
    d.efc_beta[w] = beta;
    d.efc_beta_num[w] = beta_num;
    d.efc_beta_den[w] = beta_den;
    d.efc_done[w] = done;
    d.efc_solver_niter[w] = 0;
  }
  return true;
}

}  // namespace mjh
