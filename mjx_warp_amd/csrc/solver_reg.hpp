// Register-resident constraint solver for worlds with nv <= 32
// (solver.py:26-1000: _create_context, _update_constraint, _update_gradient,
// _linesearch_iterative / _linesearch_parallel, CG beta, _done).
//
// Mapping: a team of 32 lanes per world, two worlds per wavefront.  Lane i
// owns dof i -- its entries of qacc, Ma, grad, Mgrad, search, mv, the mass
// matrix row (M is symmetric, so the row is read as a coalesced column) and
// row i of the Newton Hessian and of its Cholesky factor, all in VGPRs.  Lane
// r also owns constraint rows r, r + 32, ..., r + 32 (RS - 1) (Jaref, jv, D,
// force and the line-search quadratic): RS row sets, of which the first
// ceil(nr / 32) are live.  The world's Jacobian rows are staged in LDS with
// the odd row stride NVC + 1, so both row-wise (lane = row) and column-wise
// (lane = dof) reads are bank-conflict free; rows past the staged count are
// read from global memory, where they are read-only during the solve.
//
// Cross-lane sums are DPP butterflies inside 16-lane rows plus one swizzle
// across the two rows of a team; every lane of the team ends with the same
// bits, so per-world decisions (line-search bracket, done) stay uniform.
// Vectors that every lane needs whole are broadcast through a 32-float LDS
// buffer: LDS instructions of one wavefront execute in issue order, so a
// write followed by reads of the same wavefront needs no barrier, only a
// compiler fence.
//
// The dof dimension is padded to NVC (a multiple of 4, <= 32) with identity
// rows of M / H and zero vectors, so every dof loop has a compile-time trip
// count and no guards; padded entries never change a sum.
#pragma once

#include "linalg_reg.hpp"
#include "solver_lds.hpp"

namespace mjh {

constexpr int kRegSets = 8;                // most row sets per lane
constexpr int kRegRows = 32 * kRegSets;    // row capacity of the register solver

// LDS floats per world: broadcast buffers (vector 32, column 32, rows
// kRegRows) and rj staged Jacobian rows of stride nvc + 1 (nvc: nv rounded
// up to 4), rounded up to a float4.
__host__ __device__ inline int lds_reg_floats(int rj, int nvc) { return (64 + kRegRows + rj * (nvc + 1) + 3) & ~3; }

template <int NVC, int RS>
struct RegSolver {
  static_assert(NVC % 4 == 0 && NVC <= kRegNvMax, "padded dof count");
  static_assert(RS >= 1 && RS <= kRegSets, "row sets");
  const mjhModel& m;
  const mjhData& d;
  int w, nv, radr, nr, lane, rj;
  float* vbuf;  // [32] dof-vector broadcast
  float* fbuf;  // [32 * RS] row-scalar broadcast
  float* Jl;    // [rj][JS] staged Jacobian rows
  static constexpr int JS = NVC + 1;
  const float* M;  // qM of the world (global, read-only during the solve)
  // dof lane state
  RegChol<NVC> ch;  // Hessian row and its Cholesky factor (CG: factor of M)
  float qacc, Ma, grad, Mgrad, search, mv, qs, qas, qc;
  // row lane state: rows lane + 32 s, s < RS
  float Jaref[RS], jv[RS], D[RS], force[RS], q0[RS], q1[RS], q2[RS];

  int blk;   // CG: first dof of the lane's tree block
  int nset;  // live row sets, ceil(nr / 32) but at least 1 (team-uniform); set 0
             // is always live, which the loops state at compile time
  // the factor in ch is current: Newton's H = M + J' D(active) J depends on
  // the iterate only through the active set, so while no row changes
  // activity the previous factor is the one a rebuild would produce
  bool fact_valid;
  unsigned act_prev;  // bit s: row lane + 32 s was active at the last factorisation

  MJH_DEV int row(int s) const { return lane + 32 * s; }
  // global Jacobian row r (rows past the LDS capacity)
  MJH_DEV gfloat* Jg(int r) const { return opaque(d.efc_J + (size_t)radr * nv) + r * nv; }
  MJH_DEV int nlds() const { return nr < rj ? nr : rj; }

  MJH_DEV void put_vec(float v) {
    wave_fence();
    vbuf[lane] = v;
    wave_fence();
  }
  MJH_DEV float4 vec4(int q) const { return reinterpret_cast<const float4*>(vbuf)[q]; }

  // M[lane] . v for the vector last passed to put_vec (M symmetric: the
  // lane reads column `lane`, coalesced across the team)
  MJH_DEV float mrow_dot() const {
    float s = 0.0f;
    if (lane >= nv) return s;
    gfloat* Mc = opaque(M) + lane;
    for (int j = 0; j < nv; j++) s += Mc[j * nv] * vbuf[j];
    return s;
  }

  // J[r] . v for the vector last passed to put_vec (0 for r >= nr)
  MJH_DEV float jrow_dot(int r) const {
    float s = 0.0f;
    if (r >= nr) return s;
    if (r < rj) {
      const float* Jr = Jl + r * JS;
#pragma unroll
      for (int q = 0; q < NVC / 4; q++) {
        const float4 v = vec4(q);
        s += Jr[4 * q] * v.x;
        s += Jr[4 * q + 1] * v.y;
        s += Jr[4 * q + 2] * v.z;
        s += Jr[4 * q + 3] * v.w;
      }
    } else {
      gfloat* Jr = Jg(r);
      for (int c = 0; c < nv; c++) s += Jr[c] * vbuf[c];
    }
    return s;
  }

  // solver.py:75-173: active set, forces, cost; returns cost and gauss
  MJH_DEV void update_constraint(float* cost, float* gauss) {
    float c = 0.0f;
#pragma unroll
    for (int s = 0; s < RS; s++) {
      if (s > 0 && s >= nset) break;
      const float jar = Jaref[s];
      const bool a = jar < 0.0f;
      force[s] = a ? -1.0f * D[s] * jar : 0.0f;
      if (a) c += 0.5f * D[s] * jar * jar;
    }
    c = sum32(c);
    wave_fence();
#pragma unroll
    for (int s = 0; s < RS; s++) {
      if (s > 0 && s >= nset) break;
      fbuf[lane + 32 * s] = force[s];
    }
    wave_fence();
    float q = 0.0f;
    const int nl = nlds();
    if (lane < NVC)
      for (int r = 0; r < nl; r++) q += Jl[r * JS + lane] * fbuf[r];
    const int ci = lane < nv ? lane : 0;
    for (int r = nl; r < nr; r++) q += (lane < nv ? Jg(r)[ci] : 0.0f) * fbuf[r];
    qc = q;
    const float g = sum32(0.5f * (Ma - qs) * (qacc - qas));
    *gauss = g;
    *cost = c + g;
  }

  // solver.py:176-312; returns grad . grad and sets Mgrad
  MJH_DEV float update_gradient() {
    grad = Ma - qs - qc;
    const float gd = sum32(grad * grad);
    if (m.opt.solver == mjhSOL_NEWTON) {
      unsigned act = 0;
#pragma unroll
      for (int s = 0; s < RS; s++) {
        if (s > 0 && s >= nset) break;
        if (Jaref[s] < 0.0f) act |= 1u << s;
      }
      const float changed = sum32(act != act_prev ? 1.0f : 0.0f);
      if (fact_valid && changed == 0.0f) {
        Mgrad = ch.solve(grad);
        return gd;
      }
      act_prev = act;
      fact_valid = true;
      const int ln = opaque(lane);
      gfloat* Mc = opaque(M) + (ln < nv ? ln : 0);
#pragma unroll
      for (int j = 0; j < NVC; j++) {
        const float v = Mc[(j < nv ? j : 0) * nv];
        ch.h[j] = (ln < nv && j < nv) ? v : (j == ln ? 1.0f : 0.0f);
      }
      wave_fence();
#pragma unroll
      for (int s = 0; s < RS; s++) {
        if (s > 0 && s >= nset) break;
        fbuf[lane + 32 * s] = Jaref[s] < 0.0f ? D[s] : 0.0f;
      }
      wave_fence();
      const int nl = nlds();
      for (int r = 0; r < nl; r++) {
        const float dr = fbuf[r];
        if (dr == 0.0f) continue;
        const float* Jr = Jl + r * JS;
        const float a = lane < NVC ? dr * Jr[lane] : 0.0f;
#pragma unroll
        for (int j = 0; j < NVC; j++) ch.h[j] = fmaf(a, Jr[j], ch.h[j]);
      }
      const int c = lane < nv ? lane : 0;
      for (int r = nl; r < nr; r++) {
        const float dr = fbuf[r];
        if (dr == 0.0f) continue;
        gfloat* Jr = Jg(r);
        const float a = lane < nv ? dr * Jr[c] : 0.0f;
#pragma unroll
        for (int j = 0; j < NVC; j++) ch.h[j] = fmaf(a, j < nv ? Jr[j < nv ? j : 0] : 0.0f, ch.h[j]);
      }
      ch.factor();
    } else if (!fact_valid) {
      // CG preconditions with the factor of M, fixed during the solve
      ch.load_lower(d.qLDd + (size_t)w * nv * nv, nv, blk);
      fact_valid = true;
    }
    Mgrad = ch.solve(grad);
    return gd;
  }

  MJH_DEV V3f rows_at(float alpha) const {
    float a = 0.f, b = 0.f, c = 0.f;
#pragma unroll
    for (int s = 0; s < RS; s++)
      if (s == 0 || s < nset)
      if (Jaref[s] + alpha * jv[s] < 0.0f) {
        const V3f p = eval_pt(q0[s], q1[s], q2[s], alpha);
        a += p.a;
        b += p.b;
        c += p.c;
      }
    return V3f{sum32(a), sum32(b), sum32(c)};
  }

  // solver.py:339-664
  MJH_DEV float linesearch_iterative(float g0, float g1, float g2, float search_dot, float prev_alpha) const {
    const float gtol = m.opt.tolerance * m.opt.ls_tolerance * sqrtf(search_dot) * m.meaninertia * (float)max(1, nv);
    float pa = 0.f, pb = 0.f, pc = 0.f;
#pragma unroll
    for (int s = 0; s < RS; s++)
      if (s == 0 || s < nset)
      if (Jaref[s] < 0.0f) {
        pa += q0[s];
        pb += q1[s];
        pc += 2.0f * q2[s];
      }
    const V3f p0 = V3f{g0 + sum32(pa), g1 + sum32(pb), 2.0f * g2 + sum32(pc)};
    float lo_alpha = -safe_div(p0.b, p0.c);
    V3f lo = eval_pt(g0, g1, g2, lo_alpha);
    {
      const V3f s = rows_at(lo_alpha);
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
#pragma unroll
      for (int s = 0; s < RS; s++) {
        if (s > 0 && s >= nset) break;
        const float jar = Jaref[s], v = jv[s];
        if (jar + la * v < 0.0f) {
          const V3f p = eval_pt(q0[s], q1[s], q2[s], la);
          s0 += p.a; s1 += p.b; s2 += p.c;
        }
        if (jar + ha * v < 0.0f) {
          const V3f p = eval_pt(q0[s], q1[s], q2[s], ha);
          s3 += p.a; s4 += p.b; s5 += p.c;
        }
        if (jar + ma * v < 0.0f) {
          const V3f p = eval_pt(q0[s], q1[s], q2[s], ma);
          s6 += p.a; s7 += p.b; s8 += p.c;
        }
      }
      V3f ln = eval_pt(g0, g1, g2, la), hn = eval_pt(g0, g1, g2, ha), mn = eval_pt(g0, g1, g2, ma);
      ln = V3f{ln.a + sum32(s0), ln.b + sum32(s1), ln.c + sum32(s2)};
      hn = V3f{hn.a + sum32(s3), hn.b + sum32(s4), hn.c + sum32(s5)};
      mn = V3f{mn.a + sum32(s6), mn.b + sum32(s7), mn.c + sum32(s8)};
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
  MJH_DEV float linesearch_parallel(float g0, float g1, float g2) const {
    int best = 0;
    float best_cost = 0.0f;
    for (int k = 0; k < m.nlsp; k++) {
      const float a = m.alpha_candidate[k];
      float s0 = 0.f, s1 = 0.f, s2 = 0.f;
#pragma unroll
      for (int s = 0; s < RS; s++)
        if (s == 0 || s < nset)
        if (Jaref[s] + a * jv[s] < 0.0f) {
          s0 += q0[s];
          s1 += q1[s];
          s2 += q2[s];
        }
      const float t0 = g0 + sum32(s0), t1 = g1 + sum32(s1), t2 = g2 + sum32(s2);
      const float c = a * a * t2 + a * t1 + t0;
      if (lane == 0) {
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
    put_vec(search);
    mv = mrow_dot();
#pragma unroll
    for (int s = 0; s < RS; s++) {
      if (s > 0 && s >= nset) break;
      const float v = jrow_dot(row(s));
      const float jar = Jaref[s], dd = D[s];
      jv[s] = v;
      q0[s] = 0.5f * jar * jar * dd;
      q1[s] = v * jar * dd;
      q2[s] = 0.5f * v * v * dd;
    }
    const float g0 = sum32(lane < nv ? gauss / (float)nv : 0.0f);
    const float g1 = sum32(search * (Ma - qs));
    const float g2 = sum32(0.5f * search * mv);
    const float alpha = m.opt.ls_parallel ? linesearch_parallel(g0, g1, g2)
                                          : linesearch_iterative(g0, g1, g2, search_dot, prev_alpha);
    qacc += alpha * search;
    Ma += alpha * mv;
#pragma unroll
    for (int s = 0; s < RS; s++)
      if (s == 0 || s < nset) Jaref[s] += alpha * jv[s];
    return alpha;
  }
};

// Solves world t.world (at most 32 RS rows) with the register solver.
template <int NVC, int RS>
MJH_DEV void solve_reg_nv(const mjhModel& m, const mjhData& d, float* lds, int rj, const Team<kRegTeam>& t) {
  const int w = t.world, nv = m.nv;
  const int radr = d.world_efc[2 * w], nr = d.world_efc[2 * w + 1];
  RegSolver<NVC, RS> s{m, d, w, nv, radr, nr, t.lane, rj};
  s.vbuf = lds;
  s.ch.lane = t.lane;
  s.ch.cbuf = lds + 32;
  s.fbuf = lds + 64;
  s.Jl = lds + 64 + kRegRows;
  s.nset = nr > 32 ? (nr + 31) / 32 : 1;
  s.fact_valid = false;
  s.act_prev = 0;
  const int lane = t.lane;
  const bool dof = lane < nv;
  s.ch.rdiag = 1.0f;
  s.mv = 0.0f;
  const size_t o = (size_t)w * nv;
  // stage the world's inputs (solver.py:973 warm start)
  s.qacc = dof ? d.qacc_warmstart[o + lane] : 0.0f;
  s.qs = dof ? d.qfrc_smooth[o + lane] : 0.0f;
  s.qas = dof ? d.qacc_smooth[o + lane] : 0.0f;
  s.M = d.qMd + o * nv;
  float aref[RS];
#pragma unroll
  for (int k = 0; k < RS; k++) {
    const int r = s.row(k);
    s.D[k] = r < nr ? d.efc_D[radr + r] : 0.0f;
    s.jv[k] = 0.0f;
    aref[k] = r < nr ? d.efc_aref[radr + r] : 0.0f;
  }
  const int nst = nr < rj ? nr : rj;
  wave_fence();
  // rows [0, nst), lanes over columns; padding columns [nv, NVC) zero
  if (lane < NVC)
    for (int r = 0; r < nst; r++) s.Jl[r * s.JS + lane] = dof ? s.Jg(r)[lane] : 0.0f;
  if (m.opt.solver == mjhSOL_CG) {
    // first dof of the lane's kinematic-tree block of qLD (io.py:150-162)
    s.blk = lane;
    for_each_tree_block(m, [&](int dof0, int n) {
      if (lane >= dof0 && lane < dof0 + n) s.blk = dof0;
    });
  }
  // _create_context (solver.py:26-72)
  s.put_vec(s.qacc);
#pragma unroll
  for (int k = 0; k < RS; k++) {
    const int r = s.row(k);
    float a = 0.0f;
    if (r < nr) {
      const float arn = aref[k] / (float)nv;
      if (r < rj) {
        const float* Jr = s.Jl + r * s.JS;
#pragma unroll
        for (int c = 0; c < NVC; c++)
          if (c < nv) a += Jr[c] * s.vbuf[c] - arn;
      } else {
        gfloat* Jr = s.Jg(r);
        for (int c = 0; c < nv; c++) a += Jr[c] * s.vbuf[c] - arn;
      }
    }
    s.Jaref[k] = a;
  }
  s.Ma = s.mrow_dot();
  float cost, gauss;
  float prev_cost = __int_as_float(0x7f800000);
  s.update_constraint(&cost, &gauss);
  float grad_dot = s.update_gradient();
  s.search = -1.0f * s.Mgrad;  // the first direction does not read the previous one
  float search_dot = sum32(s.search * s.search);
  const int iters = m.opt.iterations;
  bool done = false;
  float alpha = d.efc_alpha[w], beta = d.efc_beta[w], beta_num = d.efc_beta_num[w], beta_den = d.efc_beta_den[w];
  for (int it = 0; it < iters; it++) {
    if (iters > 1 && done) break;
    alpha = s.linesearch(gauss, search_dot, alpha);
    const float pg = s.grad, pmg = s.Mgrad;
    prev_cost = cost;
    s.update_constraint(&cost, &gauss);
    grad_dot = s.update_gradient();
    if (m.opt.solver == mjhSOL_CG) {
      beta_num = sum32(s.grad * (s.Mgrad - pmg));
      beta_den = sum32(pg * pmg);
      beta = fmaxf(0.0f, beta_num / fmaxf(MJH_MINVAL, beta_den));
      s.search = -1.0f * s.Mgrad + beta * s.search;
    } else {
      s.search = -1.0f * s.Mgrad;
    }
    search_dot = sum32(s.search * s.search);
    const float scale = m.meaninertia * (float)max(1, nv);
    done = ((prev_cost - cost) / scale < m.opt.tolerance) || (sqrtf(grad_dot) / scale < m.opt.tolerance);
  }
  // write back
  if (dof) {
    d.qacc[o + lane] = s.qacc;
    d.qacc_warmstart[o + lane] = s.qacc;
    d.qfrc_constraint[o + lane] = s.qc;
    d.efc_Ma[o + lane] = s.Ma;
    d.efc_grad[o + lane] = s.grad;
    d.efc_Mgrad[o + lane] = s.Mgrad;
    d.efc_search[o + lane] = s.search;
    d.efc_mv[o + lane] = s.mv;
  }
#pragma unroll
  for (int k = 0; k < RS; k++) {
    const int r = s.row(k);
    if (r < nr) {
      d.efc_force[radr + r] = s.force[k];
      d.efc_Jaref[radr + r] = s.Jaref[k];
      d.efc_jv[radr + r] = s.jv[k];
      d.efc_active[radr + r] = s.Jaref[k] < 0.0f;
    }
  }
  if (lane == 0) {
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
}

// The register solver for world t.world with RS row sets: returns false
// (and leaves the world untouched) when its rows exceed 32 RS.  The solve
// launches take RS = 2 (64 rows, the register budget of two wavefronts per
// SIMD); a tail launch takes the larger worlds with RS = kRegSets.
template <int RS>
MJH_DEV bool solve_reg(const mjhModel& m, const mjhData& d, float* lds, int rj, const Team<kRegTeam>& t) {
  if (!t.valid) return true;
  if (d.world_efc[2 * t.world + 1] > 32 * RS) return false;
  with_nvc(m.nv, [&](auto nvc) {
    solve_reg_nv<decltype(nvc)::value, RS>(m, d, lds, rj, t);
    return 0;
  });
  return true;
}

}  // namespace mjh
