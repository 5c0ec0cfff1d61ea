// Constraint rows of one world: joint limits (slide / hinge) and pyramidal
// condim-3 contacts (constraint.py:22-240).  The world's rows form one
// contiguous block of the efc pool (one atomicAdd on nefc per wavefront):
// limit rows in joint order, then four rows per active contact in contact
// order, the same per-world order as the reference's two launches.
#pragma once

#include "collision.hpp"
#include "common.hpp"

namespace mjh {

// constraint.py:22-73 _update_efc_row
MJH_DEV void update_efc_row(const mjhModel& m, const mjhData& d, int efcid, float pos_aref, float pos_imp,
                            float invweight, const float* solref, const float* solimp, float margin,
                            bool refsafe, float Jqvel) {
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
  const float x = fabsf(pos_imp) / width;
  const float ya = (1.0f / powf(mid, power - 1.0f)) * powf(x, power);
  const float yb = 1.0f - (1.0f / powf(1.0f - mid, power - 1.0f)) * powf(1.0f - x, power);
  const float y = x < mid ? ya : yb;
  float imp = dmin + y * (dmax - dmin);
  imp = fminf(fmaxf(imp, dmin), dmax);
  if (x > 1.0f) imp = dmax;
  d.efc_D[efcid] = 1.0f / fmaxf(invweight * (1.0f - imp) / imp, MJH_MINVAL);
  d.efc_aref[efcid] = -k * imp * pos_aref - b * Jqvel;
  d.efc_pos[efcid] = pos_aref + margin;
  d.efc_margin[efcid] = margin;
}

MJH_DEV bool limit_active(const mjhModel& m, const float* qpos, int j, float* pos, float* J) {
  const float q = qpos[m.jnt_qposadr[j]];
  const float dmin = q - m.jnt_range[2 * j], dmax = m.jnt_range[2 * j + 1] - q;
  *pos = fminf(dmin, dmax) - m.jnt_margin[j];
  *J = (dmin < dmax) ? 1.0f : -1.0f;
  return *pos < 0.0f;
}

// LDS floats per world of make_constraint: per contact of a chunk of TEAM
// contacts, 5 ints (pool index, row offset, two bodies, compacted list) and
// 20 floats (two offsets, frame, two frictions, distance, invweight), and
// the J * qvel of its four rows
template <int TEAM>
__host__ __device__ constexpr int lds_constraint_floats() {
  return 29 * TEAM;
}

// The rows of one world into the block [base, base + total) of the pool
// (reserved by the caller; total counts nlim limit rows first): limit rows
// in joint order, then four rows per active contact in contact order, as
// many as the block and the pool hold.
template <int TEAM>
MJH_DEV void make_constraint_rows(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t, int base,
                                  int total, int nlim) {
  const int w = t.world, nv = m.nv;
  const bool refsafe = !(m.opt.disableflags & mjhDSBL_REFSAFE);
  const bool do_limit = !(m.opt.disableflags & mjhDSBL_LIMIT) && m.nlimited > 0;
  const bool do_contact = !(m.opt.disableflags & mjhDSBL_CONTACT) && m.opt.cone == mjhCONE_PYRAMIDAL;
  const float* qpos = WPTR(d.qpos, w, m.nq);
  const float* qvel = WPTR(d.qvel, w, nv);
  int cadr = 0, ncon = 0;
  if (t.valid && do_contact) {
    cadr = d.world_con[2 * w];
    ncon = d.world_con[2 * w + 1];
  }
  const int cap = d.njmax;
  const int nwritten = total <= 0 ? 0 : max(0, min(total, cap - base));
  if (!t.valid) return;
  if (t.lane == 0) {
    d.world_efc[2 * w] = base;
    d.world_efc[2 * w + 1] = nwritten;
    if (base + total > cap) atomicAdd(d.overflow + 2, min(total, base + total - cap));
  }
  // limit rows (constraint.py:104-141)
  int k0 = 0;
  if (do_limit)
    for (int kk = 0; kk < m.nlimited; kk += TEAM) {
      const int k = kk + t.lane;
      float pos = 0.0f, J = 0.0f;
      int j = 0;
      bool act = false;
      if (k < m.nlimited) {
        j = m.jnt_limited_slide_hinge_adr[k];
        act = limit_active(m, qpos, j, &pos, &J);
      }
      int tot;
      const int off = team_exscan_bits<TEAM, 1>(act ? 1 : 0, &tot);
      if (act && k0 + off < nwritten) {
        const int r = base + k0 + off;
        d.efc_worldid[r] = w;
        float* Jr = d.efc_J + (size_t)r * nv;
        for (int i = 0; i < nv; i++) Jr[i] = 0.0f;
        const int da = m.jnt_dofadr[j];
        Jr[da] = J;
        update_efc_row(m, d, r, pos, pos, m.dof_invweight0[da], m.jnt_solref + 2 * j, m.jnt_solimp + 5 * j,
                       m.jnt_margin[j], refsafe, J * qvel[da]);
      }
      k0 += tot;
    }
  // contact rows (constraint.py:144-210), in chunks of TEAM contacts: the
  // per-contact scalars are gathered with lanes over contacts, the Jacobian
  // rows are built one contact at a time with lanes over dofs, and the
  // impedance / reference terms are computed with lanes over rows.  A
  // contact's four rows are written only if all four fit.
  int* lci = (int*)lds;
  int* lrr = lci + TEAM;
  int* lb = lrr + TEAM;
  int* lact = lb + 2 * TEAM;
  float* lf = (float*)(lact + TEAM);
  float* ljq = lf + 20 * TEAM;
  const float* cdof = WPTR(d.cdof, w, nv * 6);
  const float* sc = WPTR(d.subtree_com, w, m.nbody * 3);
  int rrun = nlim;  // rows of this world so far
  for (int c0 = 0; c0 < ncon; c0 += TEAM) {
    const int c = c0 + t.lane;
    bool act = false;
    if (c < ncon) {
      const int ci = cadr + c;
      const float pos = d.contact_dist[ci] - d.contact_includemargin[ci];
      act = d.contact_dim[ci] == 3 && pos < 0.0f;
      if (act) {
        const int b1 = m.geom_bodyid[d.contact_geom[2 * ci]];
        const int b2 = m.geom_bodyid[d.contact_geom[2 * ci + 1]];
        const float* fr = d.contact_friction + 5 * ci;
        const float fri0 = fr[0];
        float invweight = m.body_invweight0[2 * b1] + m.body_invweight0[2 * b2];
        invweight = invweight + fri0 * fri0 * invweight;
        invweight = invweight * 2.0f * fri0 * fri0 / m.opt.impratio;
        const V3 cpos = ld3(d.contact_pos + 3 * ci);
        float* f = lf + 20 * t.lane;
        st3(f, cpos - ld3(sc + 3 * m.body_rootid[b1]));
        st3(f + 3, cpos - ld3(sc + 3 * m.body_rootid[b2]));
        st9(f + 6, ld9(d.contact_frame + 9 * ci));
        f[15] = fr[0];
        f[16] = fr[1];
        f[17] = pos;
        f[18] = invweight;
        lci[t.lane] = ci;
        lb[2 * t.lane] = b1;
        lb[2 * t.lane + 1] = b2;
      }
    }
    int nact;
    const int off = team_exscan_bits<TEAM, 1>(act ? 1 : 0, &nact);
    if (act) {
      lact[off] = t.lane;
      lrr[t.lane] = rrun + 4 * off;
    }
    sync();
    for (int a = 0; a < nact; a++) {
      const int k = lact[a];
      const int rr = lrr[k];
      if (rr + 4 > nwritten) break;
      const int r0 = base + rr;
      const int b1 = lb[2 * k], b2 = lb[2 * k + 1];
      const float* f = lf + 20 * k;
      const V3 off1 = ld3(f), off2 = ld3(f + 3);
      const V3 n = ld3(f + 6), t1 = ld3(f + 9), t2 = ld3(f + 12);
      const float fr0 = f[15], fr1 = f[16];
      float jq[4] = {0.f, 0.f, 0.f, 0.f};
      for (int i = t.lane; i < nv; i += TEAM) {
        const S6 cd = ld6(cdof + 6 * i);
        const V3 rot = V3{cd.v[0], cd.v[1], cd.v[2]}, lin = V3{cd.v[3], cd.v[4], cd.v[5]};
        V3 jac1 = lin + cross(rot, off1);
        V3 jac2 = lin + cross(rot, off2);
        if (!m.body_dof_tree[b1 * nv + i]) jac1 = V3{0.f, 0.f, 0.f};
        if (!m.body_dof_tree[b2 * nv + i]) jac2 = V3{0.f, 0.f, 0.f};
        const V3 jd = jac2 - jac1;
        const float d0 = dot(n, jd), d1 = dot(t1, jd), d2 = dot(t2, jd);
        const float J0 = d0 + d1 * fr0;
        const float J1 = d0 - d1 * fr0;
        const float J2 = d0 + d2 * fr1;
        const float J3 = d0 - d2 * fr1;
        d.efc_J[(size_t)(r0 + 0) * nv + i] = J0;
        d.efc_J[(size_t)(r0 + 1) * nv + i] = J1;
        d.efc_J[(size_t)(r0 + 2) * nv + i] = J2;
        d.efc_J[(size_t)(r0 + 3) * nv + i] = J3;
        const float qv = qvel[i];
        jq[0] += J0 * qv;
        jq[1] += J1 * qv;
        jq[2] += J2 * qv;
        jq[3] += J3 * qv;
      }
#pragma unroll
      for (int e = 0; e < 4; e++) jq[e] = team_sum<TEAM>(jq[e]);
      if (t.lane == 0) st4(ljq + 4 * a, Q4{jq[0], jq[1], jq[2], jq[3]});
    }
    sync();
    for (int q = t.lane; q < 4 * nact; q += TEAM) {
      const int k = lact[q >> 2];
      const int rr = lrr[k];
      if (rr + 4 > nwritten) continue;
      const int r = base + rr + (q & 3);
      const int ci = lci[k];
      const float* f = lf + 20 * k;
      d.efc_worldid[r] = w;
      update_efc_row(m, d, r, f[17], f[17], f[18], d.contact_solref + 2 * ci, d.contact_solimp + 5 * ci,
                     d.contact_includemargin[ci], refsafe, ljq[q]);
    }
    sync();
    rrun += 4 * nact;
  }
}

// constraint.py:213-240 make_constraint (nefc is reset by the caller):
// count the world's rows, reserve its block, write the rows
template <int TEAM>
MJH_DEV void make_constraint(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  if (m.opt.disableflags & mjhDSBL_CONSTRAINT) return;
  const int w = t.world;
  const bool do_limit = !(m.opt.disableflags & mjhDSBL_LIMIT) && m.nlimited > 0;
  const bool do_contact = !(m.opt.disableflags & mjhDSBL_CONTACT) && m.opt.cone == mjhCONE_PYRAMIDAL;
  const float* qpos = WPTR(d.qpos, w, m.nq);
  int cadr = 0, ncon = 0;
  if (t.valid && do_contact) {
    cadr = d.world_con[2 * w];
    ncon = d.world_con[2 * w + 1];
  }
  int nlim = 0, ncact = 0;
  if (t.valid) {
    if (do_limit)
      for (int k = t.lane; k < m.nlimited; k += TEAM) {
        float pos, J;
        nlim += limit_active(m, qpos, m.jnt_limited_slide_hinge_adr[k], &pos, &J) ? 1 : 0;
      }
    for (int c = t.lane; c < ncon; c += TEAM) {
      const int ci = cadr + c;
      if (d.contact_dim[ci] == 3 && d.contact_dist[ci] - d.contact_includemargin[ci] < 0.0f) ncact++;
    }
  }
  nlim = team_sum_int<TEAM>(nlim);
  ncact = team_sum_int<TEAM>(ncact);
  const int total = nlim + 4 * ncact;
  const int base = team_reserve<TEAM>(d.nefc, t.valid ? total : 0);
  make_constraint_rows(m, d, lds, t, base, total, nlim);
}

// The fused forward's collision and row reservation for models the nxn
// broadphase serves (ngeom <= 100; collision_driver.py:458-547 and the
// counting half of constraint.py:213-240).  A counting pass tests the
// candidate pairs and runs each hit pair's narrowphase without writing:
// its contacts, and those of them that become rows (condim 3 inside the
// include margin); the limit rows are counted beside.  The world's three
// pool blocks (pairs, contacts, rows) are then reserved by three atomics
// in flight together -- one round trip where collision() and
// make_constraint() take three -- and a write pass stores the pairs and
// contacts (with their parameters) exactly as nxn_broadphase and
// narrowphase do: candidate order, type-ordered pairs, contacts only of
// pairs inside the pair pool.  Returns the row block (*rbase, *rtotal,
// limit rows first, *rnlim of them) for make_constraint_rows.
template <int TEAM>
MJH_DEV void collide_and_reserve(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t, int* rbase,
                                 int* rtotal, int* rnlim) {
  const int w = t.world, ng = m.ngeom;
  const bool fp = !(m.opt.disableflags & mjhDSBL_FILTERPARENT);
  const float* gxpos = WPTR(d.geom_xpos, w, ng * 3);
  const float* gxmat = WPTR(d.geom_xmat, w, ng * 9);
  const bool table = ng <= MJH_NXN_TABLE_MAX;
  const int ncand = table ? m.nnxn : ng * (ng - 1) / 2;
  const bool rows = !(m.opt.disableflags & mjhDSBL_CONSTRAINT);
  const bool do_limit = rows && !(m.opt.disableflags & mjhDSBL_LIMIT) && m.nlimited > 0;
  const bool do_crow = rows && !(m.opt.disableflags & mjhDSBL_CONTACT) && m.opt.cone == mjhCONE_PYRAMIDAL;
  if (t.valid) nxn_stage_geoms(m, gxpos, gxmat, lds, !table, t);
  sync();
  // counting pass; the first 21 rounds keep each lane's hit (bit i) and its
  // pair's contact count (3 bits at 3 i: a pair yields at most 4 contacts)
  int npair = 0, ncon = 0, ncact = 0, nlim = 0;
  unsigned hits = 0;
  unsigned long long cnts = 0;
  if (t.valid) {
    for (int p = t.lane, i = 0; p < ncand; p += TEAM, i++) {
      int g1, g2;
      if (!nxn_candidate(m, lds, p, table, fp, true, &g1, &g2)) continue;
      npair++;
      if (i < 32) hits |= 1u << i;
      type_order(m, &g1, &g2);
      ContactSink s{&d, 0, 0, 0, w, false, 0, 0};
      if (do_crow) {
        s.dim = pair_condim(m, g1, g2);
        s.incl = pair_includemargin(m, g1, g2);
      }
      narrowphase_pair(m, gxpos, gxmat, g1, g2, s);
      ncon += s.k;
      ncact += s.nact;
      if (i < 21) cnts |= (unsigned long long)s.k << (3 * i);
    }
    if (do_limit) {
      const float* qpos = WPTR(d.qpos, w, m.nq);
      for (int k = t.lane; k < m.nlimited; k += TEAM) {
        float pos, J;
        nlim += limit_active(m, qpos, m.jnt_limited_slide_hinge_adr[k], &pos, &J) ? 1 : 0;
      }
    }
  }
  npair = team_sum_int<TEAM>(npair);
  ncon = team_sum_int<TEAM>(ncon);
  ncact = team_sum_int<TEAM>(ncact);
  nlim = team_sum_int<TEAM>(nlim);
  const int nrow = rows ? nlim + 4 * ncact : 0;
  int pbase, cbase, ebase;
  team_reserve3<TEAM>(d.ncollision, t.valid ? npair : 0, d.ncon, t.valid ? ncon : 0, d.nefc, t.valid ? nrow : 0,
                      &pbase, &cbase, &ebase);
  *rbase = ebase;
  *rtotal = nrow;
  *rnlim = nlim;
  if (!t.valid) return;
  const int cap = d.nconmax;
  const int npw = npair <= 0 ? 0 : max(0, min(npair, cap - pbase));
  const int ncw = ncon <= 0 ? 0 : max(0, min(ncon, cap - cbase));
  if (t.lane == 0) {
    d.world_pair[2 * w] = pbase;
    d.world_pair[2 * w + 1] = npw;
    if (pbase + npair > cap) atomicAdd(d.overflow + 0, min(npair, pbase + npair - cap));
    d.world_con[2 * w] = cbase;
    d.world_con[2 * w + 1] = ncw;
    if (cbase + ncon > cap) atomicAdd(d.overflow + 1, min(ncon, cbase + ncon - cap));
  }
  // write pass
  int kp = 0, kc = 0;
  for (int p0 = 0, i = 0; p0 < ncand; p0 += TEAM, i++) {
    const int p = p0 + t.lane;
    int g1 = 0, g2 = 0;
    bool ok = false;
    if (p < ncand) {
      ok = nxn_candidate(m, lds, p, table, fp, i >= 32, &g1, &g2);
      if (i < 32) ok = ((hits >> i) & 1u) != 0;
    }
    int tot;
    const int off = team_exscan_bits<TEAM, 1>(ok ? 1 : 0, &tot);
    int c = 0;
    if (ok) {
      PairWriter<TEAM> pw{m, d, pbase, cap, kp + off};
      pw.put(g1, g2, w);
      type_order(m, &g1, &g2);
      if (kp + off < npw) {
        if (i < 21) {
          c = (int)((cnts >> (3 * i)) & 7ull);
        } else {
          ContactSink s{&d, 0, 0, 0, w, false, 0, 0};
          narrowphase_pair(m, gxpos, gxmat, g1, g2, s);
          c = s.k;
        }
      }
    }
    kp += tot;
    int ctot;
    const int coff = team_exscan_bits<TEAM, 3>(c, &ctot);
    if (c > 0) {
      ContactSink s{&d, cbase, cap, kc + coff, w, true, 0, 0};
      narrowphase_pair(m, gxpos, gxmat, g1, g2, s);
      const int n = min(cbase + kc + coff + c, cap) - (cbase + kc + coff);
      if (n > 0) contact_params(m, d, g1, g2, cbase + kc + coff, n);
    }
    kc += ctot;
  }
  sync();
}

}  // namespace mjh
