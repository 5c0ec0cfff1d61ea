// Constraint rows of one world: joint limits (slide / hinge) and pyramidal
// condim-3 contacts (constraint.py:22-240).  The world's rows form one
// contiguous block of the efc pool (one atomicAdd on nefc per wavefront):
// limit rows in joint order, then four rows per active contact in contact
// order, the same per-world order as the reference's two launches.
#pragma once

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

// constraint.py:213-240 make_constraint (nefc is reset by the caller)
template <int TEAM>
MJH_DEV void make_constraint(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  if (m.opt.disableflags & mjhDSBL_CONSTRAINT) return;
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
  // count rows
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

}  // namespace mjh
