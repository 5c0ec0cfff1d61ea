// Velocity / actuation / acceleration stages, M*v, Cartesian applied forces
// and the Euler / implicitfast integrators of one world.
#pragma once

#include "smooth.hpp"

namespace mjh {

// Actuator tile groups (io.py:164-204).  The reference's loop end for the
// last group differs between call sites; `last_end` passes the value each
// caller uses (forward.py:485-489, 629-633, 380-383).
template <class F>
MJH_DEV void for_each_act_tile(const mjhModel& m, int last_end, F f) {
  for (int g = 0; g < m.nact_group; g++) {
    const int beg = m.actuator_moment_tileadr[g];
    const int end = (g == m.nact_group - 1) ? last_end : m.actuator_moment_tileadr[g + 1];
    const int tnv = m.actuator_moment_tilesize_nv[g];
    const int tnu = m.actuator_moment_tilesize_nu[g];
    for (int k = beg; k < end; k++)
      f(m.actuator_moment_offset_nv[k], m.actuator_moment_offset_nu[k], tnv, tnu);
  }
}

// forward.py:441-496 actuator_velocity = moment * qvel over the tree tiles
template <int TEAM>
MJH_DEV void actuator_velocity(const mjhModel& m, const mjhData& d, const Team<TEAM>& t) {
  if (!t.valid) return;
  const int w = t.world, nv = m.nv, nu = m.nu;
  const float* moment = WPTR(d.actuator_moment, w, nu * nv);
  const float* qvel = WPTR(d.qvel, w, nv);
  float* vel = WPTR(d.actuator_velocity, w, nu);
  // sparse models: moment * qvel over all actuators and dofs (forward.py:427-438)
  if (m.opt.is_sparse) {
    for (int a = t.lane; a < nu; a += TEAM) {
      float s = 0.0f;
      for (int j = 0; j < nv; j++) s += moment[(size_t)a * nv + j] * qvel[j];
      vel[a] = s;
    }
    return;
  }
  for_each_act_tile(m, m.nact_group, [&](int onv, int onu, int tnv, int tnu) {
    if (tnu == 0 || tnv == 0) return;
    for (int a = t.lane; a < tnu; a += TEAM) {
      float s = 0.0f;
      const float* row = moment + (onu + a) * nv + onv;
      for (int j = 0; j < tnv; j++) s += row[j] * qvel[onv + j];
      vel[onu + a] = s;
    }
  });
}

// forward.py:423-500 fwd_velocity
template <int TEAM>
MJH_DEV void fwd_velocity(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  actuator_velocity(m, d, t);
  sync();
  com_vel(m, d, lds, t);
  passive(m, d, t);
  rne(m, d, lds, t);
}

// forward.py:503-641 fwd_actuation
template <int TEAM>
MJH_DEV void fwd_actuation(const mjhModel& m, const mjhData& d, const Team<TEAM>& t) {
  const int w = t.world, nv = m.nv, nu = m.nu;
  float* qfrc = WPTR(d.qfrc_actuator, w, nv);
  if (nu == 0 || (m.opt.disableflags & mjhDSBL_ACTUATION)) {
    if (t.valid) {
      float* act_dot = WPTR(d.act_dot, w, m.na);
      for (int i = t.lane; i < m.na; i += TEAM) act_dot[i] = 0.0f;
      for (int i = t.lane; i < nv; i += TEAM) qfrc[i] = 0.0f;
    }
    sync();
    return;
  }
  float* force = WPTR(d.actuator_force, w, nu);
  if (t.valid) {
    const float* length = WPTR(d.actuator_length, w, nu);
    const float* velocity = WPTR(d.actuator_velocity, w, nu);
    const float* ctrl = WPTR(d.ctrl, w, nu);
    const bool noclamp = m.opt.disableflags & mjhDSBL_CLAMPCTRL;
    for (int a = t.lane; a < nu; a += TEAM) {
      const float* g = m.actuator_gainprm + 10 * a;
      const float* b = m.actuator_biasprm + 10 * a;
      float gain = g[0];
      gain += g[1] * length[a];
      gain += g[2] * velocity[a];
      float bias = b[0];
      bias += b[1] * length[a];
      bias += b[2] * velocity[a];
      float c = ctrl[a];
      if (m.actuator_ctrllimited[a] && !noclamp)
        c = fminf(fmaxf(c, m.actuator_ctrlrange[2 * a]), m.actuator_ctrlrange[2 * a + 1]);
      // actuator forcerange is read but never applied by the reference
      // (forward.py:539-541)
      force[a] = gain * c + bias;
    }
  }
  sync();
  if (t.valid) {
    const float* moment = WPTR(d.actuator_moment, w, nu * nv);
    if (m.opt.is_sparse) {
      // sparse models: moment' force over all actuators and dofs (forward.py:571-578)
      for (int j = t.lane; j < nv; j += TEAM) {
        float s = 0.0f;
        for (int a = 0; a < nu; a++) s += moment[(size_t)a * nv + j] * force[a];
        qfrc[j] = s;
      }
    } else {
      for_each_act_tile(m, m.nact_group, [&](int onv, int onu, int tnv, int tnu) {
        if (tnu == 0 || tnv == 0) return;
        for (int j = t.lane; j < tnv; j += TEAM) {
          float s = 0.0f;
          for (int a = 0; a < tnu; a++) s += moment[(onu + a) * nv + onv + j] * force[onu + a];
          qfrc[onv + j] = s;
        }
      });
    }
  }
  sync();
  if (t.valid)
    for (int i = t.lane; i < nv; i += TEAM) {
      const int j = m.dof_jntid[i];
      if (m.jnt_actfrclimited[j])
        qfrc[i] = fminf(fmaxf(qfrc[i], m.jnt_actfrcrange[2 * j]), m.jnt_actfrcrange[2 * j + 1]);
    }
  sync();
}

// support.py:140-172 xfrc_accumulate: qfrc += J' xfrc_applied
template <int TEAM>
MJH_DEV void xfrc_accumulate(const mjhModel& m, const mjhData& d, float* qfrc_all, const Team<TEAM>& t) {
  if (!t.valid) return;
  const int w = t.world, nv = m.nv, nbody = m.nbody;
  const float* cdof = WPTR(d.cdof, w, nv * 6);
  const float* xipos = WPTR(d.xipos, w, nbody * 3);
  const float* sc = WPTR(d.subtree_com, w, nbody * 3);
  const float* xfrc = WPTR(d.xfrc_applied, w, nbody * 6);
  float* qfrc = WPTR(qfrc_all, w, nv);
  for (int i = t.lane; i < nv; i += TEAM) {
    const S6 c = ld6(cdof + 6 * i);
    const V3 rot = V3{c.v[0], c.v[1], c.v[2]};
    const V3 lin = V3{c.v[3], c.v[4], c.v[5]};
    float acc = 0.0f;
    for (int b = m.dof_bodyid[i]; b < nbody; b++) {
      if (!m.body_dof_tree[b * nv + i]) continue;
      const V3 off = ld3(xipos + 3 * b) - ld3(sc + 3 * m.body_rootid[b]);
      const V3 f = ld3(xfrc + 6 * b), tq = ld3(xfrc + 6 * b + 3);
      acc += dot(lin, f) + dot(rot, tq) + dot(cross(rot, off), f);
    }
    qfrc[i] += acc;
  }
}

// forward.py:644-661 fwd_acceleration, in two parts: the smooth force
// qfrc_smooth = passive - bias + actuator + applied + J' xfrc_applied ...
template <int TEAM>
MJH_DEV void fwd_acceleration_force(const mjhModel& m, const mjhData& d, const Team<TEAM>& t) {
  const int w = t.world, nv = m.nv;
  float* qs = WPTR(d.qfrc_smooth, w, nv);
  if (t.valid) {
    const float* pas = WPTR(d.qfrc_passive, w, nv);
    const float* bias = WPTR(d.qfrc_bias, w, nv);
    const float* act = WPTR(d.qfrc_actuator, w, nv);
    const float* app = WPTR(d.qfrc_applied, w, nv);
    for (int i = t.lane; i < nv; i += TEAM) qs[i] = pas[i] - bias[i] + act[i] + app[i];
  }
  sync();
  xfrc_accumulate(m, d, d.qfrc_smooth, t);
  sync();
}

// ... and qacc_smooth = inv(M) qfrc_smooth with the factor qLD; lds: nv*nv + nv floats
template <int TEAM>
MJH_DEV void fwd_acceleration_solve(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  const int w = t.world, nv = m.nv;
  if (t.valid) solve_blocks(m, WPTR(d.qLDd, w, nv * nv), WPTR(d.qacc_smooth, w, nv), WPTR(d.qfrc_smooth, w, nv), lds, t);
  sync();
}

template <int TEAM>
MJH_DEV void fwd_acceleration(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  fwd_acceleration_force(m, d, t);
  fwd_acceleration_solve(m, d, lds, t);
}

// support.py:35-93 mul_m (dense): res = M vec over the tree blocks; worlds
// with skip set are left untouched
template <int TEAM>
MJH_DEV void mul_m(const mjhModel& m, const float* qM, float* res, const float* vec, const Team<TEAM>& t) {
  const int nv = m.nv;
  for_each_tree_block(m, [&](int dof, int n) {
    for (int r = t.lane; r < n; r += TEAM) {
      float s = 0.0f;
      const float* row = qM + (dof + r) * nv + dof;
      for (int c = 0; c < n; c++) s += row[c] * vec[dof + c];
      res[dof + r] = s;
    }
  });
}

// forward.py:43-162 _advance (without the host-side time update)
template <int TEAM>
MJH_DEV void advance(const mjhModel& m, const mjhData& d, const float* qacc_all, const Team<TEAM>& t) {
  if (!t.valid) return;
  const int w = t.world, nv = m.nv;
  const float dt = m.opt.timestep;
  if (m.na) {
    float* act = WPTR(d.act, w, m.na);
    const float* act_dot = WPTR(d.act_dot, w, m.na);
    for (int a = t.lane; a < m.nu; a += TEAM) {
      const int adr = m.actuator_actadr[a];
      if (adr == -1) continue;
      float x = act[adr];
      const float xd = act_dot[adr];
      if (m.actuator_dyntype[a] == mjhDYN_FILTEREXACT) {
        const float prm = m.actuator_dynprm[10 * a];
        const float tau = prm < MJH_MINVAL ? MJH_MINVAL : prm;
        x = x + xd * tau * (1.0f - expf(-dt / tau));
      } else {
        x = x + xd * dt;
      }
      // the activation clamp's result is discarded (forward.py:86)
      act[adr] = x;
    }
  }
  float* qvel = WPTR(d.qvel, w, nv);
  const float* qacc = WPTR(qacc_all, w, nv);
  for (int i = t.lane; i < nv; i += TEAM) qvel[i] = qvel[i] + qacc[i] * dt;
  sync();
  float* qpos = WPTR(d.qpos, w, m.nq);
  for (int j = t.lane; j < m.njnt; j += TEAM) {
    const int jt = m.jnt_type[j], qa = m.jnt_qposadr[j], da = m.jnt_dofadr[j];
    if (jt == mjhJNT_FREE) {
      const V3 p = ld3(qpos + qa) + dt * ld3(qvel + da);
      const Q4 q = quat_integrate(ld4(qpos + qa + 3), ld3(qvel + da + 3), dt);
      st3(qpos + qa, p);
      st4(qpos + qa + 3, q);
    } else if (jt == mjhJNT_BALL) {
      st4(qpos + qa, quat_integrate(ld4(qpos + qa), ld3(qvel + da), dt));
    } else {
      qpos[qa] = qpos[qa] + dt * qvel[da];
    }
  }
}

// forward.py:165-242 euler (dense eulerdamp: (M + dt diag(damping)) qacc =
// qfrc_smooth + qfrc_constraint, factor and solve per tree block);
// lds: nv*nv + nv floats
// forward.py:171-193 eulerdamp_sparse, for sparse models: qM_integration =
// qM + dt diag(damping) in the sparse layout, its L'DL factor and
// qfrc_integration (qacc_integration itself comes from the dense solve)
template <int TEAM>
MJH_DEV void euler_sparse_outputs(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  const int w = t.world, nv = m.nv, nM = m.nM;
  float* Mi = WPTR(d.qM_integration, w, nM);
  if (t.valid) {
    const float* S = WPTR(d.qM, w, nM);
    for (int k = t.lane; k < nM; k += TEAM) Mi[k] = S[k];
  }
  sync();
  if (t.valid)
    for (int i = t.lane; i < nv; i += TEAM) {
      Mi[m.dof_Madr[i]] += m.opt.timestep * m.dof_damping[i];
      d.qfrc_integration[w * nv + i] = d.qfrc_smooth[w * nv + i] + d.qfrc_constraint[w * nv + i];
    }
  sync();
  factor_sparse(m, Mi, WPTR(d.qLD_integration, w, nM), WPTR(d.qLDiagInv_integration, w, nv), lds, t);
}

// Whether the integrator advances with qacc itself (no implicit term to
// solve), else with qacc_integration.
MJH_DEV bool advances_with_qacc(const mjhModel& m) {
  if (m.opt.integrator == mjhINT_EULER) return m.opt.disableflags & mjhDSBL_EULERDAMP;
  const bool passive_on = !(m.opt.disableflags & mjhDSBL_PASSIVE);
  const bool act_on = !(m.opt.disableflags & mjhDSBL_ACTUATION) && m.actuator_affine_bias_gain;
  return !passive_on && !act_on;
}

template <int TEAM>
MJH_DEV void euler(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  const int w = t.world, nv = m.nv;
  if (advances_with_qacc(m)) {
    advance(m, d, d.qacc, t);
    sync();
    return;
  }
  if (m.opt.is_sparse) euler_sparse_outputs(m, d, lds, t);
  if constexpr (TEAM == kRegTeam) {
    // (M + dt diag(damping)) qacc_int = qfrc_smooth + qfrc_constraint, the
    // matrix row and its factor in registers
    if (t.valid) {
      const int lane = t.lane;
      const bool dof = lane < nv;
      int beg, end;
      tree_block(m, lane, &beg, &end);
      const float damp = dof ? m.dof_damping[lane] * m.opt.timestep : 0.0f;
      const float b = dof ? d.qfrc_smooth[w * nv + lane] + d.qfrc_constraint[w * nv + lane] : 0.0f;
      const float x = with_nvc(nv, [&](auto nvc) {
        RegChol<decltype(nvc)::value> c{lane, lds};
        c.load_sym(WPTR(d.qMd, w, nv * nv), nv, beg, end, damp);
        c.factor();
        return c.solve(b);
      });
      if (dof) d.qacc_integration[w * nv + lane] = x;
    }
    sync();
    advance(m, d, d.qacc_integration, t);
    sync();
    return;
  }
  if (t.valid) {
    const float* qM = WPTR(d.qMd, w, nv * nv);
    const float* qs = WPTR(d.qfrc_smooth, w, nv);
    const float* qc = WPTR(d.qfrc_constraint, w, nv);
    float* qacc_int = WPTR(d.qacc_integration, w, nv);
    const float dt = m.opt.timestep;
    for_each_tree_block(m, [&](int dof, int n) {
      float* A = lds;
      float* x = lds + n * n;
      for (int p = t.lane; p < n * n; p += TEAM) {
        const int r = p / n, c = p % n;
        float v = qM[(dof + r) * nv + dof + c];
        if (r == c) v += m.dof_damping[dof + r] * dt;
        A[p] = v;
      }
      for (int p = t.lane; p < n; p += TEAM) x[p] = qs[dof + p] + qc[dof + p];
      sync();
      chol_lds(A, n, n, m, t);
      cho_solve_lds(A, n, n, x, t);
      for (int p = t.lane; p < n; p += TEAM) qacc_int[dof + p] = x[p];
      sync();
    });
  }
  sync();
  advance(m, d, d.qacc_integration, t);
  sync();
}

// forward.py:245-405 implicit (implicitfast, dense); lds: nv*nv + nv floats
template <int TEAM>
MJH_DEV void implicit(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  const int w = t.world, nv = m.nv, nu = m.nu;
  const bool passive_on = !(m.opt.disableflags & mjhDSBL_PASSIVE);
  const bool act_on = !(m.opt.disableflags & mjhDSBL_ACTUATION) && m.actuator_affine_bias_gain;
  if (advances_with_qacc(m)) {
    advance(m, d, d.qacc, t);
    sync();
    return;
  }
  const float dt = m.opt.timestep;
  float* avi = WPTR(d.act_vel_integration, w, nu);
  if (act_on && t.valid) {
    const float* ctrl = WPTR(d.ctrl, w, nu);
    const float* act = WPTR(d.act, w, m.na);
    for (int a = t.lane; a < nu; a += TEAM) {
      float bv = 0.0f, gv = 0.0f;
      if (m.actuator_biastype[a] == mjhBIAS_AFFINE) bv = m.actuator_biasprm[10 * a + 2];
      if (m.actuator_gaintype[a] == mjhGAIN_AFFINE) gv = m.actuator_gainprm[10 * a + 2];
      // the reference indexes act by actuator id here (forward.py:296-297)
      const float c = (m.actuator_dyntype[a] != mjhDYN_NONE) ? act[a] : ctrl[a];
      avi[a] = bv + gv * c;
    }
  }
  sync();
  if constexpr (TEAM == kRegTeam) {
    // qM_integration = qM - dt (moment' diag(vel) moment - diag(damping)),
    // a row per lane in registers, written out, factored and solved in place
    if (t.valid) {
      const int lane = t.lane;
      const bool dof = lane < nv;
      int beg, end;
      tree_block(m, lane, &beg, &end);
      const float damp = (dof && passive_on) ? m.dof_damping[lane] * dt : 0.0f;
      const float b = dof ? d.qfrc_smooth[w * nv + lane] + d.qfrc_constraint[w * nv + lane] : 0.0f;
      if (dof) d.qfrc_integration[w * nv + lane] = b;
      const float* moment = WPTR(d.actuator_moment, w, nu * nv);
      float* qMi = WPTR(d.qM_integration, w, nv * nv);
      const float x = with_nvc(nv, [&](auto nvc) {
        constexpr int NVC = decltype(nvc)::value;
        RegChol<NVC> c{lane, lds};
        c.load_sym(WPTR(d.qMd, w, nv * nv), nv, beg, end, damp);
        if (act_on) {
          for (int a = 0; a < nu; a++) {
            const float* mr = moment + a * nv;
            const float ca = dof ? dt * mr[lane] * avi[a] : 0.0f;
            if (ca == 0.0f) continue;
#pragma unroll
            for (int j = 0; j < NVC; j++)
              if (j < nv && j >= beg && j < end) c.h[j] -= ca * mr[j];
          }
        }
        if (dof) {
#pragma unroll
          for (int j = 0; j < NVC; j++)
            if (j >= beg && j < end) qMi[lane * nv + j] = c.h[j];
        }
        c.factor();
        return c.solve(b);
      });
      if (dof) d.qacc_integration[w * nv + lane] = x;
    }
    sync();
    advance(m, d, d.qacc_integration, t);
    sync();
    return;
  }
  if (t.valid) {
    const float* qM = WPTR(d.qMd, w, nv * nv);
    const float* moment = WPTR(d.actuator_moment, w, nu * nv);
    float* qMi = WPTR(d.qM_integration, w, nv * nv);
    float* qfi = WPTR(d.qfrc_integration, w, nv);
    const float* qs = WPTR(d.qfrc_smooth, w, nv);
    const float* qc = WPTR(d.qfrc_constraint, w, nv);
    // qM_integration = qM - dt * (moment' diag(vel) moment - diag(damping))
    for_each_act_tile(m, m.nqLD_tile, [&](int onv, int onu, int tnv, int tnu) {
      if (tnv == 0) return;
      for (int p = t.lane; p < tnv * tnv; p += TEAM) {
        const int r = p / tnv, c = p % tnv;
        float q = 0.0f;
        if (act_on && tnu != 0)
          for (int a = 0; a < tnu; a++)
            q += moment[(onu + a) * nv + onv + r] * avi[onu + a] * moment[(onu + a) * nv + onv + c];
        if (passive_on && r == c) q += -m.dof_damping[onv + r];
        qMi[(onv + r) * nv + onv + c] = qM[(onv + r) * nv + onv + c] - q * dt;
      }
      for (int j = t.lane; j < tnv; j += TEAM) qfi[onv + j] = qs[onv + j] + qc[onv + j];
    });
  }
  sync();
  if (t.valid) {
    const float* qMi = WPTR(d.qM_integration, w, nv * nv);
    const float* qfi = WPTR(d.qfrc_integration, w, nv);
    float* qacc_int = WPTR(d.qacc_integration, w, nv);
    for_each_tree_block(m, [&](int dof, int n) {
      float* A = lds;
      float* x = lds + n * n;
      for (int p = t.lane; p < n * n; p += TEAM) A[p] = qMi[(dof + p / n) * nv + dof + p % n];
      for (int p = t.lane; p < n; p += TEAM) x[p] = qfi[dof + p];
      sync();
      chol_lds(A, n, n, m, t);
      cho_solve_lds(A, n, n, x, t);
      for (int p = t.lane; p < n; p += TEAM) qacc_int[dof + p] = x[p];
      sync();
    });
  }
  sync();
  advance(m, d, d.qacc_integration, t);
  sync();
}

}  // namespace mjh
