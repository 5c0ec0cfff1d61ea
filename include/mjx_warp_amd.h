/* mjx_warp_amd — C ABI of the MI355X-native batched physics step.
 *
 * The library exposes the per-stage functions of mujoco_warp's public API
 * (mujoco_warp/__init__.py:18-47) as extern "C" entry points that take a
 * model description, a data description and a HIP stream.  Both
 * descriptions are plain structs of device pointers and sizes; the caller
 * (normally the Python package over ctypes) owns every buffer, the library
 * allocates no memory inside a call (mjh_multi_create, which builds a plan,
 * aside) and never synchronises the stream.  It
 * creates, once per host thread and device, a second stream and two events
 * that it keeps until the process ends (see mjh_solve).
 *
 * Field names, shapes and meanings follow the reference's Model / Data /
 * Contact / Constraint structs (mujoco_warp/_src/types.py:218-800).  Every
 * array is world-major: per-world arrays are [nworld, n, k] contiguous,
 * contacts and constraint rows are global pools of capacity nconmax / njmax
 * tagged with a world id (types.py:615-645, 264-363).
 *
 * The array lists below are X-macros: X(kind, name, dims) where kind is
 *   R  real  (float on the device; the CPU restatement may instantiate double)
 *   I  int32
 *   B  bool  (one byte)
 * and dims is the array shape in terms of the size fields.  The same lists
 * build the C structs here, the ctypes mirrors in mjx_warp_amd/_native.py
 * and the CPU restatement's structs, so there is one source of truth.
 */
#ifndef MJX_WARP_AMD_H_
#define MJX_WARP_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MJH_ABI_VERSION 1

/* ---- enums (values of MuJoCo 3.2.7, as used by types.py:27-182) ---- */
enum mjhJoint { mjhJNT_FREE = 0, mjhJNT_BALL = 1, mjhJNT_SLIDE = 2, mjhJNT_HINGE = 3 };
enum mjhGeom { mjhGEOM_PLANE = 0, mjhGEOM_HFIELD = 1, mjhGEOM_SPHERE = 2, mjhGEOM_CAPSULE = 3,
               mjhGEOM_ELLIPSOID = 4, mjhGEOM_CYLINDER = 5, mjhGEOM_BOX = 6, mjhGEOM_MESH = 7 };
enum mjhIntegrator { mjhINT_EULER = 0, mjhINT_RK4 = 1, mjhINT_IMPLICIT = 2, mjhINT_IMPLICITFAST = 3 };
enum mjhCone { mjhCONE_PYRAMIDAL = 0, mjhCONE_ELLIPTIC = 1 };
enum mjhSolver { mjhSOL_PGS = 0, mjhSOL_CG = 1, mjhSOL_NEWTON = 2 };
enum mjhTrn { mjhTRN_JOINT = 0, mjhTRN_JOINTINPARENT = 1 };
enum mjhDyn { mjhDYN_NONE = 0, mjhDYN_INTEGRATOR = 1, mjhDYN_FILTER = 2, mjhDYN_FILTEREXACT = 3 };
enum mjhGain { mjhGAIN_FIXED = 0, mjhGAIN_AFFINE = 1 };
enum mjhBias { mjhBIAS_NONE = 0, mjhBIAS_AFFINE = 1 };
enum mjhDisable {
  mjhDSBL_CONSTRAINT = 1 << 0, mjhDSBL_EQUALITY = 1 << 1, mjhDSBL_FRICTIONLOSS = 1 << 2,
  mjhDSBL_LIMIT = 1 << 3, mjhDSBL_CONTACT = 1 << 4, mjhDSBL_PASSIVE = 1 << 5,
  mjhDSBL_GRAVITY = 1 << 6, mjhDSBL_CLAMPCTRL = 1 << 7, mjhDSBL_WARMSTART = 1 << 8,
  mjhDSBL_FILTERPARENT = 1 << 9, mjhDSBL_ACTUATION = 1 << 10, mjhDSBL_REFSAFE = 1 << 11,
  mjhDSBL_SENSOR = 1 << 12, mjhDSBL_MIDPHASE = 1 << 13, mjhDSBL_EULERDAMP = 1 << 14
};

#define MJH_MINVAL 1e-15f
#define MJH_MINIMP 0.0001f
#define MJH_MAXIMP 0.9999f

/* ---- status codes ---- */
enum mjhStatus {
  MJH_OK = 0,
  MJH_ERR_INVALID = -1,      /* bad argument / shape (reference: ValueError) */
  MJH_ERR_UNSUPPORTED = -2,  /* unsupported feature (reference: NotImplementedError) */
  MJH_ERR_HIP = -3,          /* HIP runtime error */
  MJH_ERR_CAPACITY = -4      /* a size exceeds what a kernel can stage */
};

/* ---- model ---- */
typedef struct mjhOption {
  float timestep;
  float impratio;
  float tolerance;
  float ls_tolerance;
  float gravity[3];
  int integrator;
  int cone;
  int solver;
  int iterations;
  int ls_iterations;
  int disableflags;
  int is_sparse;
  int ls_parallel;
} mjhOption;

/* size fields of the model (types.py:493-504 plus the lengths of the
 * warp-only index tables built by io.py:76-233) */
#define MJH_MODEL_SIZES(X) \
  X(nq) X(nv) X(nu) X(na) X(nbody) X(njnt) X(ngeom) X(nsite) X(nexclude) X(nmocap) \
  X(nM) X(nlsp) X(nlevel) X(nlimited) X(ntril) X(nfullm) X(nmulm) X(nqLD_update) \
  X(nqLD_update_level) X(nqLD_tile) X(nqLD_group) X(nact_tile) X(nact_group) \
  X(nmesh) X(nmeshvert) X(actuator_affine_bias_gain) X(max_level_width) X(nnxn) X(tree_dof_max)

#define MJH_MODEL_ARRAYS(X) \
  X(R, qpos0, (nq)) \
  X(R, qpos_spring, (nq)) \
  X(I, body_tree, (nbody)) \
  X(I, body_treeadr, (nlevel)) \
  X(I, actuator_moment_offset_nv, (nact_tile)) \
  X(I, actuator_moment_offset_nu, (nact_tile)) \
  X(I, actuator_moment_tileadr, (nact_group)) \
  X(I, actuator_moment_tilesize_nv, (nact_group)) \
  X(I, actuator_moment_tilesize_nu, (nact_group)) \
  X(R, alpha_candidate, (nlsp)) \
  X(I, qM_fullm_i, (nfullm)) \
  X(I, qM_fullm_j, (nfullm)) \
  X(I, qM_mulm_i, (nmulm)) \
  X(I, qM_mulm_j, (nmulm)) \
  X(I, qM_madr_ij, (nmulm)) \
  X(I, qLD_update_tree, (nqLD_update, 3)) \
  X(I, qLD_update_treeadr, (nqLD_update_level)) \
  X(I, qLD_tile, (nqLD_tile)) \
  X(I, qLD_tileadr, (nqLD_group)) \
  X(I, qLD_tilesize, (nqLD_group)) \
  X(I, body_parentid, (nbody)) \
  X(I, body_rootid, (nbody)) \
  X(I, body_weldid, (nbody)) \
  X(I, body_mocapid, (nbody)) \
  X(I, body_jntnum, (nbody)) \
  X(I, body_jntadr, (nbody)) \
  X(I, body_dofnum, (nbody)) \
  X(I, body_dofadr, (nbody)) \
  X(I, body_geomnum, (nbody)) \
  X(I, body_geomadr, (nbody)) \
  X(R, body_pos, (nbody, 3)) \
  X(R, body_quat, (nbody, 4)) \
  X(R, body_ipos, (nbody, 3)) \
  X(R, body_iquat, (nbody, 4)) \
  X(R, body_mass, (nbody)) \
  X(R, subtree_mass, (nbody)) \
  X(R, body_inertia, (nbody, 3)) \
  X(R, body_invweight0, (nbody, 2)) \
  X(I, body_contype, (nbody)) \
  X(I, body_conaffinity, (nbody)) \
  X(I, jnt_type, (njnt)) \
  X(I, jnt_qposadr, (njnt)) \
  X(I, jnt_dofadr, (njnt)) \
  X(I, jnt_bodyid, (njnt)) \
  X(I, jnt_limited, (njnt)) \
  X(B, jnt_actfrclimited, (njnt)) \
  X(R, jnt_solref, (njnt, 2)) \
  X(R, jnt_solimp, (njnt, 5)) \
  X(R, jnt_pos, (njnt, 3)) \
  X(R, jnt_axis, (njnt, 3)) \
  X(R, jnt_stiffness, (njnt)) \
  X(R, jnt_range, (njnt, 2)) \
  X(R, jnt_actfrcrange, (njnt, 2)) \
  X(R, jnt_margin, (njnt)) \
  X(I, jnt_limited_slide_hinge_adr, (nlimited)) \
  X(I, dof_bodyid, (nv)) \
  X(I, dof_jntid, (nv)) \
  X(I, dof_parentid, (nv)) \
  X(I, dof_Madr, (nv)) \
  X(R, dof_armature, (nv)) \
  X(R, dof_damping, (nv)) \
  X(R, dof_invweight0, (nv)) \
  X(I, dof_tri_row, (ntril)) \
  X(I, dof_tri_col, (ntril)) \
  X(I, geom_type, (ngeom)) \
  X(I, geom_contype, (ngeom)) \
  X(I, geom_conaffinity, (ngeom)) \
  X(I, geom_condim, (ngeom)) \
  X(I, geom_bodyid, (ngeom)) \
  X(I, geom_dataid, (ngeom)) \
  X(I, geom_priority, (ngeom)) \
  X(R, geom_solmix, (ngeom)) \
  X(R, geom_solref, (ngeom, 2)) \
  X(R, geom_solimp, (ngeom, 5)) \
  X(R, geom_size, (ngeom, 3)) \
  X(R, geom_aabb, (ngeom, 2, 3)) \
  X(R, geom_rbound, (ngeom)) \
  X(R, geom_pos, (ngeom, 3)) \
  X(R, geom_quat, (ngeom, 4)) \
  X(R, geom_friction, (ngeom, 3)) \
  X(R, geom_margin, (ngeom)) \
  X(R, geom_gap, (ngeom)) \
  X(I, site_bodyid, (nsite)) \
  X(R, site_pos, (nsite, 3)) \
  X(R, site_quat, (nsite, 4)) \
  X(I, mesh_vertadr, (nmesh)) \
  X(I, mesh_vertnum, (nmesh)) \
  X(R, mesh_vert, (nmeshvert, 3)) \
  X(I, actuator_trntype, (nu)) \
  X(I, actuator_dyntype, (nu)) \
  X(I, actuator_gaintype, (nu)) \
  X(I, actuator_biastype, (nu)) \
  X(I, actuator_trnid, (nu, 2)) \
  X(I, actuator_actadr, (nu)) \
  X(B, actuator_ctrllimited, (nu)) \
  X(B, actuator_forcelimited, (nu)) \
  X(B, actuator_actlimited, (nu)) \
  X(R, actuator_dynprm, (nu, 10)) \
  X(R, actuator_gainprm, (nu, 10)) \
  X(R, actuator_biasprm, (nu, 10)) \
  X(R, actuator_ctrlrange, (nu, 2)) \
  X(R, actuator_forcerange, (nu, 2)) \
  X(R, actuator_actrange, (nu, 2)) \
  X(R, actuator_gear, (nu, 6)) \
  X(I, exclude_signature, (nexclude)) \
  X(B, body_dof_tree, (nbody, nv)) \
  X(I, nxn_pair, (nnxn))
/* body_dof_tree[b, i] = 1 iff dof i's body is an ancestor-or-self of body b
 * (b != 0): the kinematic-tree membership that constraint.py:86-93 and
 * support.py:152-158 recompute by walking body_parentid.
 * nxn_pair: for models of at most MJH_NXN_TABLE_MAX geoms, the geom pairs
 * g1 < g2 in the nxn broadphase's order (collision_driver.py:465-478) that
 * pass the model-only part of its filter (collision_driver.py:27-49:
 * contype / conaffinity mask, distinct weld bodies), packed as
 * g1 | g2 << 15 | parent_child << 30, where parent_child marks the pairs
 * the parent filter removes unless DISABLE_FILTERPARENT is set; empty
 * (nnxn = 0) for larger models. */
#define MJH_NXN_TABLE_MAX 100

/* ---- data ---- */
/* step_desc: optional device buffer of MJH_STEP_DESC_WORDS words (ndesc)
 * that mjh_step_fused (and, for nv <= 32, mjh_solve and so mjh_step) fills
 * with the model / data description each call (by a one-thread kernel, so
 * the call stays capturable) and whose kernels then read it from device
 * memory; a model passed only as a kernel argument compiles to slower code
 * for these kernels (part of it copied to scratch memory, more registers).
 * NULL: the kernel-argument path, with the same results. */
#define MJH_STEP_DESC_WORDS 2048

#define MJH_DATA_SIZES(X) X(nworld) X(nconmax) X(njmax) X(qM_rows) X(qM_cols) X(qMd_n)

#define MJH_DATA_ARRAYS(X) \
  X(I, ncon, (1)) \
  X(I, nefc, (1)) \
  X(I, ncollision, (1)) \
  X(R, qpos, (nworld, nq)) \
  X(R, qvel, (nworld, nv)) \
  X(R, act, (nworld, na)) \
  X(R, qacc_warmstart, (nworld, nv)) \
  X(R, ctrl, (nworld, nu)) \
  X(R, qfrc_applied, (nworld, nv)) \
  X(R, xfrc_applied, (nworld, nbody, 6)) \
  X(R, mocap_pos, (nworld, nmocap, 3)) \
  X(R, mocap_quat, (nworld, nmocap, 4)) \
  X(R, qacc, (nworld, nv)) \
  X(R, act_dot, (nworld, na)) \
  X(R, xpos, (nworld, nbody, 3)) \
  X(R, xquat, (nworld, nbody, 4)) \
  X(R, xmat, (nworld, nbody, 3, 3)) \
  X(R, xipos, (nworld, nbody, 3)) \
  X(R, ximat, (nworld, nbody, 3, 3)) \
  X(R, xanchor, (nworld, njnt, 3)) \
  X(R, xaxis, (nworld, njnt, 3)) \
  X(R, geom_xpos, (nworld, ngeom, 3)) \
  X(R, geom_xmat, (nworld, ngeom, 3, 3)) \
  X(R, site_xpos, (nworld, nsite, 3)) \
  X(R, site_xmat, (nworld, nsite, 3, 3)) \
  X(R, subtree_com, (nworld, nbody, 3)) \
  X(R, cdof, (nworld, nv, 6)) \
  X(R, cinert, (nworld, nbody, 10)) \
  X(R, actuator_length, (nworld, nu)) \
  X(R, actuator_moment, (nworld, nu, nv)) \
  X(R, crb, (nworld, nbody, 10)) \
  X(R, qM, (nworld, qM_rows, qM_cols)) \
  X(R, qLD, (nworld, qM_rows, qM_cols)) \
  X(R, qLDiagInv, (nworld, nv)) \
  X(R, qMd, (nworld, qMd_n, qMd_n)) \
  X(R, qLDd, (nworld, qMd_n, qMd_n)) \
  X(R, actuator_velocity, (nworld, nu)) \
  X(R, cvel, (nworld, nbody, 6)) \
  X(R, cdof_dot, (nworld, nv, 6)) \
  X(R, qfrc_bias, (nworld, nv)) \
  X(R, qfrc_spring, (nworld, nv)) \
  X(R, qfrc_damper, (nworld, nv)) \
  X(R, qfrc_passive, (nworld, nv)) \
  X(R, actuator_force, (nworld, nu)) \
  X(R, qfrc_actuator, (nworld, nv)) \
  X(R, qfrc_smooth, (nworld, nv)) \
  X(R, qacc_smooth, (nworld, nv)) \
  X(R, qfrc_constraint, (nworld, nv)) \
  X(R, contact_dist, (nconmax)) \
  X(R, contact_pos, (nconmax, 3)) \
  X(R, contact_frame, (nconmax, 3, 3)) \
  X(R, contact_includemargin, (nconmax)) \
  X(R, contact_friction, (nconmax, 5)) \
  X(R, contact_solref, (nconmax, 2)) \
  X(R, contact_solreffriction, (nconmax, 2)) \
  X(R, contact_solimp, (nconmax, 5)) \
  X(I, contact_dim, (nconmax)) \
  X(I, contact_geom, (nconmax, 2)) \
  X(I, contact_efc_address, (nconmax)) \
  X(I, contact_worldid, (nconmax)) \
  X(I, efc_worldid, (njmax)) \
  X(R, efc_J, (njmax, nv)) \
  X(R, efc_pos, (njmax)) \
  X(R, efc_margin, (njmax)) \
  X(R, efc_D, (njmax)) \
  X(R, efc_aref, (njmax)) \
  X(R, efc_force, (njmax)) \
  X(R, efc_Jaref, (njmax)) \
  X(R, efc_Ma, (nworld, nv)) \
  X(R, efc_grad, (nworld, nv)) \
  X(R, efc_grad_dot, (nworld)) \
  X(R, efc_Mgrad, (nworld, nv)) \
  X(R, efc_search, (nworld, nv)) \
  X(R, efc_search_dot, (nworld)) \
  X(R, efc_gauss, (nworld)) \
  X(R, efc_cost, (nworld)) \
  X(R, efc_prev_cost, (nworld)) \
  X(I, efc_solver_niter, (nworld)) \
  X(I, efc_active, (njmax)) \
  X(R, efc_gtol, (nworld)) \
  X(R, efc_mv, (nworld, nv)) \
  X(R, efc_jv, (njmax)) \
  X(R, efc_quad, (njmax, 3)) \
  X(R, efc_quad_gauss, (nworld, 3)) \
  X(R, efc_h, (nworld, nv, nv)) \
  X(R, efc_alpha, (nworld)) \
  X(R, efc_prev_grad, (nworld, nv)) \
  X(R, efc_prev_Mgrad, (nworld, nv)) \
  X(R, efc_beta, (nworld)) \
  X(R, efc_beta_num, (nworld)) \
  X(R, efc_beta_den, (nworld)) \
  X(B, efc_done, (nworld)) \
  X(B, efc_ls_done, (nworld)) \
  X(R, efc_p0, (nworld, 3)) \
  X(R, efc_lo, (nworld, 3)) \
  X(R, efc_lo_alpha, (nworld)) \
  X(R, efc_hi, (nworld, 3)) \
  X(R, efc_hi_alpha, (nworld)) \
  X(R, efc_lo_next, (nworld, 3)) \
  X(R, efc_lo_next_alpha, (nworld)) \
  X(R, efc_hi_next, (nworld, 3)) \
  X(R, efc_hi_next_alpha, (nworld)) \
  X(R, efc_mid, (nworld, 3)) \
  X(R, efc_mid_alpha, (nworld)) \
  X(R, efc_cost_candidate, (nworld, nlsp)) \
  X(R, efc_quad_total_candidate, (nworld, nlsp, 3)) \
  X(R, rne_cacc, (nworld, nbody, 6)) \
  X(R, rne_cfrc, (nworld, nbody, 6)) \
  X(R, qfrc_integration, (nworld, nv)) \
  X(R, qacc_integration, (nworld, nv)) \
  X(R, act_vel_integration, (nworld, nu)) \
  X(R, qM_integration, (nworld, qM_rows, qM_cols)) \
  X(R, qLD_integration, (nworld, qM_rows, qM_cols)) \
  X(R, qLDiagInv_integration, (nworld, nv)) \
  X(R, sap_geom_sort, (nworld, ngeom, 4)) \
  X(R, sap_projection_lower, (nworld2, ngeom)) \
  X(R, sap_projection_upper, (nworld, ngeom)) \
  X(I, sap_sort_index, (nworld2, ngeom)) \
  X(I, sap_range, (nworld, ngeom)) \
  X(I, sap_cumulative_sum, (nworld_ngeom)) \
  X(I, sap_segment_index, (nworld1)) \
  X(I, collision_pair, (nconmax, 2)) \
  X(I, collision_worldid, (nconmax)) \
  X(I, world_pair, (nworld, 2)) \
  X(I, world_con, (nworld, 2)) \
  X(I, world_efc, (nworld, 2)) \
  X(I, world_flags, (nworld)) \
  X(I, step_desc, (ndesc)) \
  X(I, overflow, (4))
/* The last five arrays have no counterpart in types.py.  Each world's
 * broadphase pairs, contacts and constraint rows occupy one contiguous block
 * of the global pool, reserved with one atomic per wavefront; world_pair,
 * world_con and world_efc hold (first index, count written) of that block so
 * that later stages find a world's rows without scanning the pool.
 * For nv <= 32 the solve launch holds 64 rows per world; worlds with 65-256
 * rows are solved (and, in the fused step, integrated) by a second launch
 * that runs beside it on the library's second stream.  world_flags marks
 * worlds over every register capacity; the global-memory solver finishes
 * them in a tail launch of the same step.  overflow
 * counts pairs / contacts / rows dropped because a pool was full (the
 * reference drops contacts silently and writes rows out of bounds,
 * collision_primitive.py:64-65, constraint.py:119,162).
 * qM / qLD have the reference's layout: dense [nworld, nv, nv] (qM_rows =
 * qM_cols = nv; qLD the Cholesky factor L, M = L L') or, for sparse models,
 * MuJoCo's [nworld, 1, nM] (qM_rows = 1, qM_cols = nM; qLD the L'DL factor
 * with qLDiagInv = 1 / D).  qMd / qLDd are the dense matrix and Cholesky
 * factor every kernel works with: for dense models qMd_n = 0 and the caller
 * passes qM's and qLD's pointers as qMd and qLDd; for sparse models
 * qMd_n = nv and they are separate arrays.
 * Size aliases: nworld2 = 2*nworld, nworld1 = nworld+1,
 * nworld_ngeom = nworld*ngeom. */

#define MJH_X_FIELD_R(name) float* name;
#define MJH_X_FIELD_I(name) int32_t* name;
#define MJH_X_FIELD_B(name) uint8_t* name;
#define MJH_X_FIELD(kind, name, dims) MJH_X_FIELD_##kind(name)
#define MJH_X_SIZE(name) int32_t name;

typedef struct mjhModel {
  MJH_MODEL_SIZES(MJH_X_SIZE)
  mjhOption opt;
  float meaninertia;
  MJH_MODEL_ARRAYS(MJH_X_FIELD)
} mjhModel;

typedef struct mjhData {
  MJH_DATA_SIZES(MJH_X_SIZE)
  MJH_DATA_ARRAYS(MJH_X_FIELD)
} mjhData;

/* ---- entry points ----
 * Every entry point returns an mjhStatus.  stream is a hipStream_t (0 = the
 * null stream).  Calls are asynchronous and capture-safe (no allocation, no
 * host synchronisation, no memset nodes).  A model and a data description
 * must have been checked once with mjh_check before use. */
typedef void* mjhStream;

int mjh_abi_version(void);
const char* mjh_last_error(void);
/* validates sizes and pointers of a model/data pair (host-only, no launch) */
int mjh_check(const mjhModel* m, const mjhData* d);

/* smooth.py:32-145 kinematics */
int mjh_kinematics(const mjhModel* m, const mjhData* d, mjhStream stream);
/* smooth.py:148-245 com_pos */
int mjh_com_pos(const mjhModel* m, const mjhData* d, mjhStream stream);
/* smooth.py:248-314 crb */
int mjh_crb(const mjhModel* m, const mjhData* d, mjhStream stream);
/* smooth.py:392-395 factor_m */
int mjh_factor_m(const mjhModel* m, const mjhData* d, mjhStream stream);
/* smooth.py:693-696 solve_m: x = inv(M) y, x and y are [nworld, nv] */
int mjh_solve_m(const mjhModel* m, const mjhData* d, float* x, const float* y, mjhStream stream);
/* smooth.py:398-466 rne */
int mjh_rne(const mjhModel* m, const mjhData* d, mjhStream stream);
/* smooth.py:469-536 transmission */
int mjh_transmission(const mjhModel* m, const mjhData* d, mjhStream stream);
/* smooth.py:539-606 com_vel */
int mjh_com_vel(const mjhModel* m, const mjhData* d, mjhStream stream);
/* passive.py:27-108 passive */
int mjh_passive(const mjhModel* m, const mjhData* d, mjhStream stream);
/* support.py:35-137 mul_m: res = M vec for worlds where skip[w] == 0
 * (skip may be NULL) */
int mjh_mul_m(const mjhModel* m, const mjhData* d, float* res, const float* vec,
              const uint8_t* skip, mjhStream stream);
/* support.py:140-172 xfrc_accumulate: qfrc += J' xfrc_applied */
int mjh_xfrc_accumulate(const mjhModel* m, const mjhData* d, float* qfrc, mjhStream stream);
/* collision_driver.py:525-547 collision (broadphase + narrowphase + params) */
int mjh_collision(const mjhModel* m, const mjhData* d, mjhStream stream);
/* collision_driver.py:458-512 nxn_broadphase */
int mjh_nxn_broadphase(const mjhModel* m, const mjhData* d, mjhStream stream);
/* collision_driver.py:350-455 sap_broadphase */
int mjh_sap_broadphase(const mjhModel* m, const mjhData* d, mjhStream stream);
/* collision_primitive.py:271-274 primitive_narrowphase, plus the contact
 * parameter mixing of collision_driver.py:304-347 */
int mjh_primitive_narrowphase(const mjhModel* m, const mjhData* d, mjhStream stream);
/* constraint.py:213-240 make_constraint */
int mjh_make_constraint(const mjhModel* m, const mjhData* d, mjhStream stream);
/* solver.py:872-1000 solve */
int mjh_solve(const mjhModel* m, const mjhData* d, mjhStream stream);
/* forward.py:408-420, 423-500, 503-641, 644-661 */
int mjh_fwd_position(const mjhModel* m, const mjhData* d, mjhStream stream);
int mjh_fwd_velocity(const mjhModel* m, const mjhData* d, mjhStream stream);
int mjh_fwd_actuation(const mjhModel* m, const mjhData* d, mjhStream stream);
int mjh_fwd_acceleration(const mjhModel* m, const mjhData* d, mjhStream stream);
/* forward.py:165-242 euler, 245-405 implicit (implicitfast) */
int mjh_euler(const mjhModel* m, const mjhData* d, mjhStream stream);
int mjh_implicit(const mjhModel* m, const mjhData* d, mjhStream stream);
/* forward.py:664-679 forward: one launch per stage */
int mjh_forward(const mjhModel* m, const mjhData* d, mjhStream stream);
/* forward.py:682-695 step: one launch per stage (the plain per-kernel step).
 * Does not advance d.time, which the reference keeps on the host. */
int mjh_step(const mjhModel* m, const mjhData* d, mjhStream stream);
/* The same step as mjh_step with the stages fused: for nv <= 32 two
 * launches (every forward stage; solver and integrator) on 32-lane teams,
 * otherwise one launch per step; worlds whose constraint rows exceed the
 * fused solver's capacity finish in a launch beside it or a tail launch.  No counterpart in the
 * reference; results equal mjh_step's up to floating-point summation
 * order. */
int mjh_step_fused(const mjhModel* m, const mjhData* d, mjhStream stream);
/* forward.py:43-162 and 682-695 for nstep steps in one call */
int mjh_rollout(const mjhModel* m, const mjhData* d, int nstep, int fused, mjhStream stream);

/* ---- heterogeneous batches ----
 * No counterpart in the reference, whose batch is one Model (types.py: every
 * world of a Data shares the Model).  A plan groups several (model, data)
 * pairs -- each group its own model and its own worlds -- so that one fused
 * step advances all of them with one launch per stage over every group
 * (workgroups find their group in a device table), instead of one step call
 * per model.  Groups with nv <= 32 run on 32-lane teams, larger ones on
 * 64-lane teams; each group's Data keeps its own pools and counters. */
typedef struct mjhMulti mjhMulti;
/* Checks every group as mjh_check does and builds the plan (the one call
 * here that allocates: the group table in device memory).  The plan holds
 * the groups' pointers: their buffers must outlive it. */
int mjh_multi_create(int ngroups, const mjhModel* models, const mjhData* datas, mjhMulti** out);
/* Launch choices for one group of a plan (no counterpart in the reference).
 * team: lanes per world -- 32 (two worlds per wavefront, register solver,
 * nv <= 32 only) or 64 (one world per wavefront, LDS solver); 0 takes the
 * library's choice (32 iff nv <= 32).  rmax: the LDS solver's row capacity
 * per world (a world with more rows is solved by the global-memory tail);
 * <= 0 takes the library's.  rj: Jacobian rows per world staged in LDS
 * (all of a world's rows when it has at most rj, else none on 64 lanes; the
 * first rj on 32 lanes); < 0 takes the library's.  Every choice computes
 * the same step; they differ in speed. */
typedef struct mjhGroupTune {
  int team;
  int rmax;
  int rj;
} mjhGroupTune;
/* flags of mjh_multi_create_tuned: launch the groups in buckets, one per
 * team size and LDS occupancy class (workgroups a CU holds by the LDS of the
 * group's launches), each bucket's launches sized for its own groups; without
 * it every launch takes the largest group's LDS. */
#define MJH_MULTI_BUCKETS 1
/* Work-queue launches: each stage launch of the plan is a grid of a few
 * workgroups per SIMD that take worlds from eight counters in device
 * memory (each counter every eighth world, largest models first) until none
 * are left, instead of one workgroup per world (or pair of worlds); the same
 * results. */
#define MJH_MULTI_QUEUE 2
/* mjh_multi_create with per-group choices (`tunes`: ngroups entries, or
 * NULL for the library's) and flags. */
int mjh_multi_create_tuned(int ngroups, const mjhModel* models, const mjhData* datas, const mjhGroupTune* tunes,
                           int flags, mjhMulti** out);
/* One fused step (mjh_step_fused) of every world of every group. */
int mjh_multi_step(const mjhMulti* plan, mjhStream stream);
/* Frees the plan (after its last step has finished). */
int mjh_multi_destroy(mjhMulti* plan);

/* No counterpart in the reference.  The same step as mjh_multi_create +
 * mjh_multi_step without a plan and without any allocation, for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the group table lives in
 * a fixed device array of the library (one per GPU, held until the process
 * ends) that the call fills with one-thread kernels on `stream` before the
 * step's launches; the big-world launch runs in order on `stream` (no
 * second stream).  At most MJH_MULTI_STATIC_MAX_GROUPS groups (larger
 * counts return MJH_ERR_CAPACITY).  Every call on a GPU rewrites that GPU's
 * array: calls that overlap on one GPU (two streams, two host threads) race
 * -- order them on one stream or with events. */
#define MJH_MULTI_STATIC_MAX_GROUPS 64
int mjh_multi_step_static(int ngroups, const mjhModel* models, const mjhData* datas, mjhStream stream);

/* ---- transition derivatives and the adjoint ----
 * The reference has no gradients (its kernels overwrite arrays in place;
 * collision_convex.py:22 disables backward code generation); these follow
 * MuJoCo's mjd_transitionFD (engine_derivative_fd.c) instead.  State
 * x = (dq [nv], qvel [nv], act [na]), nx = 2 nv + na, with dq the tangent
 * space of qpos (local-frame rotation vectors for free / ball joints, as
 * mj_integratePos / mj_differentiatePos); input u = ctrl [nu]. */

/* Finite-difference Jacobians of one fused step at d's state, per world:
 * A [nworld, nx, nx] = dx'/dx, B [nworld, nx, nu] = dx'/du.  fd is a Data
 * of the same model with fd->nworld = d->nworld * (1 + S (nx + nu)) worlds
 * (S = 2 if centered, else 1), used as scratch: world w*K holds the
 * nominal state, w*K + 1 + c*S + s column c's perturbation by +eps (s = 0)
 * or -eps (s = 1).  d is read and not stepped.  A control column uses only
 * the sides that stay inside ctrlrange (zero column if neither does), as
 * mjd_transitionFD does. */
int mjh_transition_fd(const mjhModel* m, const mjhData* d, const mjhData* fd, float eps, int centered, float* A,
                      float* B, mjhStream stream);
/* One step of the adjoint recursion at d's state (d->qpos is used to turn
 * a raw qpos gradient into the tangent space):
 *   lam = A' lam_next + g_x,   grad_ctrl = B' lam_next,
 * lam / lam_next [nworld, nx], g_x from g_qpos [nworld, nq] (raw
 * coordinates), g_qvel [nworld, nv], g_act [nworld, na].  Null gradients and
 * a null lam_next count as zero; grad_ctrl may be null. */
int mjh_adjoint_step(const mjhModel* m, const mjhData* d, const float* A, const float* B, const float* g_qpos,
                     const float* g_qvel, const float* g_act, const float* lam_next, float* lam, float* grad_ctrl,
                     mjhStream stream);
/* The analytic vector-Jacobian product of one step (no counterpart in the
 * reference, which has no gradients; MuJoCo's transition derivative
 * conventions): d holds the step just taken from the state (qpos
 * [nworld, nq], qvel [nworld, nv]) -- its forward quantities at that state
 * and qvel' in d->qvel, as mjh_step_fused leaves them --, lam_next
 * [nworld, nx] is the adjoint of the successor state.  Writes
 *   lam = (dx'/dx)' lam_next (+ gadd [nworld, nx] if not null),
 *   grad_ctrl = (dx'/du)' lam_next   [nworld, nu],
 * by reverse mode through the integrator, the smooth dynamics (rne, the
 * mass matrix, springs, dampers, actuators) and implicit differentiation of
 * the converged constraint optimum with its active set held (one solve
 * with M + dt diag(damping) and one with M + J_A' D_A J_A per world);
 * contact points and frames follow the narrowphase's plane and sphere /
 * capsule pair geometry (other pairs: held frames).  Euler or
 * implicitfast, na = 0, no affine actuators, free / hinge / slide joints
 * (the caller checks joint types); a world with more active rows than fit
 * the LDS of one workgroup gets NaN gradients. */
int mjh_adjoint_vjp(const mjhModel* m, const mjhData* d, const float* qpos, const float* qvel, const float* lam_next,
                    const float* gadd, float* lam, float* grad_ctrl, mjhStream stream);
/* The adjoint's tape (no counterpart in the reference): the Data fields
 * mjh_adjoint_vjp reads, saved after a step so that the reverse pass needs
 * no recomputation.  mjh_adjoint_tape_bytes gives one slot's size for (m,
 * d) (the pools at full capacity); mjh_adjoint_tape_save copies the
 * fields of the step d just took into a slot (device memory of that size;
 * pools up to their counters); mjh_adjoint_vjp_taped is mjh_adjoint_vjp
 * reading those fields from the slot instead of d (d gives the sizes). */
int mjh_adjoint_tape_bytes(const mjhModel* m, const mjhData* d, unsigned long long* bytes);
int mjh_adjoint_tape_save(const mjhModel* m, const mjhData* d, void* slot, mjhStream stream);
int mjh_adjoint_vjp_taped(const mjhModel* m, const mjhData* d, const void* slot, const float* qpos, const float* qvel,
                          const float* lam_next, const float* gadd, float* lam, float* grad_ctrl, mjhStream stream);
/* lam [nworld, nx] as gradients with respect to d's raw state: g_qpos
 * [nworld, nq] (a rotation's tangent gradient g becomes 2 q (0, g),
 * orthogonal to q), g_qvel [nworld, nv], g_act [nworld, na]. */
int mjh_adjoint_state_grad(const mjhModel* m, const mjhData* d, const float* lam, float* g_qpos, float* g_qvel,
                           float* g_act, mjhStream stream);

#ifdef __cplusplus
}
#endif

#endif /* MJX_WARP_AMD_H_ */
