/*
 * shell_oracle.h -- C API of the CPU FP64 oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle restates, on the CPU in FP64, the
 * explicit force-and-step path of the Loop-subdivision Koiter shell solver
 * (reference: ShellSystem.C, ShellShellContact.C, Friction.C,
 * SubdivEvaluation.{h,C}).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or the
 * timed CPU baseline.  The product library (libshellhip) never links it.
 *
 * Layouts: node vectors are FP64, node-interleaved xyz (index 3*n+c);
 * triangles are int32 triples; element ids are triangle indices.
 */
#ifndef SHELL_ORACLE_H
#define SHELL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physical and numerical parameters.  Field meanings follow the reference's
 * config keys (ShellSystem.C:170-418); the build-defined extensions are
 * marked "ext". */
typedef struct {
  double E1, E2, nu, G12;          /* shell_young_modulus(2), shell_poisson_ratio, shell_shear_modulus */
  double thickness, rho;           /* thickness, shell_mass_density */
  double memb_stiff_mult, bend_stiff_mult;
  double damping_alpha, damping_beta, damping_gamma;
  double pressure_final;           /* ext: p(t) = pressure_final*min(t/pressure_ramp_time,1), LOCAL base */
  double pressure_ramp_time;       /* ext: <= 0 means constant pressure_final */
  double bulk_modulus;             /* volume pressure kappa*(V0 r(t)/V - 1) */
  double volume_final_ratio;       /* ext: r(t) = 1-(1-ratio)*min(t/volume_ramp_time,1) */
  double volume_ramp_time;
  double contact_stiffness;        /* ext: >0 replaces stiff_density*area_min */
  double mu_s, mu_d;
  double self_damping, self_friction_damping, fric_stiff_mult;
  double friction_stop_velocity, static_friction_tol;
  double cell_size_mult, cell_margin, neighbor_margin;
  double newmark_beta, newmark_gamma, dt;
  double dt_min, dt_max;           /* set_newmark_parameters clamp (ShellSystem.C:746-747); 0: none */
  double rel_loc_err_min, rel_loc_err_max;
  int32_t adapt_dt_every;          /* 0: constant dt (ShellSystem.C:957) */
  int32_t contact;                 /* 0: none, 1: linear, 2: smooth (ShellSystem.h:45) */
  int32_t quirk_fulltime;          /* 1: pressure *= fulltime (== 0), ShellSystem.C:1480 */
  int32_t quirk_volume_schedule;   /* 1: r(t) = max(1-floor(5t)*0.01, 0.5), ShellSystem.C:1474 */
  int32_t omp_faithful;            /* 1: OpenMP loops with critical assembly (CPU baseline timing) */
  int32_t volume_at_prediction;    /* ext: 1: enclosed volume of the force-evaluation configuration
                                      instead of the previous step's (ShellSystem.C:1097, 2505) */
  int32_t rebuild_trigger;         /* ext: 0: du_sum >= tolerance (ShellSystem.C:1917-1919); 1: 2 x the
                                      largest node displacement since the list's build >= tolerance */
  /* the loads of compute_external_force beyond the pressure (ShellSystem.C:1587-1786) */
  int32_t cavity_contact;          /* shell_cavity_contact */
  int32_t cavity;                  /* enum Cavity (ShellSystem.h:43): 0 box, 1 cylinder, 2 ellipsoid */
  int32_t gravity_dir;             /* coordinate of the gravity load */
  int32_t pad0;
  double cavity_R0[3];             /* Rx, Ry, Rz at t = 0 */
  double cavity_rate;              /* R(t) = R(0) exp(rate t), the form of the reference's config.cfg
                                      (Rx = 0.1*exp(-4*t)); 0: constant */
  double mu_s_cavity, mu_d_cavity, cavity_damping, cavity_friction_damping;
  double gravity, damping_rotation;
  /* in-plane growth (apply_growth, ShellSystem.C:2295-2459): static growth factors of the form
     g_k(x, y, z, t) = g0 + gt t + gx x + gy y + gz z at the element's reference point, or dynamic growth
     by the rates; growth_base 0 local, 1 cartesian, 2 cylindrical */
  int32_t growth_base, thickness_dir, grow_mass, pad1;
  double growth_g0[2], growth_gt[2], growth_gx[2], growth_gy[2], growth_gz[2];
  double metric_growth_rate, curvature_growth_rate;
  /* open meshes (ShellSystem.C:576-651, 2138-2292): the last ghost_elems triangles are the ghost layer;
     shell_boundary 0 clamped, 1 clamped weak, 2 pinned, 3 pinned weak, 4 free, 5 free weak */
  int32_t ghost_elems, shell_boundary, bc_coords, bc_every;
  double bc_stiff_mult;
} orc_params;

typedef struct {
  double t;            /* time after the step */
  double residual;     /* ||f_int+f_ext+f_c|| / ||f_ext|| at the step's force evaluation */
  double fnorm, fext_norm;
  double e_tens, e_bend, e_self, e_fric, e_kin;
  double volume, area;
  double du_max;
  int32_t n_contacts;  /* pairs with depth > 0 */
  int32_t n_pairs;     /* candidate pairs in the list */
  int32_t n_rebuilds;  /* broad-phase rebuilds so far */
  int32_t problem;     /* max severity (ShellSystem.h:48-49) */
  int32_t n_rejected;  /* rejected step attempts so far */
  double dt;           /* dt for the next step */
  double residual_deform; /* residual with the rigid-body inertial forces removed (least squares) */
  double e_cav;        /* cavity contact energy */
  int32_t n_cav;       /* nodes in cavity contact (depth > 0) */
  int32_t pad;
} orc_report;

/* meshes */
int  orc_icosphere_counts(int level, int* nv, int* ne);
int  orc_icosphere(int level, int round_g, double* X, int32_t* tri);

/* system */
void* orc_create(int nv, int ne, const double* X, const int32_t* tri, const orc_params* prm);
void  orc_destroy(void* h);
const char* orc_last_error(void);
/* the BoundaryCache list (n1 n2 n3 n4 per edge, ShellSystem.C:603-651) before the bc_criterion filter
   (nodes may be null: the count); orc_bc_keep keeps the edges whose flag is set, in order */
int orc_boundary_edges(void* h, int32_t* nodes);
int orc_bc_keep(void* h, const int8_t* keep);

int  orc_patch_size(void* h, int32_t* N /* ne */);
int  orc_patches(void* h, int32_t* idx /* ne*12, -1 padded */, int32_t* tri_reordered /* ne*3 */);
/* per element, 12-padded: phi, dphi1, dphi2, d2phi11, d2phi22, d2phi12 (6*12) then
 * A1,A2,A11,A22,A12 (15), ff1_ref (3), ff2_ref (3), JxW (1), R00,R01,R10,R11 (4), xyz (3):
 * 72+29 = 101 doubles */
int  orc_elem_tables(void* h, double* out /* ne*101 */);
int  orc_mass(void* h, double* Minv, double* C);

int  orc_set_state(void* h, const double* u, const double* v, const double* a, double t);
int  orc_get_state(void* h, double* u, double* v, double* a, double* t);

/* forces at displacement u_pred with velocities v (contact uses v), time t.
 * Writes f (undamped total), f_ext (pressure+volume+contact), energies[4] =
 * (E_tens, E_bend, E_self, E_fric); elem_area (optional). Uses the current
 * contact candidate list and updates its friction state, like the reference. */
int  orc_forces(void* h, const double* u_pred, const double* v, double t,
                double* f, double* f_ext, double* energies, double* elem_area);

/* internal force only (no external load, no contact): used by the CFL estimate and tests */
int  orc_internal_force(void* h, const double* u_pred, double* f_int);

int  orc_init_dynamics(void* h);       /* apply_initial_conditions without distortion */
int  orc_step(void* h, int n_steps, orc_report* rep);
int  orc_measure(void* h);
int  orc_properties(void* h, double* out /* 16 */);
/* per node, from the last force evaluation: the cavity spring origin (3),
 * the cavity energy and friction energy over the node's mean element area
 * (the .vtu fields ecav and efric, ShellSystem.C:1626-1706) */
int  orc_node_cavity(void* h, double* origin /* nv*3 */, double* ecav /* nv */, double* efric /* nv */);
/* per element: the growth vector G (3), the Voigt inverse growth tensor GV (9, row-major) and the last
 * strains alpha, beta (6): 18 values */
int  orc_elem_growth(void* h, double* out /* ne*18 */);
int  orc_elem_measure(void* h, double* out /* ne*10: ff1, ff2, jac, kGauss, kMean, area */);
int  orc_contact_rebuild(void* h, int force);
int  orc_pairs(void* h, uint32_t* e1, uint32_t* e2, int cap);   /* returns count, sorted */
int  orc_pair_state(void* h, double* d2, int32_t* stick, double* spring_origin, int cap);
/* smooth contact: per pair (same order) v1 w1 v2 w2 of the last closest points */
int  orc_pair_coords(void* h, double* vw, int cap);
/* SmoothShellShellContact(e1, e2) + compute() at the state's positions from
 * x (4, in/out): out[8] = d2, dc[3], point()[3], Newton iterations */
int  orc_smooth_contact(void* h, int e1, int e2, double* x, double* out);
/* the limit surface of a patch of the given valence (points X, 3 per patch
 * slot) at (xi, eta): out[18] = point, d/dxi, d/deta, xixi, etaeta, xieta */
void orc_limit_point(int valence, const double* X, double xi, double eta, double* out);
/* closest points of two limit-surface patches (points X1, X2) from x (4,
 * in/out): out[5] = d2, dc[3], Newton iterations */
void orc_smooth_closest(int valence1, const double* X1, int valence2, const double* X2, double* x, double* out);
int  orc_close_counts(void* h, int32_t* counts);
int  orc_close_sets(void* h, int32_t* start /* ne+1 */, int32_t* ids, int cap);
double orc_cfl_omega_max(void* h, int iters, uint64_t seed);

/* primitives (for tests) */
void orc_stam_basis(double v, double w, double* out /* 6*12: b, bv, bw, bvv, bww, bvw */);
int  orc_raw_elem_basis(int valence, double* out /* 6*(valence+6) in (xi,eta) */);
void orc_loop_subdiv_rows(int valence, double* S /* 12*(valence+6) */);
void orc_tri_distance(const double* p /* 18: e1 v0..v2, e2 v0..v2 */, double* out /* d2, dc[3], v[2], w[2] */);
void orc_segment_distance(const double* p /* 12 */, double* out /* diff[3], s, t */);
void orc_point_triangle(const double* p /* 12 */, double* out /* diff[3], v, w */);
void orc_friction_apply(double* io /* see oracle_geom.h */);
double orc_contact_force(double k, double h, double d);
double orc_contact_energy(double k, double h, double d);

#ifdef __cplusplus
}
#endif
#endif
