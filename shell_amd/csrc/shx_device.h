/*
 * shx_device.h -- device-side data layout of libshellhip.
 *
 * One shell per context.  Everything lives in HBM:
 *  - element tables, structure of arrays [k][ne] so that one thread per
 *    element reads consecutive addresses: patch ids (12 x int32), valence
 *    (int8), and 29 FP64 rows (R00 R01 R10 R11, JxW, A1 A2 A11 A22 A12,
 *    ff1_ref, ff2_ref, xyz);
 *  - the Loop basis at the quadrature point for each valence (6 x 12 FP64,
 *    staged in LDS by the kernels that use it);
 *  - node state u, v, a twice (ping-pong; DShell::par selects the current
 *    copy), node-interleaved xyz FP64;
 *  - per-element generalised stresses G (18 rows) written by the element
 *    kernel and gathered by the node kernel through the node incidence list,
 *    so no force is ever scattered with atomics;
 *  - the contact pair list (sorted by eid1 then eid2, ping-pong) and its
 *    per-pair friction state.
 */
#ifndef SHX_DEVICE_H
#define SHX_DEVICE_H

#include <stdint.h>

namespace shx {

/* element table rows */
enum {
  T_R00 = 0, T_R01, T_R10, T_R11, T_JXW,
  T_A1 = 5, T_A2 = 8, T_A11 = 11, T_A22 = 14, T_A12 = 17,
  T_FF1 = 20, T_FF2 = 23, T_XYZ = 26, T_ROWS = 29
};

/* generalised stress rows: G0 = JxW a3 (pressure), G1, G2 (first-derivative
 * terms), G3, G4, G5 (second-derivative terms) */
enum { G_ROWS = 18 };

/* per-shell constants (kernel argument) */
struct SConst {
  int nv, ne;
  double h, rho;
  double E1, E2, nu, G12, nu2;     /* nu2 = E1 nu / E2 (ShellSystem.C:1199) */
  double K, D;
  double stiff_density, stiff_static_friction, d2_friction;
  double pressure_final, pressure_ramp_time;
  double bulk_modulus, volume_final_ratio, volume_ramp_time;
  double contact_stiffness, mu_s, mu_d, self_damping, self_friction_damping;
  double friction_stop_velocity;
  double cell_size_mult, cell_margin;
  double beta, gamma, dt_min, dt_max, rel_loc_err_min, rel_loc_err_max;
  double conv_tol, load_end_time;
  int adapt_dt_every, contact, self_friction, quirk_fulltime, quirk_volume_schedule;
  int stop_on_converged;
  unsigned n_cells_max;
};

/* per-shell mutable scalars (device memory, written by the finalize and
 * broad-phase kernels) */
struct DShell {
  double t, dt, nm[10];
  double du_sum, avg_elem_size;
  double volume, volume0, area;
  double bmin[3], bmax[3];
  double cell_size, d_max, d2_max;
  double residual, fnorm, fext_norm, e_tens, e_bend, e_self, e_fric, e_kin, du_max;
  int par, fresh, rebuild, force_rebuild;
  int problem, err;
  int n_steps, n_rejected, n_rebuilds, converged_step, n_contacts;
  int nd[3], n_cells;
  int n_pairs, pair_buf, n_entries;
  int steps_target, done, attempts;
};

/* error bits in DShell::err */
enum { ERR_ENTRIES = 1, ERR_PAIRS = 2, ERR_ROW = 4, ERR_CELLS = 8 };

/* device pointers (kernel argument) */
struct DPtrs {
  const double* X;            /* nv*3 */
  const int32_t* pidx;        /* 12*ne */
  const int8_t* vcls;         /* ne: valence */
  const int32_t* tri;         /* 3*ne, SoA [k][ne] */
  double* tab;                /* T_ROWS*ne */
  const double* rawb;         /* [4 valences 3..6][6][12] */
  const int32_t* inc_start;   /* nv+1 */
  const int32_t* inc;         /* e*16+i */
  double* minv;               /* nv */
  double* cdamp;              /* nv */
  double* U[2];
  double* V[2];
  double* A[2];
  double* G;                  /* G_ROWS*ne */
  double* area;               /* ne */
  double* fc;                 /* nv*3 contact force accumulator */
  double* fout;               /* nv*3 debug outputs (forces mode) */
  double* fext_out;
  double* epart;              /* element kernel block partials */
  double* npart;              /* node kernel block partials */
  double* cpart;              /* contact kernel block partials */
  DShell* sh;
  /* contact */
  const int32_t* close_start; /* ne+1 */
  const int32_t* close_ids;
  double* ebox;               /* 6*ne, SoA */
  double* bpart;              /* broad-phase tile partials: 7 per tile */
  int32_t* ecell;             /* 6*ne: cell index ranges lo[3], hi[3] */
  uint32_t* cell_cnt;         /* n_cells_max+1 */
  uint32_t* cell_start;       /* n_cells_max+1 */
  int32_t* cell_elems;        /* entry capacity */
  uint32_t* scan_tmp;         /* block sums for scans */
  int32_t* found_cnt;         /* ne+1 */
  int32_t* found_e2;          /* ne*ROWCAP */
  double* found_d2;           /* ne*ROWCAP */
  int32_t* row_start[2];      /* ne+1, per pair buffer */
  int32_t* pe1[2];
  int32_t* pe2[2];
  double* pd2[2];
  int32_t* pstick[2];
  double* pso[2];             /* 3 per pair */
  int ent_cap, pair_cap;
  unsigned cell_cap;
};

constexpr int kRowCap = 64;      /* new candidates per element and rebuild */

} // namespace shx
#endif
