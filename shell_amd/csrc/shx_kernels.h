/*
 * shx_kernels.h -- kernel entry points of libshellhip (defined in
 * shx_kernels.hip, launched from shx_api.hip).
 */
#ifndef SHX_KERNELS_H
#define SHX_KERNELS_H

#include <hip/hip_runtime.h>

#include "shx_device.h"

namespace shx {

/* kernel modes: a full step attempt (predictor from the current state), a
 * plain force evaluation at the current u (shx_forces), or the initial
 * acceleration a = M^-1 f (apply_initial_conditions) */
enum { M_STEP = 0, M_FORCES = 1, M_INIT = 2 };

/* element-force precision (values of SHX_FP64 / SHX_FP32_SPLIT) */
enum { SHX_PREC_FP64 = 0, SHX_PREC_FP32 = 1 };

__global__ void k_precompute(DPtrs P);
__global__ void k_mass(DPtrs P);
__global__ void k_mass_grow(DPtrs P);
__global__ void k_grow(DPtrs P, int mode);
__global__ void k_elem_grow(DPtrs P, int mode);
__global__ void k_elem_w3(DPtrs P, int mode);
__global__ void k_elem_w4(DPtrs P, int mode);
__global__ void k_fin_elem(DPtrs P, int mode);
__global__ void k_contact_w2(DPtrs P, int mode);
__global__ void k_contact_w3(DPtrs P, int mode);
__global__ void k_contact_w4(DPtrs P, int mode);
__global__ void k_node_w4(DPtrs P, int mode);
__global__ void k_node_w6(DPtrs P, int mode);
__global__ void k_node_w8(DPtrs P, int mode);
__global__ void k_node_loads(DPtrs P, int mode);
__global__ void k_node_bc(DPtrs P, int mode);
__global__ void k_bc_strong(DPtrs P);
__global__ void k_precompute32(DPtrs P);
__global__ void k_elem32(DPtrs P, int mode);
__global__ void k_fin_node(DPtrs P, int mode);
__global__ void k_kd(DPtrs P);
__global__ void k_rot(DPtrs P, int mode);
__global__ void k_bp_box(DPtrs P, int posmode, int list);
__global__ void k_bp_grid(DPtrs P, int list);
__global__ void k_bp_count(DPtrs P, int posmode, int list);
__global__ void k_bp_scan(DPtrs P, int list);
__global__ void k_bp_scatter(DPtrs P, int list);
__global__ void k_bp_pairs_l8(DPtrs P, int posmode, int list);
__global__ void k_bp_pairs_l16(DPtrs P, int posmode, int list);
__global__ void k_bp_pairs_l32(DPtrs P, int posmode, int list);
__global__ void k_bp_exact_w2(DPtrs P, int posmode, int list);
__global__ void k_bp_exact_w3(DPtrs P, int posmode, int list);
__global__ void k_bp_exact_w4(DPtrs P, int posmode, int list);
__global__ void k_contact_smooth(DPtrs P, int mode);
__global__ void k_bp_exact_smooth(DPtrs P, int posmode, int list);
__global__ void k_bp_smooth_new(DPtrs P, int posmode, int list);
__global__ void k_bp_host_plan(DPtrs P);
__global__ void k_bp_host_done(DPtrs P);

} // namespace shx
#endif
