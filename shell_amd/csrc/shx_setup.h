/*
 * shx_setup.h -- host-side tables of libshellhip: the Loop basis at the
 * single quadrature point per valence and the limit-surface weights at an
 * arbitrary point of a patch, and the per-shell copy of the patches the GPU
 * preprocessor (shx_topo.hip) forms, for the host's launch planning.
 */
#ifndef SHX_SETUP_H
#define SHX_SETUP_H

#include <cstdint>
#include <string>
#include <vector>

namespace shx {

constexpr int kMaxPatch = 12;   /* patch size for valence 6; valence 5 gives 11 */

/* one shell's patches, downloaded from the GPU preprocessor */
struct Topology {
  int nv = 0, ne = 0;
  std::vector<int32_t> tri;          /* ne*3, rotated so the irregular vertex is node 0 */
  std::vector<int32_t> patch;        /* ne*12 element-major, -1 padded */
  std::vector<uint8_t> valence;      /* valence of node 0 of each element */
};

/* Loop basis of a patch with an irregular vertex of the given valence (6:
 * regular) at (xi, eta) = (1/3, 1/3): out[k*12 + j], k = value, d/dxi,
 * d/deta, d2/dxi2, d2/deta2, d2/dxideta, j over the patch (zero padded). */
bool loop_basis_at_centroid(int valence, double out[6 * kMaxPatch]);

/* Loop limit-surface weights of a patch at barycentric (xi, eta) of its
 * element (SubdivEvaluation::evaluate / evaluate_deriv, SubdivEvaluation.C:
 * 14-167): w[0] position, w[1] d/dxi, w[2] d/deta over the patch slots.
 * A regular patch uses the box splines directly; an irregular one
 * subdivides n times toward its irregular vertex (Stam 1998) and evaluates
 * the box splines of the sub-patch that holds the point. */
bool subdiv_weights(int valence, double xi, double eta, double w[3][kMaxPatch]);

/* the limit-surface tables of shx_smooth.h (kLimitTabSize doubles: box
 * spline monomials, then P_k Ab A^(n-1) per valence 3..5, k, n), built on
 * first use */
const double* limit_table();


} // namespace shx
#endif
