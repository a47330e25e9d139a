/*
 * shx_setup.h -- host-side mesh preparation of libshellhip: subdivision
 * patches, Loop basis at the single quadrature point, node incidence lists
 * and topological contact exclusion.  Integer/graph work done once per shell;
 * the per-element geometry (reinit_mod, reference metrics, masses) is
 * computed on the GPU from what this produces.
 */
#ifndef SHX_SETUP_H
#define SHX_SETUP_H

#include <cstdint>
#include <string>
#include <vector>

namespace shx {

constexpr int kMaxPatch = 12;   /* patch size for valence 6; valence 5 gives 11 */

struct Topology {
  int nv = 0, ne = 0;
  std::vector<int32_t> tri;          /* ne*3, rotated so the irregular vertex is node 0 */
  std::vector<int32_t> patch;        /* ne*12 element-major, -1 padded */
  std::vector<uint8_t> valence;      /* valence of node 0 of each element */
  /* node -> (element, local patch index) entries, encoded e*16+i, element order */
  std::vector<int32_t> inc_start, inc;
  /* node -> incident triangles (vertex incidence) */
  std::vector<int32_t> vtx_start, vtx;
};

/* Builds patches in the one-ring order whose regular-case renumbering is
 * SubdivEvaluation.C:10: [v0, ring of v0 from v1 to v2 turning away from v2,
 * then the vertices at lattice offsets (1,1), (2,0), (2,-1), (0,2), (-1,2)].
 * Fails (returns false, err set) unless the mesh is a closed oriented
 * 2-manifold whose triangles have at most one vertex of valence != 6. */
bool build_topology(int nv, int ne, const int32_t* tri_in, Topology& T, std::string& err);

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

/* elements whose distance on the undeformed mesh is below cutoff, closed
 * under neighbourhood (build_topological_neighborhood, ShellSystem.C:684-741);
 * CSR sorted by element id. */
void build_close_sets(const Topology& T, const double* X, double cutoff, std::vector<int32_t>& start,
                      std::vector<int32_t>& ids);

} // namespace shx
#endif
