/*
 * shx_topo.h -- the GPU topology preprocessor (shx_topo.hip): vertex and
 * patch incidence, mesh validation, one-ring patches and contact exclusion
 * sets of every shell of a batch, in global element and node ids.
 */
#ifndef SHX_TOPO_H
#define SHX_TOPO_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace shx {

constexpr int kTopoCloseMax = 128;   /* contact exclusion set of one element (capacity) */

/* error bits per shell (TopoPtrs::err[2 s]); err[2 s + 1] the first element */
enum {
  TOPO_NONMANIFOLD = 1,   /* a directed edge held by more than one triangle */
  TOPO_OPEN = 2,          /* an edge without its reverse: the mesh is not closed */
  TOPO_IRREGULAR = 4,     /* a triangle with more than one vertex of valence != 6 */
  TOPO_VALENCE = 8,       /* an irregular vertex of valence outside 3..6 */
  TOPO_RING = 16,         /* the one-ring walk did not close */
  TOPO_CLOSE_CAP = 32     /* a contact exclusion set larger than kTopoCloseMax */
};

struct TopoPtrs {
  int nE, nV;                 /* padded global element / node counts */
  const int32_t* eshell;      /* nE */
  const int32_t* tin;         /* 3*nE: input triangles, global node ids; -1 for padding */
  const int8_t* ghost;        /* nE: 1 for a ghost triangle of an open mesh's ghost layer: part of the
                                 adjacency (valences, rings), without a patch or an exclusion set, and
                                 never in another element's exclusion set */
  const double* X;            /* 3*nV: undeformed node positions */
  const int32_t* contact;     /* per shell: self-contact on (exclusion sets needed) */
  const double* cutoff;       /* per shell: (1 + neighbor_margin) h */
  int32_t* err;               /* 2 per shell */
  int32_t *vcnt, *vcur, *vstart, *vinc;     /* vertex incidence */
  int32_t* tri;               /* 3*nE: rotated triangles (irregular vertex first) */
  int32_t* patch;             /* 12*nE element-major, -1 padded */
  int8_t* valence;            /* nE */
  int32_t *icnt, *icur, *istart, *inc;      /* patch incidence (element << 4 | slot) */
  int32_t *ccnt, *cstart, *cids;            /* contact exclusion sets */
};

__global__ void k_topo_valence(TopoPtrs T);
__global__ void k_topo_scan(const int32_t* in, int32_t* out, int n);
__global__ void k_topo_vfill(TopoPtrs T);
__global__ void k_topo_vsort(TopoPtrs T);
__global__ void k_topo_patch(TopoPtrs T);
__global__ void k_topo_ifill(TopoPtrs T);
__global__ void k_topo_isort(TopoPtrs T);
__global__ void k_topo_close(TopoPtrs T, int pass);

} // namespace shx
#endif
