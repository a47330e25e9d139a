/*
 * shx_topo.hip -- the topology half of build_cache on the GPU
 * (ShellSystem.C:490-571, 657-741): for every shell of a batch at once,
 * the vertex incidence (nodes_to_elem_map), the validation of the mesh (a
 * closed, consistently oriented 2-manifold whose triangles have at most one
 * vertex of valence != 6, valences 3..6), each element's rotation to its
 * irregular vertex and its 12-node one-ring patch (libMesh's
 * prepare_subdiv_mesh / find_one_ring order, whose regular-case renumbering
 * is SubdivEvaluation.C:10), the node -> (element, slot) patch incidence,
 * and the contact exclusion sets (build_topological_neighborhood,
 * ShellSystem.C:684-741) on the undeformed mesh.
 *
 * Element and node ids are global over the batch (shell ranges padded to
 * kAlign); padded elements hold tin = -1 and padded nodes no incidence.  The
 * ghost triangles of an open mesh (after its real ones) are adjacency only:
 * they close the rings of the boundary vertices and carry no patch.
 * CSR lists are filled through atomic cursors and then sorted per row, so
 * their contents do not depend on the order of the atomics.  Compiled with
 * -ffp-contract=off: the close-set distances round like the host's and the
 * oracle's.
 */
#include <hip/hip_runtime.h>

#include "shx_math.h"
#include "shx_topo.h"

namespace shx {

namespace {

/* third vertex of the triangle that holds the directed edge a -> b (the
 * opposite vertex across it), from a's vertex incidence; -1 if there is
 * none; *count: triangles holding it (1 on a manifold) */
__device__ int opposite(const TopoPtrs& T, int a, int b, int* count)
{
  int c = -1, n = 0;
  for (int q = T.vstart[a]; q < T.vstart[a + 1]; ++q) {
    const int* t = T.tin + 3 * (size_t)T.vinc[q];
    for (int k = 0; k < 3; ++k)
      if (t[k] == a && t[(k + 1) % 3] == b) { c = t[(k + 2) % 3]; ++n; }
  }
  if (count) *count = n;
  return c;
}

__device__ void fail(const TopoPtrs& T, int e, int code)
{
  const int s = T.eshell[e];
  atomicOr(&T.err[2 * s], code);
  atomicMin(&T.err[2 * s + 1], e);
}

/* ascending insertion sort of a short row */
__device__ void sort_row(int32_t* r, int n)
{
  for (int i = 1; i < n; ++i) {
    const int32_t x = r[i];
    int j = i - 1;
    while (j >= 0 && r[j] > x) { r[j + 1] = r[j]; --j; }
    r[j + 1] = x;
  }
}

__device__ void corners(const TopoPtrs& T, int e, v3 P[3])
{
  for (int k = 0; k < 3; ++k) {
    const int n = T.tri[3 * (size_t)e + k];
    P[k] = mk(T.X[3 * (size_t)n], T.X[3 * (size_t)n + 1], T.X[3 * (size_t)n + 2]);
  }
}

} // namespace

/* vertex counts (valence) of every node */
__global__ void __launch_bounds__(256) k_topo_valence(TopoPtrs T)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= T.nE || T.tin[3 * (size_t)e] < 0) return;
  for (int k = 0; k < 3; ++k) atomicAdd(&T.vcnt[T.tin[3 * (size_t)e + k]], 1);
}

/* exclusive prefix sum of n counts into out[0..n] by one block of 1024
 * threads, 1024 values per round in order */
__global__ void __launch_bounds__(1024) k_topo_scan(const int32_t* in, int32_t* out, int n)
{
  __shared__ int32_t s[1024];
  __shared__ int32_t carry;
  const int t = threadIdx.x;
  if (t == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += 1024) {
    const int x = base + t < n ? in[base + t] : 0;
    s[t] = x;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int y = t >= off ? s[t - off] : 0;
      __syncthreads();
      s[t] += y;
      __syncthreads();
    }
    if (base + t < n) out[base + t] = carry + s[t] - x;
    __syncthreads();
    if (t == 1023) carry += s[1023];
    __syncthreads();
  }
  if (t == 0) out[n] = carry;
}

/* vertex incidence rows (ascending element ids after k_topo_vsort) */
__global__ void __launch_bounds__(256) k_topo_vfill(TopoPtrs T)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= T.nE || T.tin[3 * (size_t)e] < 0) return;
  for (int k = 0; k < 3; ++k) {
    const int n = T.tin[3 * (size_t)e + k];
    T.vinc[T.vstart[n] + atomicAdd(&T.vcur[n], 1)] = e;
  }
}

__global__ void __launch_bounds__(256) k_topo_vsort(TopoPtrs T)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= T.nV) return;
  sort_row(T.vinc + T.vstart[n], T.vstart[n + 1] - T.vstart[n]);
}

/* per element: validation, rotation to the irregular vertex, one-ring patch
 * [v0, v1, the ring of v0 from v1 to v2 turning away from v2, then the
 * vertices at lattice offsets (1,1), (2,0), (2,-1), (0,2), (-1,2)] */
__global__ void __launch_bounds__(256) k_topo_patch(TopoPtrs T)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= T.nE || T.tin[3 * (size_t)e] < 0 || T.ghost[e]) return;
  const int* t = T.tin + 3 * (size_t)e;
  int nirr = 0, at = 0;
  for (int k = 0; k < 3; ++k) {
    const int a = t[k], b = t[(k + 1) % 3];
    int c1, c2;
    opposite(T, a, b, &c1);
    opposite(T, b, a, &c2);
    if (c1 != 1) { fail(T, e, TOPO_NONMANIFOLD); return; }
    if (c2 != 1) { fail(T, e, c2 == 0 ? TOPO_OPEN : TOPO_NONMANIFOLD); return; }
    const int val = T.vstart[a + 1] - T.vstart[a];
    if (val != 6) { ++nirr; at = k; }
  }
  if (nirr > 1) { fail(T, e, TOPO_IRREGULAR); return; }
  const int v0 = t[at], v1 = t[(at + 1) % 3], v2 = t[(at + 2) % 3];
  const int N = T.vstart[v0 + 1] - T.vstart[v0];
  if (N < 3 || N > 6) { fail(T, e, TOPO_VALENCE); return; }
  int p[12];
  for (int i = 0; i < 12; ++i) p[i] = -1;
  p[0] = v0;
  p[1] = v1;
  /* every lookup is of an edge of another triangle, validated by that
   * triangle's own thread: a missing one ends the walk here */
  for (int k = 2; k <= N; ++k) {
    p[k] = opposite(T, p[k - 1], v0, nullptr);
    if (p[k] < 0) { fail(T, e, TOPO_OPEN); return; }
  }
  if (p[N] != v2) { fail(T, e, TOPO_RING); return; }
  const int q11 = opposite(T, v2, v1, nullptr);
  if (q11 < 0) { fail(T, e, TOPO_OPEN); return; }
  const int q20 = opposite(T, q11, v1, nullptr);
  if (q20 < 0) { fail(T, e, TOPO_OPEN); return; }
  p[N + 1] = q11;
  p[N + 2] = q20;
  p[N + 3] = opposite(T, q20, v1, nullptr);
  p[N + 4] = opposite(T, v2, q11, nullptr);
  if (p[N + 3] < 0 || p[N + 4] < 0) { fail(T, e, TOPO_OPEN); return; }
  p[N + 5] = opposite(T, v2, p[N + 4], nullptr);
  if (p[N + 5] < 0) { fail(T, e, TOPO_OPEN); return; }
  T.tri[3 * (size_t)e] = v0;
  T.tri[3 * (size_t)e + 1] = v1;
  T.tri[3 * (size_t)e + 2] = v2;
  for (int i = 0; i < 12; ++i) T.patch[12 * (size_t)e + i] = p[i];
  T.valence[e] = (int8_t)N;
  for (int i = 0; i < N + 6; ++i) atomicAdd(&T.icnt[p[i]], 1);
}

/* patch incidence rows: (element << 4 | slot), element order after k_topo_isort */
__global__ void __launch_bounds__(256) k_topo_ifill(TopoPtrs T)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= T.nE || T.tin[3 * (size_t)e] < 0 || T.ghost[e]) return;
  for (int i = 0; i < 12; ++i) {
    const int n = T.patch[12 * (size_t)e + i];
    if (n >= 0) T.inc[T.istart[n] + atomicAdd(&T.icur[n], 1)] = (e << 4) | i;
  }
}

__global__ void __launch_bounds__(256) k_topo_isort(TopoPtrs T)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= T.nV) return;
  sort_row(T.inc + T.istart[n], T.istart[n + 1] - T.istart[n]);
}

/* contact exclusion set of one element (build_topological_neighborhood,
 * ShellSystem.C:684-741): the elements sharing a vertex with it, then the
 * closure of their neighbours that lie within the cutoff of it (triangle
 * distance on the undeformed mesh), sorted.  The set does not depend on the
 * order of the search: a neighbour is taken iff it is within the cutoff and
 * attached to the set.  pass 0 counts, pass 1 writes at cstart. */
__global__ void __launch_bounds__(256) k_topo_close(TopoPtrs T, int pass)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= T.nE || T.tin[3 * (size_t)e] < 0) return;
  const int s = T.eshell[e];
  if (!T.contact[s] || T.ghost[e]) { if (pass == 0) T.ccnt[e] = 0; return; }
  const double cut2 = T.cutoff[s] * T.cutoff[s];
  int32_t set[kTopoCloseMax];
  int n = 0;
  auto has = [&](int c) {
    for (int i = 0; i < n; ++i) if (set[i] == c) return true;
    return false;
  };
  for (int k = 0; k < 3; ++k) {
    const int v = T.tri[3 * (size_t)e + k];
    for (int q = T.vstart[v]; q < T.vstart[v + 1]; ++q) {
      const int c = T.vinc[q];
      if (c != e && !T.ghost[c] && !has(c)) {
        if (n == kTopoCloseMax) { fail(T, e, TOPO_CLOSE_CAP); return; }
        set[n++] = c;
      }
    }
  }
  v3 Pe[3];
  corners(T, e, Pe);
  /* breadth-first over the set as a queue */
  for (int head = 0; head < n; ++head) {
    const int a = set[head];
    for (int k = 0; k < 3; ++k) {
      const int v = T.tri[3 * (size_t)a + k];
      for (int q = T.vstart[v]; q < T.vstart[v + 1]; ++q) {
        const int c = T.vinc[q];
        if (c == e || T.ghost[c] || has(c)) continue;
        v3 Pc[3];
        corners(T, c, Pc);
        tri_hit hit;
        tri_tri(Pe, Pc, hit);
        if (hit.d2 < cut2) {
          if (n == kTopoCloseMax) { fail(T, e, TOPO_CLOSE_CAP); return; }
          set[n++] = c;
        }
      }
    }
  }
  if (pass == 0) { T.ccnt[e] = n; return; }
  sort_row(set, n);
  int32_t* out = T.cids + T.cstart[e];
  for (int i = 0; i < n; ++i) out[i] = set[i];
}

} // namespace shx
