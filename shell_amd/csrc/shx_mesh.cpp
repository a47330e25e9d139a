/*
 * shx_mesh.cpp -- the mesh preparation of the reference's driver on the
 * host (shell.C:146-230): reading and writing OFF triangle meshes, the
 * unit icosphere of a given level (recursive midpoint subdivision of the
 * icosahedron with projection to R = 1, the construction that produced the
 * reference's icosasphere3.off / icosasphere4.off, face order and printed
 * coordinates alike), uniform midpoint refinement (MeshRefinement::
 * uniformly_refine of a triangle mesh), and the driver's projection, scale,
 * rotation and translation of the nodes.
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "shellhip.h"

namespace {

/* midpoint subdivision: every triangle (a, b, c) becomes (a, ab, ca),
 * (b, bc, ab), (c, ca, bc), (ab, bc, ca); edge midpoints are numbered in the
 * order they are first met, after the existing nodes */
void midpoint_refine(std::vector<double>& X, std::vector<int32_t>& T, bool project_unit)
{
  std::map<std::pair<int, int>, int> mid;
  auto midpoint = [&](int a, int b) {
    const std::pair<int, int> key(std::min(a, b), std::max(a, b));
    const auto it = mid.find(key);
    if (it != mid.end()) return it->second;
    double p[3];
    for (int q = 0; q < 3; ++q) p[q] = (X[3 * a + q] + X[3 * b + q]) / 2.0;
    if (project_unit) {
      const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
      for (double& x : p) x /= r;
    }
    const int id = (int)(X.size() / 3);
    X.insert(X.end(), p, p + 3);
    mid[key] = id;
    return id;
  };
  std::vector<int32_t> T2;
  T2.reserve(4 * T.size());
  for (size_t f = 0; f < T.size(); f += 3) {
    const int a = T[f], b = T[f + 1], c = T[f + 2];
    const int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
    const int32_t quad[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
    T2.insert(T2.end(), quad, quad + 12);
  }
  T.swap(T2);
}

/* "%g" printing and reading back: the reference's mesh files carry six
 * significant digits */
double round_g(double x)
{
  char b[64];
  std::snprintf(b, sizeof(b), "%g", x);
  return std::strtod(b, nullptr);
}

/* The ghost layer of an open mesh (the reference's prepare_subdiv_mesh with
 * ghosts, shell.C:228-230: libMesh's MeshTools::Subdivision
 * add_boundary_ghosts, not part of the reference's sources), so that every
 * real element has a complete one-ring patch:
 *  - every boundary edge a -> b of a real triangle (a, b, c) gets the ghost
 *    triangle (b, a, m) with m = a + b - c, c mirrored through the edge's
 *    midpoint;
 *  - every boundary vertex V, with r real triangles, outgoing boundary edge
 *    V -> A (ghost node mA) and incoming B -> V (ghost node mB), gets
 *    k = 4 - r more ghost triangles closing its ring at valence 6: k - 1 new
 *    nodes made by mirroring the last triangle around V (h = V + prev -
 *    prevprev, starting from (A, V, mA)), then (V, h_j, h_j-1) and finally
 *    (V, mB, h_k-1), or (V, mB, mA) when k = 1.
 * Real nodes and triangles keep their numbers; ghost nodes and triangles
 * follow, in the order made (boundary edges in element and side order, then
 * boundary vertices in ascending order).  Returns an empty string or the
 * reason the mesh cannot take a ghost layer. */
std::string add_ghosts(int nv, int ne, const double* X, const int32_t* tri, std::vector<double>& Xo,
                       std::vector<int32_t>& To)
{
  Xo.assign(X, X + 3 * (size_t)nv);
  To.assign(tri, tri + 3 * (size_t)ne);
  std::map<std::pair<int, int>, int> edge;       /* directed edge -> triangle */
  for (int e = 0; e < ne; ++e)
    for (int k = 0; k < 3; ++k) {
      const int a = tri[3 * e + k], b = tri[3 * e + (k + 1) % 3];
      if (a < 0 || a >= nv || b < 0 || b >= nv) return "vertex index out of range";
      if (!edge.emplace(std::make_pair(a, b), e).second) return "non-manifold or inconsistently oriented mesh";
    }
  auto add_node = [&](const double* p) {
    for (int q = 0; q < 3; ++q) Xo.push_back(p[q]);
    return (int)(Xo.size() / 3) - 1;
  };
  std::map<std::pair<int, int>, int> mirror;     /* boundary edge a -> b: its ghost node */
  std::map<int, int> out_to, in_from;            /* boundary vertex: its boundary edges' other ends */
  for (int e = 0; e < ne; ++e)
    for (int k = 0; k < 3; ++k) {
      const int a = tri[3 * e + k], b = tri[3 * e + (k + 1) % 3], c = tri[3 * e + (k + 2) % 3];
      if (edge.count(std::make_pair(b, a))) continue;
      double m[3];
      for (int q = 0; q < 3; ++q) m[q] = Xo[3 * a + q] + Xo[3 * b + q] - Xo[3 * c + q];
      const int g = add_node(m);
      To.push_back(b); To.push_back(a); To.push_back(g);
      mirror[std::make_pair(a, b)] = g;
      if (!out_to.emplace(a, b).second || !in_from.emplace(b, a).second)
        return "the boundary touches itself at a vertex";
    }
  std::vector<int> nreal(nv, 0);
  for (int e = 0; e < ne; ++e)
    for (int k = 0; k < 3; ++k) ++nreal[tri[3 * e + k]];
  for (const auto& kv : out_to) {
    const int V = kv.first, A = kv.second;
    const auto ib = in_from.find(V);
    if (ib == in_from.end()) return "the boundary touches itself at a vertex";
    const int B = ib->second, mA = mirror[std::make_pair(V, A)], mB = mirror[std::make_pair(B, V)];
    const int k = 4 - nreal[V];
    if (k < 1)
      return "a boundary vertex with " + std::to_string(nreal[V]) +
             " elements (more than 3: a reflex corner) takes no regular ghost ring";
    int prev = mA, prevprev = A;
    for (int j = 1; j < k; ++j) {
      double h[3];
      for (int q = 0; q < 3; ++q) h[q] = Xo[3 * V + q] + Xo[3 * prev + q] - Xo[3 * prevprev + q];
      const int hn = add_node(h);
      To.push_back(V); To.push_back(hn); To.push_back(prev);
      prevprev = prev;
      prev = hn;
    }
    To.push_back(V); To.push_back(mB); To.push_back(prev);
  }
  return "";
}

} // namespace

extern "C" {

int shx_mesh_icosphere(int level, int round6, int* nv, int* ne, double* X, int32_t* tri)
{
  if (level < 0 || level > 9 || !nv || !ne) return SHX_E_ARG;
  const int NV = 10 * (1 << (2 * level)) + 2, NE = 20 * (1 << (2 * level));
  *nv = NV;
  *ne = NE;
  if (!X || !tri) return SHX_OK;
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  const double V[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                           {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  const int32_t F[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                            {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                            {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
  std::vector<double> P;
  for (const auto& p : V) {
    const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    for (int q = 0; q < 3; ++q) P.push_back(p[q] / r);
  }
  std::vector<int32_t> T(&F[0][0], &F[0][0] + 60);
  for (int l = 0; l < level; ++l) midpoint_refine(P, T, true);
  for (size_t i = 0; i < P.size(); ++i) X[i] = round6 ? round_g(P[i]) : P[i];
  std::memcpy(tri, T.data(), sizeof(int32_t) * T.size());
  return SHX_OK;
}

int shx_mesh_read_off(const char* text, int* nv, int* ne, double* X, int32_t* tri)
{
  if (!text || !nv || !ne) return SHX_E_ARG;
  std::istringstream in(text);
  std::string tok;
  /* header "OFF" (possibly with counts on the same line), comments '#' */
  auto next = [&](std::string& out) -> bool {
    while (in >> out) {
      if (out[0] == '#') { std::string rest; std::getline(in, rest); continue; }
      return true;
    }
    return false;
  };
  if (!next(tok) || tok.compare(0, 3, "OFF") != 0) return SHX_E_MESH;
  std::string a, b, c;
  if (!next(a) || !next(b) || !next(c)) return SHX_E_MESH;
  const long NV = std::strtol(a.c_str(), nullptr, 10), NF = std::strtol(b.c_str(), nullptr, 10);
  if (NV <= 0 || NF <= 0) return SHX_E_MESH;
  *nv = (int)NV;
  *ne = (int)NF;
  if (!X || !tri) return SHX_OK;
  for (long i = 0; i < 3 * NV; ++i) {
    if (!next(tok)) return SHX_E_MESH;
    char* end = nullptr;
    X[i] = std::strtod(tok.c_str(), &end);
    if (end == tok.c_str()) return SHX_E_MESH;
  }
  for (long f = 0; f < NF; ++f) {
    if (!next(tok)) return SHX_E_MESH;
    if (std::strtol(tok.c_str(), nullptr, 10) != 3) return SHX_E_MESH;   /* triangles only */
    for (int k = 0; k < 3; ++k) {
      if (!next(tok)) return SHX_E_MESH;
      const long id = std::strtol(tok.c_str(), nullptr, 10);
      if (id < 0 || id >= NV) return SHX_E_MESH;
      tri[3 * f + k] = (int32_t)id;
    }
  }
  return SHX_OK;
}

int shx_mesh_write_off(int nv, int ne, const double* X, const int32_t* tri, char* out, size_t cap, size_t* len)
{
  if (nv <= 0 || ne <= 0 || !X || !tri) return SHX_E_ARG;
  std::string s = "OFF\n" + std::to_string(nv) + " " + std::to_string(ne) + " 0\n";
  char b[128];
  for (int i = 0; i < nv; ++i) {
    std::snprintf(b, sizeof(b), "%g %g %g\n", X[3 * i], X[3 * i + 1], X[3 * i + 2]);
    s += b;
  }
  for (int f = 0; f < ne; ++f) {
    std::snprintf(b, sizeof(b), "3 %d %d %d\n", tri[3 * f], tri[3 * f + 1], tri[3 * f + 2]);
    s += b;
  }
  if (len) *len = s.size();
  if (out && cap > s.size()) std::memcpy(out, s.c_str(), s.size() + 1);
  return SHX_OK;
}

int shx_mesh_refine(int nv, int ne, const double* X, const int32_t* tri, int levels, int* nv_out, int* ne_out,
                    double* X_out, int32_t* tri_out)
{
  if (nv <= 0 || ne <= 0 || !X || !tri || levels < 0 || levels > 8 || !nv_out || !ne_out) return SHX_E_ARG;
  std::vector<double> P(X, X + 3 * (size_t)nv);
  std::vector<int32_t> T(tri, tri + 3 * (size_t)ne);
  for (int l = 0; l < levels; ++l) midpoint_refine(P, T, false);
  *nv_out = (int)(P.size() / 3);
  *ne_out = (int)(T.size() / 3);
  if (X_out) std::memcpy(X_out, P.data(), sizeof(double) * P.size());
  if (tri_out) std::memcpy(tri_out, T.data(), sizeof(int32_t) * T.size());
  return SHX_OK;
}

int shx_mesh_transform(int nv, double* X, const double* project, const double* scale, const double* rotate,
                       const double* translate)
{
  if (nv <= 0 || !X) return SHX_E_ARG;
  /* projection onto the ellipsoid / cylinder / plane of radii `project`
   * (zero radii leave their coordinate alone), shell.C:195-208 */
  if (project && (project[0] != 0 || project[1] != 0 || project[2] != 0))
    for (int i = 0; i < nv; ++i) {
      double n[3] = {0, 0, 0};
      for (int q = 0; q < 3; ++q)
        if (project[q] != 0) n[q] = X[3 * i + q] / project[q];
      const double f = 1 / std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      for (int q = 0; q < 3; ++q)
        if (project[q] != 0) X[3 * i + q] *= f;
    }
  /* MeshTools::Modification::scale (shell.C:210-215) */
  if (scale && (scale[0] != 1 || scale[1] != 1 || scale[2] != 1))
    for (int i = 0; i < nv; ++i)
      for (int q = 0; q < 3; ++q) X[3 * i + q] *= scale[q];
  /* MeshTools::Modification::rotate by the Euler angles phi, theta, psi
   * (z, x, z convention; the config gives them in turns, shell.C:217-219):
   * x' = R_z(psi) R_x(theta) R_z(phi) x */
  if (rotate && (std::fmod(rotate[0], 1) != 0 || std::fmod(rotate[1], 1) != 0 || std::fmod(rotate[2], 1) != 0)) {
    const double pi2 = 6.283185307179586476925287;
    const double p = rotate[0] * pi2, t = rotate[1] * pi2, s = rotate[2] * pi2;
    const double cp = std::cos(p), sp = std::sin(p), ct = std::cos(t), st = std::sin(t), cs = std::cos(s),
                 ss = std::sin(s);
    const double R[3][3] = {{cp * cs - sp * ct * ss, sp * cs + cp * ct * ss, st * ss},
                            {-cp * ss - sp * ct * cs, -sp * ss + cp * ct * cs, st * cs},
                            {sp * st, -cp * st, ct}};
    for (int i = 0; i < nv; ++i) {
      const double x[3] = {X[3 * i], X[3 * i + 1], X[3 * i + 2]};
      for (int q = 0; q < 3; ++q) X[3 * i + q] = R[q][0] * x[0] + R[q][1] * x[1] + R[q][2] * x[2];
    }
  }
  /* MeshTools::Modification::translate (shell.C:221-223) */
  if (translate && (translate[0] != 0 || translate[1] != 0 || translate[2] != 0))
    for (int i = 0; i < nv; ++i)
      for (int q = 0; q < 3; ++q) X[3 * i + q] += translate[q];
  return SHX_OK;
}

int shx_mesh_add_ghosts(int nv, int ne, const double* X, const int32_t* tri, int* nv_out, int* ne_out,
                        double* X_out, int32_t* tri_out)
{
  if (nv <= 0 || ne <= 0 || !X || !tri || !nv_out || !ne_out) return SHX_E_ARG;
  std::vector<double> P;
  std::vector<int32_t> T;
  const std::string err = add_ghosts(nv, ne, X, tri, P, T);
  if (!err.empty()) return SHX_E_MESH;
  *nv_out = (int)(P.size() / 3);
  *ne_out = (int)(T.size() / 3);
  if (X_out) std::memcpy(X_out, P.data(), sizeof(double) * P.size());
  if (tri_out) std::memcpy(tri_out, T.data(), sizeof(int32_t) * T.size());
  return SHX_OK;
}

} // extern "C"
