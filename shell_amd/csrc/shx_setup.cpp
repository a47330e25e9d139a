/*
 * shx_setup.cpp -- host-side Loop basis tables of libshellhip (see
 * shx_setup.h).
 */
#include "shx_setup.h"
#include "shx_math.h"
#include "shx_smooth.h"

#include <algorithm>
#include <initializer_list>
#include <cmath>
#include <vector>

namespace shx {

namespace {

/* Twelve quartic box splines of the regular Loop patch (J. Stam, SIGGRAPH
 * 1998 course notes; libMesh FESubdiv::regular_shape) as 12*B(u,v,w) with
 * u = 1-v-w: entries {coefficient, power of u, power of v, power of w}.
 * Box-spline order: lattice offsets (0,-1), (-1,0), (1,-1), (0,0), (-1,1),
 * (2,-1), (1,0), (0,1), (-1,2), (2,0), (1,1), (0,2) for a triangle with
 * corners (0,0), (1,0), (0,1). */
struct BT { int8_t c, pu, pv, pw; };
/* the twelve term lists, each closed by a zero entry */
const BT kBoxTerms[] = {
  {1, 4, 0, 0}, {2, 3, 1, 0}, {0, 0, 0, 0},
  {1, 4, 0, 0}, {2, 3, 0, 1}, {0, 0, 0, 0},
  {1, 4, 0, 0}, {2, 3, 0, 1}, {6, 3, 1, 0}, {6, 2, 1, 1}, {12, 2, 2, 0}, {6, 1, 2, 1}, {6, 1, 3, 0}, {2, 0, 3, 1}, {1, 0, 4, 0}, {0, 0, 0, 0},
  {6, 4, 0, 0}, {24, 3, 0, 1}, {24, 2, 0, 2}, {8, 1, 0, 3}, {1, 0, 0, 4}, {24, 3, 1, 0}, {60, 2, 1, 1}, {36, 1, 1, 2}, {6, 0, 1, 3}, {24, 2, 2, 0}, {36, 1, 2, 1}, {12, 0, 2, 2}, {8, 1, 3, 0}, {6, 0, 3, 1}, {1, 0, 4, 0}, {0, 0, 0, 0},
  {1, 4, 0, 0}, {6, 3, 0, 1}, {12, 2, 0, 2}, {6, 1, 0, 3}, {1, 0, 0, 4}, {2, 3, 1, 0}, {6, 2, 1, 1}, {6, 1, 1, 2}, {2, 0, 1, 3}, {0, 0, 0, 0},
  {2, 1, 3, 0}, {1, 0, 4, 0}, {0, 0, 0, 0},
  {1, 4, 0, 0}, {6, 3, 0, 1}, {12, 2, 0, 2}, {6, 1, 0, 3}, {1, 0, 0, 4}, {8, 3, 1, 0}, {36, 2, 1, 1}, {36, 1, 1, 2}, {8, 0, 1, 3}, {24, 2, 2, 0}, {60, 1, 2, 1}, {24, 0, 2, 2}, {24, 1, 3, 0}, {24, 0, 3, 1}, {6, 0, 4, 0}, {0, 0, 0, 0},
  {1, 4, 0, 0}, {8, 3, 0, 1}, {24, 2, 0, 2}, {24, 1, 0, 3}, {6, 0, 0, 4}, {6, 3, 1, 0}, {36, 2, 1, 1}, {60, 1, 1, 2}, {24, 0, 1, 3}, {12, 2, 2, 0}, {36, 1, 2, 1}, {24, 0, 2, 2}, {6, 1, 3, 0}, {8, 0, 3, 1}, {1, 0, 4, 0}, {0, 0, 0, 0},
  {2, 1, 0, 3}, {1, 0, 0, 4}, {0, 0, 0, 0},
  {2, 0, 3, 1}, {1, 0, 4, 0}, {0, 0, 0, 0},
  {2, 1, 0, 3}, {1, 0, 0, 4}, {6, 1, 1, 2}, {6, 0, 1, 3}, {6, 1, 2, 1}, {12, 0, 2, 2}, {2, 1, 3, 0}, {6, 0, 3, 1}, {1, 0, 4, 0}, {0, 0, 0, 0},
  {1, 0, 0, 4}, {2, 0, 1, 3}, {0, 0, 0, 0},
};

const BT* box_terms(int k)
{
  const BT* t = kBoxTerms;
  for (int q = 0; q < k; ++q) { while (t->c != 0) ++t; ++t; }
  return t;
}

/* expand 12*B into monomials v^a w^b (a+b <= 4) with exact integer
 * arithmetic: u^k = sum binom(k;i,j) (-v)^i (-w)^j */
void box_monomials(int k, long long M[5][5])
{
  for (int a = 0; a < 5; ++a) for (int b = 0; b < 5; ++b) M[a][b] = 0;
  static const long long binom[5][5] = {{1}, {1, 1}, {1, 2, 1}, {1, 3, 3, 1}, {1, 4, 6, 4, 1}};
  for (const BT* t = box_terms(k); t->c != 0; ++t)
    for (int i = 0; i <= t->pu; ++i)
      for (int j = 0; i + j <= t->pu; ++j) {
        /* multinomial pu!/(i! j! (pu-i-j)!) = binom(pu,i)*binom(pu-i,j) */
        long long m = binom[t->pu][i] * binom[t->pu - i][j];
        if ((i + j) & 1) m = -m;
        M[t->pv + i][t->pw + j] += t->c * m;
      }
}

/* value and the five derivatives of the 12 box splines at (v, w) */
void box_eval(double v, double w, double out[6][12])
{
  double pv[5], pw[5];
  pv[0] = pw[0] = 1;
  for (int q = 1; q < 5; ++q) { pv[q] = pv[q - 1] * v; pw[q] = pw[q - 1] * w; }
  auto P = [](const double* p, int e) { return e < 0 ? 0.0 : p[e]; };
  for (int k = 0; k < 12; ++k) {
    long long M[5][5];
    box_monomials(k, M);
    double r[6] = {0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 5; ++a)
      for (int b = 0; a + b < 5; ++b) {
        if (!M[a][b]) continue;
        const double c = (double)M[a][b];
        r[0] += c * P(pv, a) * P(pw, b);
        r[1] += c * a * P(pv, a - 1) * P(pw, b);
        r[2] += c * b * P(pv, a) * P(pw, b - 1);
        r[3] += c * a * (a - 1) * P(pv, a - 2) * P(pw, b);
        r[4] += c * b * (b - 1) * P(pv, a) * P(pw, b - 2);
        r[5] += c * a * b * P(pv, a - 1) * P(pw, b - 1);
      }
    for (int q = 0; q < 6; ++q) out[q][k] = r[q] / 12.0;
  }
}

} // namespace

bool loop_basis_at_centroid(int valence, double out[6 * kMaxPatch])
{
  if (valence < 3 || valence > 6) return false;
  const int N = valence, P = N + 6;
  std::fill(out, out + 6 * kMaxPatch, 0.0);
  double b[6][12];
  const double c = 1.0 / 3.0;
  if (N == 6) {
    /* patch slot -> box spline: v0 (0,0), v1 (1,0), ring (1,-1), (0,-1),
     * (-1,0), (-1,1), v2 (0,1), then (1,1), (2,0), (2,-1), (0,2), (-1,2) */
    static const int slot_box[12] = {3, 6, 2, 0, 1, 4, 7, 10, 9, 5, 11, 8};
    box_eval(c, c, b);
    for (int q = 0; q < 6; ++q)
      for (int j = 0; j < 12; ++j) out[q * kMaxPatch + j] = b[q][slot_box[j]];
    return true;
  }
  /* Irregular v0: one Loop refinement; the centroid is the centroid of the
   * central sub-triangle, whose 12-point patch is regular.  Sub-triangle
   * parameters (1-2xi, 1-2eta); one factor -2 per derivative.
   * Fine points as combinations of patch slots. */
  const int V0 = 0, V1 = 1, V2 = N, Q11 = N + 1, Q20 = N + 2, Q2m = N + 3, Q02 = N + 4, Qm2 = N + 5;
  const int Ra = 2, Rb = std::min(3, N), Rz = N - 1, Ry = std::max(N - 2, 1);
  double S[12][kMaxPatch] = {};
  auto E = [&](int row, int a, int bb, int cc, int d) { /* edge point */
    S[row][a] += 0.375; S[row][bb] += 0.375; S[row][cc] += 0.125; S[row][d] += 0.125;
  };
  auto R6 = [&](int row, int ctr, std::initializer_list<int> nb) { /* regular vertex point */
    S[row][ctr] += 0.625;
    for (int q : nb) S[row][q] += 0.0625;
  };
  E(0, V2, Q11, V1, Q02);
  E(1, V1, Q11, V2, Q20);
  R6(2, V2, {Q11, Q02, Qm2, Rz, V0, V1});
  E(3, V1, V2, V0, Q11);
  R6(4, V1, {Q20, Q11, V2, V0, Ra, Q2m});
  E(5, V2, Rz, V0, Qm2);
  E(6, V0, V2, V1, Rz);
  E(7, V0, V1, V2, Ra);
  E(8, V1, Ra, V0, Q2m);
  E(9, V0, Rz, V2, Ry);
  {
    const double cc = 0.375 + 0.25 * std::cos(2.0 * M_PI / N);
    const double beta = (0.625 - cc * cc) / N; /* Loop (1987) */
    S[10][V0] += 1.0 - N * beta;
    for (int k = 1; k <= N; ++k) S[10][k] += beta;
  }
  E(11, V0, Ra, V1, Rb);
  box_eval(1 - 2 * c, 1 - 2 * c, b);
  const double f[6] = {1, -2, -2, 4, 4, 4};
  for (int q = 0; q < 6; ++q)
    for (int j = 0; j < P; ++j) {
      double s = 0;
      for (int i = 0; i < 12; ++i) s += S[i][j] * b[q][i];
      out[q * kMaxPatch + j] = f[q] * s;
    }
  return true;
}

namespace {

/* One Loop refinement of an irregular patch: the N+12 fine points as
 * combinations of the N+6 patch points (slots v0, ring 1..N with ring 1 =
 * v1 and ring N = v2, then Q11 Q20 Q2m Q02 Qm2 at lattice (1,1) (2,0)
 * (2,-1) (0,2) (-1,2)).  Fine points: 0 the vertex point of v0 (Loop's
 * beta_N), 1..N the edge points of v0 and each ring vertex, then (v1,v2),
 * v1, (v1,ring 2), v2, (v2,ring N-1), (v1,Q11), (v1,Q20), (v1,Q2m),
 * (v2,Q11), (v2,Q02), (v2,Qm2).  The first N+6 fine points are the next
 * level's irregular patch in the same slot order (Stam's subdivision matrix
 * A; all N+12 are his extended matrix).  This is the numbering the
 * selection matrices P_k of SubdivEvaluation.h:31-77 address. */
void loop_extended(int N, double Ab[kMaxPatch + 6][kMaxPatch])
{
  for (int i = 0; i < kMaxPatch + 6; ++i)
    for (int j = 0; j < kMaxPatch; ++j) Ab[i][j] = 0;
  const int V0 = 0, V1 = 1, V2 = N, Q11 = N + 1, Q20 = N + 2, Q2m = N + 3, Q02 = N + 4, Qm2 = N + 5;
  const int Ra = 2, Rz = N - 1;
  auto E = [&](int row, int a, int b, int c, int d) {
    Ab[row][a] += 0.375; Ab[row][b] += 0.375; Ab[row][c] += 0.125; Ab[row][d] += 0.125;
  };
  auto R6 = [&](int row, int ctr, std::initializer_list<int> nb) {
    Ab[row][ctr] += 0.625;
    for (int q : nb) Ab[row][q] += 0.0625;
  };
  const double cc = 0.375 + 0.25 * std::cos(2.0 * M_PI / N);
  const double beta = (0.625 - cc * cc) / N; /* Loop (1987) */
  Ab[0][V0] += 1.0 - N * beta;
  for (int k = 1; k <= N; ++k) Ab[0][k] += beta;
  for (int j = 1; j <= N; ++j) {
    const int prev = j == 1 ? N : j - 1, next = j == N ? 1 : j + 1;
    E(j, V0, j, prev, next);
  }
  E(N + 1, V1, V2, V0, Q11);
  R6(N + 2, V1, {V0, Q20, Q11, V2, Q2m, Ra});
  E(N + 3, V1, Ra, V0, Q2m);
  R6(N + 4, V2, {V0, V1, Q11, Q02, Qm2, Rz});
  E(N + 5, V2, Rz, V0, Qm2);
  E(N + 6, V1, Q11, V2, Q20);
  E(N + 7, V1, Q20, Q11, Q2m);
  E(N + 8, V1, Q2m, Ra, Q20);
  E(N + 9, V2, Q11, V1, Q02);
  E(N + 10, V2, Q02, Q11, Qm2);
  E(N + 11, V2, Qm2, Q02, Rz);
}

} // namespace

const double* limit_table()
{
  /* built once: the box splines' monomial coefficients and, per valence
   * 3..5, sub-patch k and level n, M = P_k Ab A^(n-1) with rows of A the
   * first N+6 rows of Ab, multiplied out left to right */
  static const std::vector<double> tab = [] {
    std::vector<double> t(kLimitTabSize, 0.0);
    for (int k = 0; k < 12; ++k) {
      long long M[5][5];
      box_monomials(k, M);
      int m = 0;
      for (int a = 0; a < 5; ++a)
        for (int b = 0; a + b < 5; ++b, ++m) t[k * kBoxMono + m] = (double)M[a][b];
    }
    for (int N = 3; N <= 5; ++N) {
      const int P = N + 6;
      /* P_k: box spline i takes fine point sel[k][i] (SubdivEvaluation.h:31-77) */
      const int sel[3][12] = {{2, 0, N + 3, 1, N, N + 8, N + 2, N + 1, N + 4, N + 7, N + 6, N + 9},
                              {N + 9, N + 6, N + 4, N + 1, N + 2, N + 5, N, 1, N + 3, N - 1, 0, 2},
                              {0, N - 1, 1, N, N + 5, N + 2, N + 1, N + 4, N + 11, N + 6, N + 9, N + 10}};
      double Ab[kMaxPatch + 6][kMaxPatch];
      loop_extended(N, Ab);
      for (int k = 0; k < 3; ++k) {
        double M[12][kMaxPatch], T[12][kMaxPatch];
        for (int i = 0; i < 12; ++i)
          for (int j = 0; j < P; ++j) M[i][j] = Ab[sel[k][i]][j];
        for (int n = 1; n <= kLimitLevels; ++n) {
          if (n > 1) {
            for (int i = 0; i < 12; ++i)
              for (int j = 0; j < P; ++j) {
                double s = 0;
                for (int q = 0; q < P; ++q) s += M[i][q] * Ab[q][j];
                T[i][j] = s;
              }
            for (int i = 0; i < 12; ++i)
              for (int j = 0; j < P; ++j) M[i][j] = T[i][j];
          }
          double* o = t.data() + limit_tab_off(N, k, n);
          for (int i = 0; i < 12; ++i)
            for (int j = 0; j < P; ++j) o[i * P + j] = M[i][j];
        }
      }
    }
    return t;
  }();
  return tab.data();
}

/* SubdivEvaluation.C:37-80, 112-165, 170-187 through the tables */
bool subdiv_weights(int valence, double xi, double eta, double w[3][kMaxPatch])
{
  if (valence < 3 || valence > 6) return false;
  limit_weights(limit_table(), valence, xi, eta, w);
  return true;
}

} // namespace shx
