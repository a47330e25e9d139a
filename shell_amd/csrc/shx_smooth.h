/*
 * shx_smooth.h -- the Loop limit surface at any point of an element and the
 * closest points of two limit-surface elements (smooth shell-shell contact,
 * SmoothShellShellContact.C:21-204), shared by the gfx950 contact kernels
 * and the host entry points; compiled with -ffp-contract=off on both sides.
 *
 * Limit surface (SubdivEvaluation.C:14-187): a regular patch (valence 6) is
 * the 12 quartic box splines of the patch; an element at an irregular
 * vertex of valence N subdivides n times toward it and evaluates the box
 * splines of the regular sub-patch k that holds the point, i.e. the 12 x
 * (N+6) matrix P_k Ab A^(n-1) (Stam 1998) applied to the patch.  Those
 * matrices are tabulated once for N = 3, 4, 5, k = 0..2, n = 1..34 (the
 * reference's max_A_exponent, reached at its irregular_tol = 1e-10).
 *
 * Closest points: minimise f(x) = |p2(v2, w2) - p1(v1, w1)|^2 over x = (v1,
 * w1, v2, w2) with each (v, w) in its element's triangle (v >= 0, w >= 0,
 * v + w <= 1), the problem SmoothShellShellContact.C:23-31 states.  The
 * reference's solver (ConstrainedMinimizer.h) is not part of its sources;
 * this is an active-set Newton method of its own: Newton steps on the face
 * of the feasible set the active constraints leave free (exact Hessian from
 * the surfaces' second derivatives; where it is not positive definite, its
 * eigenvalues taken by magnitude and the step's part of small or negative
 * curvature extended while f falls), the
 * step cut where it meets a constraint (which becomes active), an Armijo
 * backtracking line search outside the quadratic regime, a direct test of
 * an irregular vertex the iterate approaches, and, once
 * stationary on a face, the release of the constraint with the most
 * negative multiplier.  The start is the pair's last minimiser (new pairs:
 * the element centroids, SmoothShellShellContact.C:16).
 */
#ifndef SHX_SMOOTH_H
#define SHX_SMOOTH_H

#include "shx_math.h"

namespace shx {

constexpr int kLimitLevels = 34;       /* max_A_exponent, SubdivEvaluation.h:11-13 */
constexpr int kBoxMono = 15;           /* monomials v^a w^b, a + b <= 4 */
/* table: the 12 box splines as 12 x their monomial coefficients (row k,
 * a-major, b ascending), then per valence N = 3, 4, 5, sub-patch k = 0..2 and
 * level n = 1..34 the row-major 12 x (N+6) matrix P_k Ab A^(n-1) */
constexpr int kLimitTabSize = 12 * kBoxMono + 3 * kLimitLevels * 12 * (9 + 10 + 11);
constexpr int kSmoothIters = 40;       /* Newton iterations of one closest-point search */

SHX_HD int limit_tab_off(int N, int k, int n)
{
  const int base = 12 * kBoxMono + (N == 3 ? 0 : N == 4 ? 3 * kLimitLevels * 12 * 9
                                                          : 3 * kLimitLevels * 12 * (9 + 10));
  return base + (k * kLimitLevels + (n - 1)) * 12 * (N + 6);
}

/* value (q = 0), d/dv, d/dw (1, 2), d2/dv2, d2/dw2, d2/dvdw (3..5) of the 12
 * box splines at (v, w); order 0 value only, 1 up to the first derivatives */
SHX_HD void box_splines(const double* tab, double v, double w, int order, double b[6][12])
{
  double pv[5], pw[5];
  pv[0] = 1;
  pw[0] = 1;
  for (int q = 1; q < 5; ++q) { pv[q] = pv[q - 1] * v; pw[q] = pw[q - 1] * w; }
  for (int k = 0; k < 12; ++k) {
    double r[6] = {0, 0, 0, 0, 0, 0};
    int m = 0;
    for (int a = 0; a < 5; ++a)
      for (int c = 0; a + c < 5; ++c, ++m) {
        const double co = tab[k * kBoxMono + m];
        if (co == 0) continue;
        r[0] += co * pv[a] * pw[c];
        if (order >= 1) {
          if (a > 0) r[1] += co * a * pv[a - 1] * pw[c];
          if (c > 0) r[2] += co * c * pv[a] * pw[c - 1];
        }
        if (order >= 2) {
          if (a > 1) r[3] += co * (a * (a - 1)) * pv[a - 2] * pw[c];
          if (c > 1) r[4] += co * (c * (c - 1)) * pv[a] * pw[c - 2];
          if (a > 0 && c > 0) r[5] += co * (a * c) * pv[a - 1] * pw[c - 1];
        }
      }
    for (int q = 0; q < 6; ++q) b[q][k] = r[q] / 12.0;
  }
}

/* the same for one box spline k, from the powers pv, pw of v and w: r[q] */
SHX_HD void box_spline_one(const double* tab, int k, const double* pv, const double* pw, int order, double r[6])
{
  for (int q = 0; q < 6; ++q) r[q] = 0;
  int m = 0;
  for (int a = 0; a < 5; ++a)
    for (int c = 0; a + c < 5; ++c, ++m) {
      const double co = tab[k * kBoxMono + m];
      if (co == 0) continue;
      r[0] += co * pv[a] * pw[c];
      if (order >= 1) {
        if (a > 0) r[1] += co * a * pv[a - 1] * pw[c];
        if (c > 0) r[2] += co * c * pv[a] * pw[c - 1];
      }
      if (order >= 2) {
        if (a > 1) r[3] += co * (a * (a - 1)) * pv[a - 2] * pw[c];
        if (c > 1) r[4] += co * (c * (c - 1)) * pv[a] * pw[c - 2];
        if (a > 0 && c > 0) r[5] += co * (a * c) * pv[a - 1] * pw[c - 1];
      }
    }
  for (int q = 0; q < 6; ++q) r[q] = r[q] / 12.0;
}

/* rescale_irregular_coords and the sub-patch choice of SubdivEvaluation.C:
 * 39-60, 110-135, 170-187: n is floor(1 - log2(xi + eta)), the count of
 * exact doublings that bring xi + eta into (1/2, 1]; (xi, eta) become the
 * sub-patch's coordinates and jfac the factor per derivative order.  The
 * reference's factor pow(-2, n) for k = 1 has the wrong sign for even n
 * (x' = 1 - 2^n xi); -2^n is the derivative of the evaluated surface. */
SHX_HD void limit_locate(double& xi, double& eta, int& k, int& n, double& jfac)
{
  const double tol = 1e-10;      /* irregular_tol, SubdivEvaluation.h:11 */
  if (xi + eta < tol) { xi = xi > tol ? xi : tol; eta = eta > 0 ? eta : 0; }
  n = 1;
  double s = xi + eta, p2 = 1;
  while (s <= 0.5 && n < kLimitLevels) { s *= 2; p2 *= 2; ++n; }
  xi *= p2;
  eta *= p2;
  jfac = 2 * p2;
  if (xi > 0.5) { k = 0; xi = 2 * xi - 1; eta = 2 * eta; }
  else if (eta > 0.5) { k = 2; xi = 2 * xi; eta = 2 * eta - 1; }
  else { k = 1; xi = 1 - 2 * xi; eta = 1 - 2 * eta; jfac = -jfac; }
}

/* regular-patch renumbering of the box splines, SubdivEvaluation.C:10 */
SHX_HD int limit_cvi(int j)
{
  const int c[12] = {3, 6, 2, 0, 1, 4, 7, 10, 9, 5, 11, 8};
  return c[j];
}

/* limit surface of a patch with positions X[0 .. N+5] at (xi, eta):
 * d[0] the point, d[1], d[2] the tangents, d[3..5] the second derivatives
 * (xixi, etaeta, xieta); order 0 fills d[0], 1 d[0..2], 2 all */
struct Limit {
  v3 d[6];
};

SHX_HD void limit_eval(const double* tab, int N, const v3* X, double xi, double eta, int order, Limit& L)
{
  /* one box spline at a time (no 6 x 12 table in registers) */
  const int nq = order == 0 ? 1 : order == 1 ? 3 : 6;
  for (int q = 0; q < 6; ++q) L.d[q] = mk(0, 0, 0);
  double pv[5], pw[5], r[6];
  if (N == 6) {
    pv[0] = 1;
    pw[0] = 1;
    for (int q = 1; q < 5; ++q) { pv[q] = pv[q - 1] * xi; pw[q] = pw[q - 1] * eta; }
    for (int j = 0; j < 12; ++j) {
      box_spline_one(tab, limit_cvi(j), pv, pw, order, r);
      for (int q = 0; q < nq; ++q) axpy(L.d[q], X[j], r[q]);
    }
    return;
  }
  int k, n;
  double jfac;
  limit_locate(xi, eta, k, n, jfac);
  pv[0] = 1;
  pw[0] = 1;
  for (int q = 1; q < 5; ++q) { pv[q] = pv[q - 1] * xi; pw[q] = pw[q - 1] * eta; }
  const double* M = tab + limit_tab_off(N, k, n);
  const int P = N + 6;
  for (int i = 0; i < 12; ++i) {
    v3 Y = mk(0, 0, 0);
    for (int j = 0; j < P; ++j) axpy(Y, X[j], M[i * P + j]);
    box_spline_one(tab, i, pv, pw, order, r);
    for (int q = 0; q < nq; ++q) axpy(L.d[q], Y, r[q]);
  }
  const double j2 = jfac * jfac;
  for (int q = 1; q < nq; ++q) L.d[q] = scl(L.d[q], q < 3 ? jfac : j2);
}

/* weights of the same evaluation over the patch slots: w[q][j], q = value,
 * d/dxi, d/deta */
SHX_HD void limit_weights(const double* tab, int N, double xi, double eta, double w[3][12])
{
  double b[6][12];
  for (int q = 0; q < 3; ++q)
    for (int j = 0; j < 12; ++j) w[q][j] = 0;
  if (N == 6) {
    box_splines(tab, xi, eta, 1, b);
    for (int q = 0; q < 3; ++q)
      for (int j = 0; j < 12; ++j) w[q][j] = b[q][limit_cvi(j)];
    return;
  }
  int k, n;
  double jfac;
  limit_locate(xi, eta, k, n, jfac);
  box_splines(tab, xi, eta, 1, b);
  const double* M = tab + limit_tab_off(N, k, n);
  const int P = N + 6;
  const double f[3] = {1.0, jfac, jfac};
  for (int q = 0; q < 3; ++q)
    for (int j = 0; j < P; ++j) {
      double s = 0;
      for (int i = 0; i < 12; ++i) s += M[i * P + j] * b[q][i];
      w[q][j] = f[q] * s;
    }
}

/* ------------------------------------------------ closest points */

/* the two patches of a pair: valences and positions */
struct SmoothPair {
  const double* tab;
  int N[2];
  v3 X[2][12];
};

/* active constraints of one element, bits: 1 v = 0, 2 w = 0, 4 v + w = 1 */
SHX_HD void smooth_snap(double& v, double& w, int act)
{
  if (act & 1) v = 0;
  if (act & 2) w = 0;
  if (act & 4) {
    if (act & 1) w = 1;
    else if (act & 2) v = 1;
    else v = 1 - w;
  }
}

/* constraints an element's (v, w) meets, within tol */
SHX_HD int smooth_touching(double v, double w, double tol)
{
  int a = 0;
  if (v <= tol) a |= 1;
  if (w <= tol) a |= 2;
  if (v + w >= 1 - tol) a |= 4;
  if (a == 7) a = (v <= w) ? 5 : 6;   /* the three cannot hold at once: keep the nearer vertex */
  return a;
}

/* f, and with order >= 1 the gradient g, with order 2 the Hessian H (4 x 4
 * row-major) at x; p1, p2 the two points */
SHX_HD double smooth_eval(const SmoothPair& S, const double x[4], int order, double g[4], double H[16], v3& p1,
                          v3& p2)
{
  Limit L1, L2;
  limit_eval(S.tab, S.N[0], S.X[0], x[0], x[1], order, L1);
  limit_eval(S.tab, S.N[1], S.X[1], x[2], x[3], order, L2);
  p1 = L1.d[0];
  p2 = L2.d[0];
  const v3 r = sub(p2, p1);
  if (order >= 1) {
    /* r = p2 - p1: dr/dx = (-p1_v, -p1_w, p2_v, p2_w) */
    const v3 J[4] = {neg(L1.d[1]), neg(L1.d[2]), L2.d[1], L2.d[2]};
    for (int a = 0; a < 4; ++a) g[a] = 2 * dot(J[a], r);
    if (order >= 2) {
      for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c) H[4 * a + c] = 2 * dot(J[a], J[c]);
      /* + 2 r . d2r: -p1'' on the first block, +p2'' on the second */
      const double s1[3] = {-dot(r, L1.d[3]), -dot(r, L1.d[4]), -dot(r, L1.d[5])};
      const double s2[3] = {dot(r, L2.d[3]), dot(r, L2.d[4]), dot(r, L2.d[5])};
      H[0] += 2 * s1[0]; H[5] += 2 * s1[1]; H[1] += 2 * s1[2]; H[4] += 2 * s1[2];
      H[10] += 2 * s2[0]; H[15] += 2 * s2[1]; H[11] += 2 * s2[2]; H[14] += 2 * s2[2];
    }
  }
  return nrm2(r);
}

/* Cholesky solve of (A + mu I) y = b for m <= 4; false if not positive
 * definite (a pivot at or below tiny) */
SHX_HD bool smooth_chol(const double A[16], int m, double mu, const double b[4], double y[4])
{
  double Lm[4][4];
  double dmax = 0;
  for (int i = 0; i < m; ++i) dmax = fmax(dmax, fabs(A[5 * i]));
  const double tiny = 1e-13 * dmax;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = A[4 * i + j] + (i == j ? mu : 0.0);
      for (int k = 0; k < j; ++k) s -= Lm[i][k] * Lm[j][k];
      if (i == j) {
        if (!(s > tiny)) return false;
        Lm[i][i] = sqrt(s);
      } else {
        Lm[i][j] = s / Lm[j][j];
      }
    }
  double z[4];
  for (int i = 0; i < m; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= Lm[i][k] * z[k];
    z[i] = s / Lm[i][i];
  }
  for (int i = m - 1; i >= 0; --i) {
    double s = z[i];
    for (int k = i + 1; k < m; ++k) s -= Lm[k][i] * y[k];
    y[i] = s / Lm[i][i];
  }
  return true;
}

/* eigenvalues lam and eigenvectors (columns of V, stride 4) of the
 * symmetric m x m matrix A (stride 4, overwritten), m <= 4: cyclic Jacobi
 * rotations (Golub & Van Loan, sec. 8.5), at most 12 sweeps */
SHX_HD void smooth_eig(double A[16], int m, double lam[4], double V[16])
{
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) V[4 * i + j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 12; ++sweep) {
    double off = 0, dia = 0;
    for (int i = 0; i < m; ++i) {
      dia += A[5 * i] * A[5 * i];
      for (int j = i + 1; j < m; ++j) off += A[4 * i + j] * A[4 * i + j];
    }
    if (!(off > 1e-32 * dia)) break;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = A[4 * p + q];
        if (apq == 0) continue;
        const double th = (A[5 * q] - A[5 * p]) / (2 * apq);
        const double at = fabs(th);
        double t = at > 1e150 ? 0.5 / at : 1.0 / (at + sqrt(th * th + 1));
        if (th < 0) t = -t;
        const double c = 1 / sqrt(t * t + 1), sn = t * c;
        for (int k = 0; k < m; ++k) {   /* columns p, q */
          const double akp = A[4 * k + p], akq = A[4 * k + q];
          A[4 * k + p] = c * akp - sn * akq;
          A[4 * k + q] = sn * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {   /* rows p, q */
          const double apk = A[4 * p + k], aqk = A[4 * q + k];
          A[4 * p + k] = c * apk - sn * aqk;
          A[4 * q + k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {
          const double vkp = V[4 * k + p], vkq = V[4 * k + q];
          V[4 * k + p] = c * vkp - sn * vkq;
          V[4 * k + q] = sn * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < m; ++i) lam[i] = A[5 * i];
}

/* the longest feasible step along d from x over the inactive constraints,
 * and the constraint (element be, bit bbit) that stops it (be = -1: none) */
SHX_HD double smooth_maxstep(const double x[4], const double d[4], const int act[2], int& be, int& bbit)
{
  double amax = 1.7976931348623157e308;
  be = -1;
  bbit = 0;
  for (int e = 0; e < 2; ++e) {
    const double v = x[2 * e], w = x[2 * e + 1], dv = d[2 * e], dw = d[2 * e + 1];
    if (!(act[e] & 1) && dv < 0 && -v / dv < amax) { amax = -v / dv; be = e; bbit = 1; }
    if (!(act[e] & 2) && dw < 0 && -w / dw < amax) { amax = -w / dw; be = e; bbit = 2; }
    if (!(act[e] & 4) && dv + dw > 0 && (1 - v - w) / (dv + dw) < amax) {
      amax = (1 - v - w) / (dv + dw);
      be = e;
      bbit = 4;
    }
  }
  return amax > 0 ? amax : 0;
}

/* the closest points: x (feasible start) becomes the minimiser; returns
 * |p2 - p1|^2 with dc = p2 - p1 at it; *iters the Newton iterations used */
SHX_HD double smooth_closest(const SmoothPair& S, double x[4], v3& dc, int* iters)
{
  const double kStep = 1e-13;    /* a Newton step below this (barycentric units) ends the face search */
  int act[2];
  for (int e = 0; e < 2; ++e) {
    double& v = x[2 * e];
    double& w = x[2 * e + 1];
    v = v > 0 ? v : 0;
    w = w > 0 ? w : 0;
    if (v + w > 1) { if (v >= w) v = 1 - w; else w = 1 - v; }
    act[e] = smooth_touching(v, w, 1e-14);
    smooth_snap(v, w, act[e]);
  }
  int it = 0;
  bool tried[2] = {false, false};
  v3 p1, p2;
  for (; it < kSmoothIters; ++it) {
    /* at an irregular vertex the parametrisation degenerates (the tangents
     * vanish like (2 lambda)^n) and iterates only creep toward it: an
     * element whose point is within xi + eta < 1e-3 of its irregular vertex
     * tries the vertex itself (not again once a multiplier released it) */
    for (int e = 0; e < 2; ++e) {
      if (S.N[e] == 6 || tried[e] || act[e] == 3 || !(x[2 * e] + x[2 * e + 1] < 1e-3)) continue;
      double xs[4] = {x[0], x[1], x[2], x[3]};
      xs[2 * e] = 0;
      xs[2 * e + 1] = 0;
      v3 q1, q2;
      const double fv = smooth_eval(S, xs, 0, nullptr, nullptr, q1, q2);
      const double fc = smooth_eval(S, x, 0, nullptr, nullptr, q1, q2);
      if (fv <= fc) {
        x[2 * e] = 0;
        x[2 * e + 1] = 0;
        act[e] = 3;
      }
    }
    double g[4], H[16];
    const double f = smooth_eval(S, x, 2, g, H, p1, p2);
    /* free directions: per element none, one (0,1) / (1,0) / (1,-1), or both axes */
    double Z[4][4];
    int m = 0;
    for (int e = 0; e < 2; ++e) {
      const int a = act[e];
      if (a == 3 || a == 5 || a == 6) continue;
      double dirs[2][2];
      int nd = 0;
      if (a == 0) { dirs[0][0] = 1; dirs[0][1] = 0; dirs[1][0] = 0; dirs[1][1] = 1; nd = 2; }
      else if (a == 1) { dirs[0][0] = 0; dirs[0][1] = 1; nd = 1; }
      else if (a == 2) { dirs[0][0] = 1; dirs[0][1] = 0; nd = 1; }
      else { dirs[0][0] = 1; dirs[0][1] = -1; nd = 1; }
      for (int q = 0; q < nd; ++q) {
        for (int c = 0; c < 4; ++c) Z[m][c] = 0;
        Z[m][2 * e] = dirs[q][0];
        Z[m][2 * e + 1] = dirs[q][1];
        ++m;
      }
    }
    double d[4] = {0, 0, 0, 0}, dn[4] = {0, 0, 0, 0};   /* step; its part of small or negative curvature */
    bool pure = true;
    double slope = 0;
    if (m > 0) {
      double gz[4], Hz[16], dz[4];
      for (int i = 0; i < m; ++i) {
        double s = 0;
        for (int c = 0; c < 4; ++c) s += Z[i][c] * g[c];
        gz[i] = -s;
        for (int j = 0; j < m; ++j) {
          double t = 0;
          for (int a = 0; a < 4; ++a) {
            if (Z[i][a] == 0) continue;
            for (int c = 0; c < 4; ++c) t += Z[i][a] * H[4 * a + c] * Z[j][c];
          }
          Hz[4 * i + j] = t;
        }
      }
      if (!smooth_chol(Hz, m, 0.0, gz, dz)) {
        /* not positive definite on the face: Newton on |eigenvalues|
         * (floored at 1e-3 of the largest, so that rounding-sized
         * curvatures do not steer the step), so that negative-curvature
         * directions are descended rather than climbed */
        pure = false;
        double lam[4], Vz[16];
        smooth_eig(Hz, m, lam, Vz);
        double lmax = 0;
        for (int i = 0; i < m; ++i) lmax = fmax(lmax, fabs(lam[i]));
        const double floor_l = lmax > 0 ? 1e-3 * lmax : 1.0;
        double dzn[4] = {0, 0, 0, 0};
        for (int i = 0; i < m; ++i) dz[i] = 0;
        for (int i = 0; i < m; ++i) {
          double c = 0;
          for (int j = 0; j < m; ++j) c += Vz[4 * j + i] * gz[j];
          c /= fmax(fabs(lam[i]), floor_l);
          for (int j = 0; j < m; ++j) dz[j] += c * Vz[4 * j + i];
          if (!(lam[i] > floor_l))
            for (int j = 0; j < m; ++j) dzn[j] += c * Vz[4 * j + i];
        }
        for (int i = 0; i < m; ++i)
          for (int c = 0; c < 4; ++c) dn[c] += Z[i][c] * dzn[i];
      }
      for (int i = 0; i < m; ++i)
        for (int c = 0; c < 4; ++c) d[c] += Z[i][c] * dz[i];
      for (int c = 0; c < 4; ++c) slope += g[c] * d[c];
    }
    double dinf = 0;
    for (int c = 0; c < 4; ++c) dinf = fmax(dinf, fabs(d[c]));
    /* a step whose predicted decrease of f (the Newton decrement) is below
     * f's rounding cannot be resolved: the face search is done */
    const bool resolved = !(-slope <= 1e-13 * f);
    if (m > 0 && dinf >= kStep && slope < 0 && resolved) {
      int be, bbit;
      const double amax = smooth_maxstep(x, d, act, be, bbit);
      double alpha = amax < 1 ? amax : 1.0;
      double gam = 0;            /* extension along dn beyond the full step */
      int ge = -1, gbit = 0;
      if (!(pure && alpha * dinf <= 1e-3)) {
        /* Armijo backtracking on f (outside the quadratic regime) */
        int bt = 0;
        double fa = f;
        for (; bt < 60; ++bt) {
          double xt[4];
          for (int c = 0; c < 4; ++c) xt[c] = x[c] + alpha * d[c];
          v3 q1, q2;
          fa = smooth_eval(S, xt, 0, nullptr, nullptr, q1, q2);
          if (fa <= f + 1e-4 * alpha * slope) break;
          alpha *= 0.5;
        }
        if (bt == 60) break;     /* no decrease left to find: the current point stands */
        /* the full step taken, the part of small or negative curvature may
         * have been short: continue along it from x + d, doubling while f
         * keeps falling, up to the first constraint (the rest of the step
         * is Newton's and stays as it is) */
        if (!pure && bt == 0 && alpha == 1.0) {
          double x1[4];
          for (int c = 0; c < 4; ++c) x1[c] = x[c] + d[c];
          const double gmax = smooth_maxstep(x1, dn, act, ge, gbit);
          double gt = 1.0;
          while (gam < gmax) {
            const double g2 = gt < gmax ? gt : gmax;
            double xt[4];
            for (int c = 0; c < 4; ++c) xt[c] = x1[c] + g2 * dn[c];
            v3 q1, q2;
            const double f2 = smooth_eval(S, xt, 0, nullptr, nullptr, q1, q2);
            if (!(f2 < fa)) break;
            gam = g2;
            fa = f2;
            gt = 2 * g2;
          }
          if (!(gam == gmax)) ge = -1;
        }
      }
      for (int c = 0; c < 4; ++c) x[c] += alpha * d[c] + gam * dn[c];
      if (gam > 0) {
        if (ge >= 0) {
          act[ge] |= gbit;
          if (act[ge] == 7) act[ge] = gbit == 1 ? 5 : 6;
        }
      } else if (alpha == amax && be >= 0) {
        act[be] |= bbit;
        if (act[be] == 7) act[be] = bbit == 1 ? 5 : 6;
      }
      for (int e = 0; e < 2; ++e) {
        double& v = x[2 * e];
        double& w = x[2 * e + 1];
        v = v > 0 ? v : 0;
        w = w > 0 ? w : 0;
        if (v + w > 1) { if (v >= w) v = 1 - w; else w = 1 - v; }
        smooth_snap(v, w, act[e]);
      }
      continue;
    }
    /* stationary on the face: take the last small step, then the
     * multipliers of the active constraints (g = sum lambda_i a_i with
     * a = (1,0), (0,1), (-1,-1) for v >= 0, w >= 0, v + w <= 1) */
    if (m > 0 && slope < 0)
      for (int c = 0; c < 4; ++c) x[c] += d[c];
    for (int e = 0; e < 2; ++e) {
      double& v = x[2 * e];
      double& w = x[2 * e + 1];
      v = v > 0 ? v : 0;
      w = w > 0 ? w : 0;
      if (v + w > 1) { if (v >= w) v = 1 - w; else w = 1 - v; }
      smooth_snap(v, w, act[e]);
    }
    double gs = 0;
    for (int c = 0; c < 4; ++c) gs = fmax(gs, fabs(g[c]));
    double lmin = -1e-10 * gs;
    int re = -1, rbit = 0;
    for (int e = 0; e < 2; ++e) {
      const double gv = g[2 * e], gw = g[2 * e + 1];
      double lam[3] = {0, 0, 0};   /* bits 1, 2, 4 */
      switch (act[e]) {
        case 1: lam[0] = gv; break;
        case 2: lam[1] = gw; break;
        case 4: lam[2] = -0.5 * (gv + gw); break;
        case 3: lam[0] = gv; lam[1] = gw; break;
        case 5: lam[2] = -gw; lam[0] = gv - gw; break;
        case 6: lam[2] = -gv; lam[1] = gw - gv; break;
        default: break;
      }
      for (int q = 0; q < 3; ++q)
        if ((act[e] & (1 << q)) && lam[q] < lmin) { lmin = lam[q]; re = e; rbit = 1 << q; }
    }
    if (re < 0) { ++it; break; }
    if (act[re] == 3) tried[re] = true;
    act[re] &= ~rbit;
  }
  if (iters) *iters = it;
  double gd[4], Hd[16];
  const double d2 = smooth_eval(S, x, 0, gd, Hd, p1, p2);
  dc = sub(p2, p1);
  return d2;
}

} // namespace shx
#endif
