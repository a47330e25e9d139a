/*
 * shx_math.h -- small-vector math, narrow-phase distance, contact law and
 * friction shared by the gfx950 kernels and the host setup code of
 * libshellhip.  Compiled with -ffp-contract=off on both sides so that the
 * host and the device round every operation identically: the broad-phase
 * pair decisions (d2 <= d2_max) are then the same wherever they are made.
 *
 * Distance routines: closed forms of C. Ericson, "Real-Time Collision
 * Detection" (2005), sec. 5.1.9 (segment-segment) and 5.1.5 (point-triangle),
 * with the output conventions the reference's callers expect
 * (ShellShellContact.C:87-145; declarations Contact.h:22-34, whose
 * definitions the reference does not ship).
 */
#ifndef SHX_MATH_H
#define SHX_MATH_H

#include <math.h>

#if defined(__HIPCC__)
#define SHX_HD __host__ __device__ inline
#else
#define SHX_HD inline
#endif

namespace shx {

struct v3 {
  double x, y, z;
};

SHX_HD v3 mk(double x, double y, double z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
SHX_HD v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
SHX_HD v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
SHX_HD v3 scl(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
SHX_HD v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
SHX_HD double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SHX_HD v3 crs(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
SHX_HD double nrm2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
SHX_HD double nrm(v3 a) { return sqrt(nrm2(a)); }
/* a += b*s with one rounding per product and per sum */
SHX_HD void axpy(v3& a, v3 b, double s) { a.x += b.x * s; a.y += b.y * s; a.z += b.z * s; }

SHX_HD double clamp01(double a) { return a < 0 ? 0 : (a > 1 ? 1 : a); }

SHX_HD void seg_seg(v3 p1, v3 q1, v3 p2, v3 q2, v3& diff, double& s, double& t)
{
  const double EPS = 1e-300;
  const v3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
  const double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  if (a <= EPS && e <= EPS) { s = 0; t = 0; }
  else if (a <= EPS) { s = 0; t = clamp01(f / e); }
  else {
    const double c = dot(d1, r);
    if (e <= EPS) { t = 0; s = clamp01(-c / a); }
    else {
      const double b = dot(d1, d2);
      const double denom = a * e - b * b;
      s = (denom != 0) ? clamp01((b * f - c * e) / denom) : 0;
      t = (b * s + f) / e;
      if (t < 0) { t = 0; s = clamp01(-c / a); }
      else if (t > 1) { t = 1; s = clamp01((b - c) / a); }
    }
  }
  const v3 c1 = add(p1, scl(d1, s)), c2 = add(p2, scl(d2, t));
  diff = sub(c2, c1);
}

SHX_HD void pt_tri(v3 p, v3 a, v3 b, v3 c, v3& diff, double& v, double& w)
{
  const v3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  v3 q;
  if (d1 <= 0 && d2 <= 0) { v = 0; w = 0; q = a; }
  else {
    const v3 bp = sub(p, b);
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { v = 1; w = 0; q = b; }
    else {
      const double vc = d1 * d4 - d3 * d2;
      if (vc <= 0 && d1 >= 0 && d3 <= 0) { v = d1 / (d1 - d3); w = 0; q = add(a, scl(ab, v)); }
      else {
        const v3 cp = sub(p, c);
        const double d5 = dot(ab, cp), d6 = dot(ac, cp);
        if (d6 >= 0 && d5 <= d6) { v = 0; w = 1; q = c; }
        else {
          const double vb = d5 * d2 - d1 * d6;
          if (vb <= 0 && d2 >= 0 && d6 <= 0) { v = 0; w = d2 / (d2 - d6); q = add(a, scl(ac, w)); }
          else {
            const double va = d3 * d6 - d5 * d4;
            if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
              const double ww = (d4 - d3) / ((d4 - d3) + (d5 - d6));
              v = 1 - ww; w = ww; q = add(b, scl(sub(c, b), ww));
            } else {
              const double denom = 1 / (va + vb + vc);
              v = vb * denom; w = vc * denom;
              q = add(add(a, scl(ab, v)), scl(ac, w));
            }
          }
        }
      }
    }
  }
  diff = sub(q, p);
}

/* minimum distance between two triangles: 9 edge-edge then 6 vertex-face
 * tests, keeping the first strict minimum (ShellShellContact.C:87-145).
 * dc points from triangle 0 to triangle 1; (v,w) are barycentric on each. */
struct tri_hit {
  double d2;
  v3 dc;
  double v0, w0, v1, w1;
};

SHX_HD void tri_tri(const v3 P0[3], const v3 P1[3], tri_hit& h)
{
  h.d2 = 1.7976931348623157e308;
  h.dc = mk(0, 0, 0);
  h.v0 = h.w0 = h.v1 = h.w1 = 0;
  for (int i = 0; i < 3; ++i) {
    const int i1 = i == 2 ? 0 : i + 1;
    for (int j = 0; j < 3; ++j) {
      const int j1 = j == 2 ? 0 : j + 1;
      v3 d;
      double s, t;
      seg_seg(P0[i], P0[i1], P1[j], P1[j1], d, s, t);
      const double d2 = nrm2(d);
      if (d2 < h.d2) {
        h.d2 = d2;
        h.dc = d;
        h.v0 = (i == 0) * s + (i == 1) * (1 - s);
        h.w0 = (i == 1) * s + (i == 2) * (1 - s);
        h.v1 = (j == 0) * t + (j == 1) * (1 - t);
        h.w1 = (j == 1) * t + (j == 2) * (1 - t);
      }
    }
  }
  for (int side = 0; side < 2; ++side) {
    const v3* A = side == 0 ? P0 : P1;
    const v3* B = side == 0 ? P1 : P0;
    for (int j = 0; j < 3; ++j) {
      v3 d;
      double v, w;
      pt_tri(A[j], B[0], B[1], B[2], d, v, w);
      const double d2 = nrm2(d);
      if (d2 < h.d2) {
        h.d2 = d2;
        h.dc = side == 0 ? d : neg(d);
        h.v0 = (side == 0) * (j == 1) + (side == 1) * v;
        h.w0 = (side == 0) * (j == 2) + (side == 1) * w;
        h.v1 = (side == 1) * (j == 1) + (side == 0) * v;
        h.w1 = (side == 1) * (j == 2) + (side == 0) * w;
      }
    }
  }
}

/* log-barrier penalty, ShellSystem.h:387-401 */
SHX_HD double contact_force(double k, double h, double d) { return k * (h - d) * 0.5 * (1 + h / d); }
SHX_HD double contact_energy(double k, double h, double d)
{
  const double depth = h - d;
  return 0.5 * k * (0.5 * depth * depth - depth * h + h * h * log(h / d));
}

/* Margolis stick-slip friction, Friction.C:4-60 */
SHX_HD void friction(double& Efric, v3& fc, double fn, v3 n, v3 v, double vn, double v_stop, double ks,
                     double mu_s, double mu_d, double damping, int& stick, v3& origin, v3 cp)
{
  const v3 vt = sub(v, scl(n, vn));
  const double avt = nrm(vt);
  const double v_eps = 1e-8 * v_stop;
  if (mu_s > 0 && avt <= v_stop) {
    if (!stick) {
      stick = 1;
      const double pre = mu_d * fn / (ks * (avt + v_eps)) - damping;
      origin = sub(cp, scl(vt, pre));
    }
    v3 us = sub(cp, origin);
    us = sub(us, scl(n, dot(us, n)));
    Efric = 0.5 * ks * nrm2(us);
    fc = scl(us, -ks);
    fc = sub(fc, scl(vt, damping * ks));
  }
  const double mf = mu_s * fn;
  if (avt > v_stop || nrm2(fc) >= mf * mf) {
    origin = mk(0, 0, 0);
    stick = 0;
    Efric = 0;
    fc = scl(vt, -mu_d * fn / (avt + v_eps));
  }
}

} // namespace shx
#endif
