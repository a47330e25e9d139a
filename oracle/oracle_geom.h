/*
 * oracle_geom.h -- vector type, narrow-phase distance routines, contact law
 * and friction of the CPU FP64 oracle.  TEST INFRASTRUCTURE ONLY (see
 * shell_oracle.h).
 *
 * The reference declares Contact::segment_segment_distance and
 * Contact::point_triangle_distance (Contact.h:22-34) but ships no Contact.C,
 * so their arithmetic is not available.  They are restated here from the
 * published closed forms in C. Ericson, "Real-Time Collision Detection"
 * (2005): ClosestPtSegmentSegment (sec. 5.1.9) and ClosestPtPointTriangle
 * (sec. 5.1.5).  Output conventions are the ones the callers in
 * ShellShellContact.C:87-145 rely on:
 *   segment_segment_distance(p10,p11,p20,p21, diff, st): diff = c2 - c1 with
 *     c1 = p10 + st[0]*(p11-p10), c2 = p20 + st[1]*(p21-p20);
 *   point_triangle_distance(p0, p1,p2,p3, diff, vw): diff = q - p0 with
 *     q = p1 + vw[0]*(p2-p1) + vw[1]*(p3-p1).
 * Compiled with -ffp-contract=off: every product and sum is rounded once, in
 * the order written.
 */
#ifndef ORACLE_GEOM_H
#define ORACLE_GEOM_H

#include <cmath>
#include <limits>

namespace orc {

struct Pt {
  double x[3];
  Pt() { x[0] = x[1] = x[2] = 0; }
  Pt(double a, double b, double c) { x[0] = a; x[1] = b; x[2] = c; }
  double& operator()(int i) { return x[i]; }
  double operator()(int i) const { return x[i]; }
  Pt operator+(const Pt& o) const { return Pt(x[0] + o.x[0], x[1] + o.x[1], x[2] + o.x[2]); }
  Pt operator-(const Pt& o) const { return Pt(x[0] - o.x[0], x[1] - o.x[1], x[2] - o.x[2]); }
  Pt operator-() const { return Pt(-x[0], -x[1], -x[2]); }
  Pt operator*(double s) const { return Pt(x[0] * s, x[1] * s, x[2] * s); }
  Pt operator/(double s) const { return Pt(x[0] / s, x[1] / s, x[2] / s); }
  /* libMesh TypeVector semantics: vector*vector is the dot product */
  double operator*(const Pt& o) const { return x[0] * o.x[0] + x[1] * o.x[1] + x[2] * o.x[2]; }
  Pt& operator+=(const Pt& o) { x[0] += o.x[0]; x[1] += o.x[1]; x[2] += o.x[2]; return *this; }
  Pt& operator-=(const Pt& o) { x[0] -= o.x[0]; x[1] -= o.x[1]; x[2] -= o.x[2]; return *this; }
  Pt& operator*=(double s) { x[0] *= s; x[1] *= s; x[2] *= s; return *this; }
  void add_scaled(const Pt& o, double s) { x[0] += o.x[0] * s; x[1] += o.x[1] * s; x[2] += o.x[2] * s; }
  Pt cross(const Pt& o) const {
    return Pt(x[1] * o.x[2] - x[2] * o.x[1], x[2] * o.x[0] - x[0] * o.x[2], x[0] * o.x[1] - x[1] * o.x[0]);
  }
  double size_sq() const { return x[0] * x[0] + x[1] * x[1] + x[2] * x[2]; }
  double size() const { return std::sqrt(size_sq()); }
  void zero() { x[0] = x[1] = x[2] = 0; }
};
inline Pt operator*(double s, const Pt& p) { return p * s; }

inline double clamp01(double a) { return a < 0 ? 0 : (a > 1 ? 1 : a); }

/* Ericson 5.1.9, ClosestPtSegmentSegment; EPS guards degenerate segments. */
inline void segment_segment_distance(const Pt& p1, const Pt& q1, const Pt& p2, const Pt& q2,
                                     Pt& diff, double* st)
{
  const double EPS = 1e-300;
  const Pt d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
  const double a = d1 * d1, e = d2 * d2, f = d2 * r;
  double s, t;
  if (a <= EPS && e <= EPS) { s = 0; t = 0; }
  else if (a <= EPS) { s = 0; t = clamp01(f / e); }
  else {
    const double c = d1 * r;
    if (e <= EPS) { t = 0; s = clamp01(-c / a); }
    else {
      const double b = d1 * d2;
      const double denom = a * e - b * b;
      s = (denom != 0) ? clamp01((b * f - c * e) / denom) : 0;
      t = (b * s + f) / e;
      if (t < 0) { t = 0; s = clamp01(-c / a); }
      else if (t > 1) { t = 1; s = clamp01((b - c) / a); }
    }
  }
  const Pt c1 = p1 + d1 * s, c2 = p2 + d2 * t;
  diff = c2 - c1;
  st[0] = s; st[1] = t;
}

/* Ericson 5.1.5, ClosestPtPointTriangle (Voronoi-region form) */
inline void point_triangle_distance(const Pt& p, const Pt& a, const Pt& b, const Pt& c,
                                    Pt& diff, double* vw)
{
  const Pt ab = b - a, ac = c - a, ap = p - a;
  const double d1 = ab * ap, d2 = ac * ap;
  double v, w;
  Pt q;
  if (d1 <= 0 && d2 <= 0) { v = 0; w = 0; q = a; }
  else {
    const Pt bp = p - b;
    const double d3 = ab * bp, d4 = ac * bp;
    if (d3 >= 0 && d4 <= d3) { v = 1; w = 0; q = b; }
    else {
      const double vc = d1 * d4 - d3 * d2;
      if (vc <= 0 && d1 >= 0 && d3 <= 0) { v = d1 / (d1 - d3); w = 0; q = a + ab * v; }
      else {
        const Pt cp = p - c;
        const double d5 = ab * cp, d6 = ac * cp;
        if (d6 >= 0 && d5 <= d6) { v = 0; w = 1; q = c; }
        else {
          const double vb = d5 * d2 - d1 * d6;
          if (vb <= 0 && d2 >= 0 && d6 <= 0) { v = 0; w = d2 / (d2 - d6); q = a + ac * w; }
          else {
            const double va = d3 * d6 - d5 * d4;
            if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
              const double ww = (d4 - d3) / ((d4 - d3) + (d5 - d6));
              v = 1 - ww; w = ww; q = b + (c - b) * ww;
            } else {
              const double denom = 1 / (va + vb + vc);
              v = vb * denom; w = vc * denom;
              q = a + ab * v + ac * w;
            }
          }
        }
      }
    }
  }
  diff = q - p;
  vw[0] = v; vw[1] = w;
}

/* Triangle-triangle distance as ShellShellContact::compute (ShellShellContact.C:35-145):
 * 9 edge-edge tests, then 6 vertex-face tests, strict '<' keeps the first minimum.
 * P[e][n] are the three vertices of element e. */
struct TriContact {
  double d2;
  Pt dc;
  double v[2], w[2];
};

inline void tri_tri_contact(const Pt P[2][3], TriContact& c)
{
  static const int mod3[3] = {1, 2, 0};
  c.d2 = std::numeric_limits<double>::max();
  double st[2];
  Pt this_dc;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      segment_segment_distance(P[0][i], P[0][mod3[i]], P[1][j], P[1][mod3[j]], this_dc, st);
      const double this_d2 = this_dc.size_sq();
      if (this_d2 < c.d2) {
        c.d2 = this_d2;
        c.dc = this_dc;
        c.v[0] = (i == 0) * st[0] + (i == 1) * (1 - st[0]);
        c.w[0] = (i == 1) * st[0] + (i == 2) * (1 - st[0]);
        c.v[1] = (j == 0) * st[1] + (j == 1) * (1 - st[1]);
        c.w[1] = (j == 1) * st[1] + (j == 2) * (1 - st[1]);
      }
    }
  double vw[2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 3; ++j) {
      point_triangle_distance(P[i][j], P[1 - i][0], P[1 - i][1], P[1 - i][2], this_dc, vw);
      const double this_d2 = this_dc.size_sq();
      if (this_d2 < c.d2) {
        c.d2 = this_d2;
        c.dc = (i == 0 ? this_dc : -this_dc);
        c.v[0] = (i == 0) * (j == 1) + (i == 1) * vw[0];
        c.w[0] = (i == 0) * (j == 2) + (i == 1) * vw[1];
        c.v[1] = (i == 1) * (j == 1) + (i == 0) * vw[0];
        c.w[1] = (i == 1) * (j == 2) + (i == 0) * vw[1];
      }
    }
}

/* ShellSystem.h:387-401 */
inline double shell_shell_contact_force(double stiffness, double thickness, double distance)
{
  return stiffness * (thickness - distance) * 0.5 * (1 + thickness / distance);
}
inline double shell_shell_contact_energy(double stiffness, double thickness, double distance)
{
  const double depth = thickness - distance;
  return 0.5 * stiffness * (0.5 * depth * depth - depth * thickness + thickness * thickness * std::log(thickness / distance));
}

/* Friction::apply, Friction.C:4-60 (Margolis stick-slip) */
inline void friction_apply(double& Efric, Pt& f_contact, double f_normal, const Pt& n, const Pt& v,
                           double abs_v_normal, double v_stop, double spring_stiffness, double mu_s,
                           double mu_d, double damping, bool& stick, Pt& spring_origin,
                           const Pt& contact_point)
{
  const Pt v_normal = abs_v_normal * n;
  const Pt v_tangential = v - v_normal;
  const double abs_v_tangential = v_tangential.size();
  const double v_eps = 1e-8 * v_stop;
  if (mu_s > 0 && abs_v_tangential <= v_stop) {
    if (!stick) {
      stick = true;
      const double prefactor = mu_d * f_normal / (spring_stiffness * (abs_v_tangential + v_eps)) - damping;
      spring_origin = contact_point - prefactor * v_tangential;
    }
    Pt u_spring = contact_point - spring_origin;
    u_spring -= (u_spring * n) * n;
    Efric = 0.5 * spring_stiffness * u_spring.size_sq();
    f_contact = -spring_stiffness * u_spring;
    f_contact -= damping * spring_stiffness * v_tangential;
  }
  const double mf = mu_s * f_normal;
  if (abs_v_tangential > v_stop || f_contact.size_sq() >= mf * mf) {
    spring_origin.zero();
    stick = false;
    Efric = 0;
    f_contact = -mu_d * f_normal / (abs_v_tangential + v_eps) * v_tangential;
  }
}

} // namespace orc
#endif
