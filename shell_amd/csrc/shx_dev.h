/*
 * shx_dev.h -- device helpers shared by the kernel translation units of
 * libshellhip (shx_kernels.hip, shx_elem.hip).
 */
#ifndef SHX_DEV_H
#define SHX_DEV_H

#include <hip/hip_runtime.h>

#include "shx_device.h"
#include "shx_math.h"

namespace shx {

/* blocks of one shell for a kernel family: V_TILE 256-element tiles,
 * V_LANE / V_LANE8 / V_LANE32 pair-search blocks at 16 / 8 / 32 lanes per
 * element, V_ELEM element-kernel blocks, V_NODE node-kernel blocks */
enum { V_TILE = 0, V_LANE = 1, V_ELEM = 2, V_NODE = 3, V_ELEM32 = 4, V_LANE8 = 5, V_LANE32 = 6 };

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

/* x of another lane by a DPP move (both 32-bit halves); lanes of rows
 * outside RMASK get 0 */
template <int CTRL, int RMASK>
__device__ __forceinline__ double dpp_f64(double x)
{
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, RMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, RMASK, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

/* sum (MAX false) or max (MAX true, of values >= 0) over the 64 lanes of a
 * wave with every lane active, in a fixed order, without LDS: lane pairs and
 * quads (quad_perm), half rows and rows (row_half_mirror, row_mirror), then
 * rows 1 and 3 add the last lane of rows 0 and 2 (row_bcast:15) and rows 2
 * and 3 that of row 1 (row_bcast:31); lane 63 holds the result, which every
 * lane returns */
template <bool MAX>
__device__ __forceinline__ double wave_reduce(double x)
{
  auto op = [](double a, double b) { return MAX ? fmax(a, b) : a + b; };
  x = op(x, dpp_f64<0xB1, 0xF>(x));    /* quad_perm [1,0,3,2] */
  x = op(x, dpp_f64<0x4E, 0xF>(x));    /* quad_perm [2,3,0,1] */
  x = op(x, dpp_f64<0x141, 0xF>(x));   /* row_half_mirror */
  x = op(x, dpp_f64<0x140, 0xF>(x));   /* row_mirror */
  x = op(x, dpp_f64<0x142, 0xA>(x));   /* row_bcast:15 */
  x = op(x, dpp_f64<0x143, 0xC>(x));   /* row_bcast:31 */
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ v3 ld3(const double* p, int n) { return mk(p[3 * n], p[3 * n + 1], p[3 * n + 2]); }
__device__ __forceinline__ void st3(double* p, int n, v3 a) { p[3 * n] = a.x; p[3 * n + 1] = a.y; p[3 * n + 2] = a.z; }
__device__ __forceinline__ v3 ldrow3(const double* t, int row, int nE, int e)
{
  return mk(t[(size_t)row * nE + e], t[(size_t)(row + 1) * nE + e], t[(size_t)(row + 2) * nE + e]);
}

/* u_pred = u; u_pred.add(dt, v); u_pred.add(a8, a)  (ShellSystem.C:973-975),
 * each product and sum rounded once in every translation unit (the element
 * kernels are compiled with contraction, the others without) */
__device__ __forceinline__ double pred1(double u, double v, double a, double dt, double a8)
{
  double x = u;
  x = __dadd_rn(__dmul_rn(dt, v), x);
  x = __dadd_rn(__dmul_rn(a8, a), x);
  return x;
}
__device__ __forceinline__ v3 pred3(v3 u, v3 v, v3 a, double dt, double a8)
{
  return mk(pred1(u.x, v.x, a.x, dt, a8), pred1(u.y, v.y, a.y, dt, a8), pred1(u.z, v.z, a.z, dt, a8));
}

/* |a|^2 with each product and sum rounded once (as the CPU oracle forms
 * it), also in a translation unit compiled with contraction */
__device__ __forceinline__ double nrm2_rn(v3 a)
{
  return __dadd_rn(__dadd_rn(__dmul_rn(a.x, a.x), __dmul_rn(a.y, a.y)), __dmul_rn(a.z, a.z));
}

/* reparametrised basis of slot i of an element (reinit_mod,
 * ShellSystem.C:3466-3478) from the raw basis rb (6 x 12) and R */
struct Basis6 { double phi, d1, d2, d11, d22, d12; };
__device__ __forceinline__ Basis6 basis_at(const double* rb, int i, double R00, double R01, double R10,
                                           double R11)
{
  Basis6 b;
  const double dx = rb[12 + i], de = rb[24 + i];
  const double dxx = rb[36 + i], dee = rb[48 + i], dxe = rb[60 + i];
  b.phi = rb[i];
  b.d1 = dx * R00 + de * R10;
  b.d2 = dx * R01 + de * R11;
  b.d11 = dxx * R00 * R00 + dee * R10 * R10 + 2 * dxe * R00 * R10;
  b.d22 = dxx * R01 * R01 + dee * R11 * R11 + 2 * dxe * R01 * R11;
  b.d12 = dxx * R00 * R01 + dee * R10 * R11 + dxe * R01 * R10 + dxe * R00 * R11;
  return b;
}

__device__ __forceinline__ void load_rawb(const double* g, double* s)
{
  for (int q = threadIdx.x; q < 4 * 72; q += blockDim.x) s[q] = g[q];
}

/* clear the bucket counts of one shell's cell grid: the nb blocks lb =
 * 0..nb-1 of a kernel's range for the shell each take a strided share */
__device__ __forceinline__ void clear_buckets(const DPtrs& P, const SConst& C, int lb, int nb)
{
  const unsigned nbk = 1u << C.hb_bits;
  for (unsigned b = (unsigned)lb * blockDim.x + threadIdx.x; b < nbk; b += (unsigned)nb * blockDim.x)
    P.bcnt[C.bbase + b] = 0;
}

/* Exact, order-independent sums of contact forces.  A contribution x is
 * rounded once to the nearest multiple of 2^-S and added to a 128-bit two's
 * complement accumulator (w[0] the low word, w[1] the high word): an atomic
 * add on the low word returns its old value, from which the carry out of
 * that 64-bit add follows, and the high word receives x's high word plus
 * the carry (skipped when that is 0).  Integer addition is associative, so
 * the total is the exact sum of the rounded contributions whatever the
 * order of the adds.  Returns false (nothing added) for a non-finite x or
 * |x| >= 2^(126-S). */
__device__ __forceinline__ bool fx_add(unsigned long long* w, double x, int S)
{
  const double y = ldexp(x, S);
  unsigned long long lo;
  long long hi;
  if (fabs(y) < 9.2e18) {                     /* below 2^63: one 64-bit integer */
    const long long q = __double2ll_rn(y);
    lo = (unsigned long long)q;
    hi = q < 0 ? -1 : 0;
  } else if (fabs(y) < 8.5e37) {              /* below 2^126: y is a multiple of 2^11, split exactly */
    const double hd = floor(y * 0x1p-64);
    lo = (unsigned long long)(y - hd * 0x1p64);
    hi = (long long)hd;
  } else {
    return false;                             /* also NaN */
  }
  if (lo != 0) {
    const unsigned long long old = atomicAdd(w, lo);
    hi += (old + lo < lo) ? 1 : 0;            /* carry out of the low word */
  }
  if (hi != 0) atomicAdd(w + 1, (unsigned long long)hi);
  return true;
}

/* the accumulator's value in FP64: magnitude (no cancellation), then sign */
__device__ __forceinline__ double fx_get(unsigned long long lo, unsigned long long hi, int S)
{
  const bool neg = (long long)hi < 0;
  if (neg) {                                  /* two's complement negation of the 128-bit value */
    lo = ~lo + 1ull;
    hi = ~hi + (lo == 0 ? 1ull : 0ull);
  }
  double m = (double)lo;
  if (hi) m = ldexp((double)hi, 64) + m;
  m = ldexp(m, -S);
  return neg ? -m : m;
}

/* FP32 small-vector helpers (device only) */
struct v3f { float x, y, z; };
__device__ __forceinline__ v3f mkf(float x, float y, float z) { v3f r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ v3f addf(v3f a, v3f b) { return mkf(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3f subf(v3f a, v3f b) { return mkf(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3f sclf(v3f a, float s) { return mkf(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dotf(v3f a, v3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3f crsf(v3f a, v3f b)
{
  return mkf(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ void axpyf(v3f& a, v3f b, float s) { a.x += b.x * s; a.y += b.y * s; a.z += b.z * s; }
__device__ __forceinline__ v3f ldrow3f(const float* t, int row, int nE, int e)
{
  return mkf(t[(size_t)row * nE + e], t[(size_t)(row + 1) * nE + e], t[(size_t)(row + 2) * nE + e]);
}

} // namespace shx
#endif
