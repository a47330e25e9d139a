/*
 * shx_elem.hip -- the element-force kernels of libshellhip (FP64 and FP32
 * split): patch gather, metrics, strains, stresses and the per-block node
 * records the node kernel sums.
 *
 * This translation unit is compiled with floating-point contraction
 * (fused multiply-add): the element force is checked against the CPU oracle
 * to a relative 1e-10, not bit for bit.  The predictor (pred3) rounds each
 * operation as the other kernels do.
 */
#pragma clang fp contract(fast)
#include <hip/hip_runtime.h>

#include "shx_device.h"
#include "shx_kernels.h"
#include "shx_math.h"
#include "shx_dev.h"

namespace shx {

/* The block's (element, slot) entries (local element * 16 + slot, the
 * element's valence class - 3 in bits 12-13; at most kMaxBlockEnt) and per
 * block node its first entry (block-relative) and slot-group counts, staged
 * in LDS for the record passes.  Block nodes come in descending entry count:
 * thread t sums records t and 511 - t (rec_of). */
__device__ __forceinline__ void stage_entries(const DPtrs& P, int b0, int nbn, uint16_t* s_ent, uint16_t* s_nst,
                                              uint32_t* s_ngc)
{
  const int k0 = P.rent_start[b0], nent = P.rent_start[b0 + nbn] - k0;
  for (int k = threadIdx.x; k < nent; k += blockDim.x) s_ent[k] = P.rent[k0 + k];
  for (int j = threadIdx.x; j < nbn; j += blockDim.x) {
    s_nst[j] = (uint16_t)(P.rent_start[b0 + j] - k0);
    s_ngc[j] = P.rent_gcnt[b0 + j];
  }
}

__device__ __forceinline__ int rec_of(int t, int m) { return m == 0 ? t : 2 * kElemBlock - 1 - t; }

/* The element force.  Per element: gather u_pred over the patch, form the
 * deformed metric (compute_deformed_metrics, ShellSystem.C:1407-1460),
 * membrane and bending strains, orthotropic H, stress resultants and the
 * p-vectors (subtract_internal_force, ShellSystem.C:1164-1233), and the
 * element's generalised stresses G; the force on patch slot i is
 *   f_i = -(phi1 G1 + phi2 G2 - phi11 G3 - phi22 G4 - phi12 G5) + p phi G0,
 * the reference's f -= JxW (Bm^T n + Bb^T m) regrouped by basis derivative
 * (ShellSystem.C:1236-1282) plus the LOCAL pressure load (ShellSystem.C:
 * 1570-1576), G0 = JxW a3.
 *
 * A block covers kElemPerBlock elements, one per thread.  It first stages
 * its patch nodes (bn_ids: u_pred and the position the enclosed volume
 * uses) and its (element, slot) entry lists in LDS with coalesced loads, so
 * the patch gather and the record sums read LDS only; max |u_pred - u| is
 * taken over the staged nodes (every node of the shell is a patch node of
 * some block).  After the stress phase, in four passes of three slots, the
 * element threads put q_i in LDS and one thread per block node sums its
 * entries (fixed order) into the node's block record: sum q_i and
 * sum phi_i G0 (G0 = JxW a3 of the element, in LDS once).  The node kernel
 * adds a node's records (one to three), so no contribution is written per
 * (element, slot) and no atomics are used.  Block partials (energies, volume terms, max
 * |du|^2, max |u_pred - u_build|^2) go to epart for k_fin_elem.  The blocks of a shell whose contact
 * list was rebuilt at the end of the previous step clear its cell buckets.
 * Launch block b computes element block bmap[V_ELEM][b] (the blocks of the
 * active shells, set by the host). */
template <bool GROW>
__device__ __forceinline__ void elem_body(const DPtrs& P, int mode)
{
  __shared__ double s_rb[4 * 72];
  __shared__ double s_w[kElemRec][kElemBlock / 64];
  __shared__ double s_buf[kElemLds];     /* staged nodes, then per element three q_i and JxW a3 */
  __shared__ uint16_t s_ent[kMaxBlockEnt], s_nst[kMaxBlockNodes];
  __shared__ uint32_t s_ngc[kMaxBlockNodes];
  const int blk = uni(P.bmap[GROW ? kMapElemGrow : V_ELEM][blockIdx.x]);
  const int sid = uni(P.eshell[blk * kElemPerBlock]);
  const SConst& C = P.cs[sid];
  const DShell& sh = P.sh[sid];
  if (sh.bclear && C.contact) clear_buckets(P, C, blk - C.blkE, C.nbE);
  if (sh.done) return;
  load_rawb(P.rawb, s_rb);
  {
  const int nE = P.nE;
  const int t = threadIdx.x;
  const int e = blk * kElemPerBlock + t;
  const bool valid = e - C.ebase < C.ne;
  const bool predict = (mode == M_STEP);
  const double dt = sh.dt, a8 = sh.nm[8];
  const int par = sh.par;
  const double* U = P.U[par];
  const double* V = P.V[par];
  const double* Aa = P.A[par];
  const double* T = P.tab;
  const int b0 = P.bn_start[blk], nbn = P.bn_start[blk + 1] - b0;
  const bool track = predict && C.contact && C.rebuild_trigger;
  double dev = 0;                        /* max |u_pred - u_build|^2 of the block's nodes (trigger 1) */
  double dumax = 0;                      /* max |u_pred - u|^2 of the block's nodes */
  stage_entries(P, b0, nbn, s_ent, s_nst, s_ngc);
  for (int j = t; j < nbn; j += kElemBlock) {
    const int n = P.bn_ids[b0 + j];
    const v3 u = ld3(U, n);
    const v3 up = predict ? pred3(u, ld3(V, n), ld3(Aa, n), dt, a8) : u;
    const v3 x = add(ld3(P.X, n), C.volume_at_prediction ? up : u);
    if (n - C.vbase < C.nvr) {           /* a ghost node's increment is not checked (ShellSystem.C:2609) */
      if (predict) dumax = fmax(dumax, nrm2_rn(sub(up, u)));
      if (track) dev = fmax(dev, nrm2_rn(sub(up, ld3(P.urb, n))));
    }
    double* sn = s_buf + kNodeStage * j;
    sn[0] = up.x; sn[1] = up.y; sn[2] = up.z;
    sn[3] = x.x; sn[4] = x.y; sn[5] = x.z;
  }
  __syncthreads();
  double Et = 0, Eb = 0, det = 0, oarea = 0, kgs = 0, kmn = 0, narea = 0;
  v3 du1 = mk(0, 0, 0), du2 = du1, du11 = du1, du22 = du1, du12 = du1;
  double R00 = 0, R01 = 0, R10 = 0, R11 = 0;
  int N = 6;
  if (valid) {
    R00 = T[(size_t)T_R00 * nE + e]; R01 = T[(size_t)T_R01 * nE + e];
    R10 = T[(size_t)T_R10 * nE + e]; R11 = T[(size_t)T_R11 * nE + e];
    N = P.vcls[e];
    const double* rb = s_rb + 72 * (N - 3);
    v3 xa = du1, xb = du1, xc = du1;
    /* the rotation to the element frame is linear: the patch sums are
     * formed with the reference-coordinate derivatives and rotated once,
     * phi_{i,1} = R00 phi_{i,xi} + R10 phi_{i,eta}, etc. (basis_at) */
    v3 sx = du1, se = du1, sxx = du1, see = du1, sxe = du1;
#pragma unroll 2
    for (int i = 0; i < 12; ++i) {
      const double* sn = s_buf + kNodeStage * P.lidx[(size_t)i * nE + e];
      const v3 up = mk(sn[0], sn[1], sn[2]);
      axpy(sx, up, rb[12 + i]);
      axpy(se, up, rb[24 + i]);
      axpy(sxx, up, rb[36 + i]);
      axpy(see, up, rb[48 + i]);
      axpy(sxe, up, rb[60 + i]);
      if (i == 0 || i == 1 || i == N) {
        const v3 x = mk(sn[3], sn[4], sn[5]);
        if (i == 0) xa = x;
        else if (i == 1) xb = x;
        else xc = x;
      }
    }
    du1 = add(scl(sx, R00), scl(se, R10));
    du2 = add(scl(sx, R01), scl(se, R11));
    du11 = add(add(scl(sxx, R00 * R00), scl(see, R10 * R10)), scl(sxe, 2 * R00 * R10));
    du22 = add(add(scl(sxx, R01 * R01), scl(see, R11 * R11)), scl(sxe, 2 * R01 * R11));
    du12 = add(add(scl(sxx, R00 * R01), scl(see, R10 * R11)), scl(sxe, R01 * R10 + R00 * R11));
    /* triad(var, n) = p_n(var), TypeTensor::det (ShellSystem.C:2523-2531) */
    det = xa.x * (xb.y * xc.z - xc.y * xb.z) - xb.x * (xa.y * xc.z - xc.y * xa.z) +
          xc.x * (xa.y * xb.z - xb.y * xa.z);
  }
  v3 ka3 = mk(0, 0, 0), kG1 = ka3, kG2 = ka3;
  double kJ = 0, ks3 = 0, ks4 = 0, ks5 = 0;
  if (valid) {
    const v3 A1 = ldrow3(T, T_A1, nE, e), A2 = ldrow3(T, T_A2, nE, e);
    const v3 A11 = ldrow3(T, T_A11, nE, e), A22 = ldrow3(T, T_A22, nE, e), A12 = ldrow3(T, T_A12, nE, e);
    /* the reference metrics ff1_ref, ff2_ref (compute_reference_metrics,
     * ShellSystem.C:1317-1379) are formed here from the reference tangents
     * rather than read from the table: 48 bytes less per element */
    v3 A3 = crs(A1, A2);
    A3 = scl(A3, 1 / nrm(A3));
    const v3 g = mk(dot(A1, A1), dot(A2, A2), dot(A1, A2));
    const v3 fr2 = mk(dot(A11, A3), dot(A22, A3), dot(A12, A3));
    const v3 a1 = add(A1, du1);
    const v3 a2 = add(A2, du2);
    v3 a3 = crs(a1, a2);
    const double jac = nrm(a3);
    a3 = scl(a3, 1 / jac);
    const v3 a11 = add(A11, du11);
    const v3 a22 = add(A22, du22);
    const v3 a12 = add(A12, du12);
    const double ff1[3] = {dot(a1, a1), dot(a2, a2), dot(a1, a2)};
    const double ff2[3] = {dot(a11, a3), dot(a22, a3), dot(a12, a3)};
    /* gf1, gf2: the forms the strains and the curvature measurement see; with
     * in-plane growth the deformed forms through the inverse growth tensor,
     * GV ff (ShellSystem.C:1176-1186, 2747-2756); the p-vectors keep ff2 */
    double gf1[3] = {ff1[0], ff1[1], ff1[2]}, gf2[3] = {ff2[0], ff2[1], ff2[2]};
    double GV[9];
    const bool grows = GROW && C.grow;
    if (grows) {
      for (int q = 0; q < 9; ++q) GV[q] = P.gGV[(size_t)q * nE + e];
      for (int r = 0; r < 3; ++r) {
        gf1[r] = GV[3 * r] * ff1[0] + GV[3 * r + 1] * ff1[1] + GV[3 * r + 2] * ff1[2];
        gf2[r] = GV[3 * r] * ff2[0] + GV[3 * r + 1] * ff2[1] + GV[3 * r + 2] * ff2[2];
      }
    }
    double alpha[3] = {0.5 * (gf1[0] - g.x), 0.5 * (gf1[1] - g.y), 0.5 * (gf1[2] - g.z)};
    double beta[3] = {fr2.x - gf2[0], fr2.y - gf2[1], fr2.z - gf2[2]};
    alpha[2] *= 2;
    beta[2] *= 2;
    if (grows && C.grow == 2 && !sh.measure_only)
      for (int k = 0; k < 3; ++k) {
        P.gstr[(size_t)k * nE + e] = alpha[k];
        P.gstr[(size_t)(3 + k) * nE + e] = beta[k];
      }
    /* orthotropic H, ShellSystem.C:1199-1213 */
    double H[9];
    H[0] = C.E1 * g.y * g.y;
    H[1] = H[3] = C.E1 * (C.nu * g.y * g.x + (1 - C.nu) * g.z * g.z);
    H[2] = H[6] = C.E1 * -g.y * g.z;
    H[4] = C.E2 * g.x * g.x;
    H[5] = H[7] = C.E2 * -g.x * g.z;
    H[8] = C.G12 * ((1 - C.nu * C.nu2) * g.y * g.x + (1 - C.nu * C.nu2) * ((1 + C.nu) / (1 - C.nu)) * g.z * g.z);
    const double invdet = 1 / (g.x * g.y - g.z * g.z);
    for (int q = 0; q < 9; ++q) H[q] *= invdet * invdet;
    double nn[3], mm[3];
    for (int r = 0; r < 3; ++r) {
      nn[r] = C.K * (H[r * 3] * alpha[0] + H[r * 3 + 1] * alpha[1] + H[r * 3 + 2] * alpha[2]);
      mm[r] = C.D * (H[r * 3] * beta[0] + H[r * 3 + 1] * beta[1] + H[r * 3 + 2] * beta[2]);
    }
    const double JxW = T[(size_t)T_JXW * nE + e];
    Et = 0.5 * JxW * (nn[0] * alpha[0] + nn[1] * alpha[1] + nn[2] * alpha[2]);
    Eb = 0.5 * JxW * (mm[0] * beta[0] + mm[1] * beta[1] + mm[2] * beta[2]);
    if (grows) {
      /* Bm^T GVT n + Bb^T GVT m with GVT the Voigt-corrected transpose of GV
       * (ShellSystem.C:1182-1186, 1270-1275): the resultants carried by GVT */
      const double GT[9] = {GV[0], GV[3], 2 * GV[6], GV[1], GV[4], 2 * GV[7], 0.5 * GV[2], 0.5 * GV[5], GV[8]};
      double n2[3], m2[3];
      for (int r = 0; r < 3; ++r) {
        n2[r] = GT[3 * r] * nn[0] + GT[3 * r + 1] * nn[1] + GT[3 * r + 2] * nn[2];
        m2[r] = GT[3 * r] * mm[0] + GT[3 * r + 1] * mm[1] + GT[3 * r + 2] * mm[2];
      }
      for (int r = 0; r < 3; ++r) { nn[r] = n2[r]; mm[r] = m2[r]; }
    }
    const v3 a2xa3 = crs(a2, a3), a3xa1 = crs(a3, a1);
    const double invjac = 1 / jac;
    const v3 p11 = scl(add(crs(a11, a2), scl(a2xa3, ff2[0])), invjac);
    const v3 p12 = scl(add(crs(a22, a2), scl(a2xa3, ff2[1])), invjac);
    const v3 p13 = scl(add(crs(a12, a2), scl(a2xa3, ff2[2])), invjac);
    const v3 p21 = scl(add(crs(a1, a11), scl(a3xa1, ff2[0])), invjac);
    const v3 p22 = scl(add(crs(a1, a22), scl(a3xa1, ff2[1])), invjac);
    const v3 p23 = scl(add(crs(a1, a12), scl(a3xa1, ff2[2])), invjac);
    const double m2x2 = 2 * mm[2];
    v3 g1 = add(add(add(add(scl(a1, nn[0]), scl(a2, nn[2])), scl(p11, mm[0])), scl(p12, mm[1])), scl(p13, m2x2));
    v3 g2 = add(add(add(add(scl(a2, nn[1]), scl(a1, nn[2])), scl(p21, mm[0])), scl(p22, mm[1])), scl(p23, m2x2));
    const v3 G[6] = {scl(a3, JxW), scl(g1, JxW), scl(g2, JxW), scl(a3, JxW * mm[0]), scl(a3, JxW * mm[1]),
                     scl(a3, JxW * m2x2)};
    /* G0 and G3..G5 are a3 scaled: kept as a3 and the four scales */
    ka3 = a3; kG1 = G[1]; kG2 = G[2];
    kJ = JxW; ks3 = JxW * mm[0]; ks4 = JxW * mm[1]; ks5 = JxW * m2x2;
    const double new_area = 0.5 * jac;
    oarea = (mode == M_INIT || C.volume_at_prediction) ? new_area : P.area[e];
    P.area[e] = new_area;
    narea = new_area;
    /* curvatures of the deformed surface, measure_properties (ShellSystem.C:
     * 2746-2751): weight jac * qw (qw = 1/2), over det(ff1) */
    const double wd = new_area / (gf1[0] * gf1[1] - gf1[2] * gf1[2]);
    kgs = wd * (gf2[0] * gf2[1] - gf2[2] * gf2[2]);
    kmn = wd * 0.5 * (gf2[0] * gf1[1] + gf2[1] * gf1[0] - 2 * gf2[2] * gf1[2]);
    if (mode == M_FORCES) {
      /* the per-element measurement vectors (ShellSystem.C:1299-1300, 2783-2784) */
      double* em = P.emeas + 4 * (size_t)e;
      em[0] = Eb / new_area; em[1] = Et / new_area; em[2] = kgs / new_area; em[3] = kmn / new_area;
    }
  }
  /* block records in four passes of three patch slots: each element thread
   * writes q_i (f_i -= q_i) of its slots 3g .. 3g+2 to LDS, then each
   * block-node thread adds its record's entries of those slots (in entry
   * order) to sum q_i and sum phi_i G0 (f_i += p phi_i G0).  With the
   * rotation folded into per-element vectors,
   *   q_i = phi_{i,xi} Vx + phi_{i,eta} Ve
   *         - a3 (phi_{i,xixi} cxx + phi_{i,etaeta} cee + phi_{i,xieta} cxe),
   * the same sum as G1 phi_{i,1} + G2 phi_{i,2} - G3 phi_{i,11} - ... */
  const double* rb = s_rb + 72 * (N - 3);
  double* s_g0 = s_buf + kQArea;
  v3 kVx = ka3, kVe = ka3;
  double cxx = 0, cee = 0, cxe = 0;
  __syncthreads();   /* the staged nodes are consumed; s_buf now takes q_i and JxW a3 */
  if (valid) {
    const v3 kG0 = scl(ka3, kJ);
    s_g0[3 * t] = kG0.x; s_g0[3 * t + 1] = kG0.y; s_g0[3 * t + 2] = kG0.z;
    kVx = add(scl(kG1, R00), scl(kG2, R01));
    kVe = add(scl(kG1, R10), scl(kG2, R11));
    cxx = ks3 * (R00 * R00) + ks4 * (R01 * R01) + ks5 * (R00 * R01);
    cee = ks3 * (R10 * R10) + ks4 * (R11 * R11) + ks5 * (R10 * R11);
    cxe = ks3 * (2 * R00 * R10) + ks4 * (2 * R01 * R11) + ks5 * (R01 * R10 + R00 * R11);
  }
  v3 rq[2] = {mk(0, 0, 0), mk(0, 0, 0)}, rp[2] = {mk(0, 0, 0), mk(0, 0, 0)};
#pragma unroll 1
  for (int g = 0; g < 4; ++g) {
    if (valid) {
      double* qs = s_buf + kQRec * t;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int i = 3 * g + k;
        const double c = rb[36 + i] * cxx + rb[48 + i] * cee + rb[60 + i] * cxe;
        const v3 q = sub(add(scl(kVx, rb[12 + i]), scl(kVe, rb[24 + i])), scl(ka3, c));
        qs[3 * k] = q.x; qs[3 * k + 1] = q.y; qs[3 * k + 2] = q.z;
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int j = rec_of(t, m);
      if (j < nbn) {
        const uint32_t cnt = s_ngc[j];
        int k0 = s_nst[j];
        for (int gg = 0; gg < g; ++gg) k0 += (cnt >> (8 * gg)) & 255;
        const int k1 = k0 + ((cnt >> (8 * g)) & 255);
        for (int k = k0; k < k1; ++k) {
          const int ent = s_ent[k], le = (ent >> 4) & 255, i = ent & 15;
          const double* qs = s_buf + kQRec * le + 3 * (i - 3 * g);
          const double* g0 = s_g0 + 3 * le;
          const double phi = s_rb[72 * (ent >> 12) + i];
          rq[m] = add(rq[m], mk(qs[0], qs[1], qs[2]));
          rp[m] = add(rp[m], scl(mk(g0[0], g0[1], g0[2]), phi));
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const int j = rec_of(t, m);
    if (j < nbn) {
      double* rec = P.brec + (size_t)kBRec * (b0 + j);
      rec[0] = rq[m].x; rec[1] = rq[m].y; rec[2] = rq[m].z; rec[3] = rp[m].x; rec[4] = rp[m].y; rec[5] = rp[m].z;
    }
  }
  /* block partial: wave reductions (DPP), then the block's waves in order */
  double rr[kElemRec] = {Et, Eb, dumax, det, oarea, dev, kgs, kmn, narea};
#pragma unroll
  for (int c = 0; c < kElemRec; ++c) rr[c] = (c == 2 || c == 5) ? wave_reduce<true>(rr[c]) : wave_reduce<false>(rr[c]);
  if ((t & 63) == 0)
#pragma unroll
    for (int c = 0; c < kElemRec; ++c) s_w[c][t / 64] = rr[c];
  __syncthreads();
  if (t < kElemRec) {
    double acc = s_w[t][0];
    for (int w = 1; w < kElemBlock / 64; ++w)
      acc = (t == 2 || t == 5) ? fmax(acc, s_w[t][w]) : acc + s_w[t][w];
    P.epart[(size_t)blk * kElemRec + t] = acc;
  }
  }
}

/* launch-shape variants (waves per SIMD), picked by the context's shx_shapes */
__global__ void __launch_bounds__(kElemBlock, 3) k_elem_w3(DPtrs P, int mode) { elem_body<false>(P, mode); }
__global__ void __launch_bounds__(kElemBlock, 4) k_elem_w4(DPtrs P, int mode) { elem_body<false>(P, mode); }
/* with in-plane growth (batches where a shell grows) */
__global__ void __launch_bounds__(kElemBlock, 3) k_elem_grow(DPtrs P, int mode) { elem_body<true>(P, mode); }

/* FP32 split element force: the same method as k_elem, with the metric and
 * curvature changes formed in FP32 from the displacement gradients relative
 * to the reference surface, so that no difference of nearly equal numbers
 * is rounded:
 *   a_a.a_b - A_a.A_b = A_a.du_b + du_a.A_b + du_a.du_b,
 *   n = a1 x a2 = N + dn,  dn = A1 x du2 + du1 x A2 + du1 x du2,
 *   |n| - |N| = (2 N.dn + dn.dn) / (|n| + |N|),
 *   a3 - A3 = dn / |n| - N (|n| - |N|) / (|n| |N|),
 *   ff2_ref - ff2 = -(A_ab.(a3 - A3) + du_ab.a3).
 * The displacements enter the FP32 gradient sums relative to the patch's
 * first node (formed in FP64).  The predictor, the element vertex positions
 * of the enclosed volume and all reductions stay FP64.  Slot contributions
 * and JxW a3 enter the block records in FP32 (brec32). */
__global__ void __launch_bounds__(kElemBlock) k_elem32(DPtrs P, int mode)
{
  __shared__ float s_rb[4 * 72];
  __shared__ double s_w[kElemRec][kElemBlock / 64];
  __shared__ double s_buf[kElemLds];     /* staged nodes (FP64), then per element three q_i and JxW a3 (FP32) */
  __shared__ uint16_t s_ent[kMaxBlockEnt], s_nst[kMaxBlockNodes];
  __shared__ uint32_t s_ngc[kMaxBlockNodes];
  const int blk = uni(P.bmap[V_ELEM32][blockIdx.x]);
  const int sid = uni(P.eshell[blk * kElemPerBlock]);
  const SConst& C = P.cs[sid];
  const DShell& sh = P.sh[sid];
  if (sh.bclear && C.contact) clear_buckets(P, C, blk - C.blkE, C.nbE);
  if (sh.done) return;
  for (int q = threadIdx.x; q < 4 * 72; q += blockDim.x) s_rb[q] = (float)P.rawb[q];
  const int nE = P.nE;
  const int t = threadIdx.x;
  const int e = blk * kElemPerBlock + t;
  const bool valid = e - C.ebase < C.ne;
  const bool predict = (mode == M_STEP);
  const double dt = sh.dt, a8 = sh.nm[8];
  const int par = sh.par;
  const double* U = P.U[par];
  const double* V = P.V[par];
  const double* Aa = P.A[par];
  const float* T = P.tab32;
  const int b0 = P.bn_start[blk], nbn = P.bn_start[blk + 1] - b0;
  const bool track = predict && C.contact && C.rebuild_trigger;
  double dev = 0;                        /* max |u_pred - u_build|^2 of the block's nodes (trigger 1) */
  double dumax = 0;                      /* max |u_pred - u|^2 of the block's nodes */
  stage_entries(P, b0, nbn, s_ent, s_nst, s_ngc);
  for (int j = t; j < nbn; j += kElemBlock) {
    const int n = P.bn_ids[b0 + j];
    const v3 u = ld3(U, n);
    const v3 up = predict ? pred3(u, ld3(V, n), ld3(Aa, n), dt, a8) : u;
    const v3 x = add(ld3(P.X, n), C.volume_at_prediction ? up : u);
    if (n - C.vbase < C.nvr) {           /* a ghost node's increment is not checked (ShellSystem.C:2609) */
      if (predict) dumax = fmax(dumax, nrm2_rn(sub(up, u)));
      if (track) dev = fmax(dev, nrm2_rn(sub(up, ld3(P.urb, n))));
    }
    double* sn = s_buf + kNodeStage * j;
    sn[0] = up.x; sn[1] = up.y; sn[2] = up.z;
    sn[3] = x.x; sn[4] = x.y; sn[5] = x.z;
  }
  __syncthreads();
  double Et = 0, Eb = 0, det = 0, oarea = 0, kgs = 0, kmn = 0, narea = 0;
  v3f du1 = mkf(0, 0, 0), du2 = du1, du11 = du1, du22 = du1, du12 = du1;
  float R00 = 0, R01 = 0, R10 = 0, R11 = 0;
  int N = 6;
  if (valid) {
    R00 = T[(size_t)F_R00 * nE + e]; R01 = T[(size_t)F_R01 * nE + e];
    R10 = T[(size_t)F_R10 * nE + e]; R11 = T[(size_t)F_R11 * nE + e];
    N = P.vcls[e];
    const float* rb = s_rb + 72 * (N - 3);
    /* displacements relative to the patch's first node, formed in FP64 and
     * then rounded: the gradient sums are unchanged (the basis derivatives
     * sum to zero over the patch) and the rounding acts on |u_i - u_0| */
    const double* s0 = s_buf + kNodeStage * P.lidx[e];
    const v3 u0 = mk(s0[0], s0[1], s0[2]);
    v3 xa = mk(0, 0, 0), xb = xa, xc = xa;
    /* reference-coordinate sums, rotated once (as k_elem) */
    v3f sx = du1, se = du1, sxx = du1, see = du1, sxe = du1;
#pragma unroll 2
    for (int i = 0; i < 12; ++i) {
      const double* sn = s_buf + kNodeStage * P.lidx[(size_t)i * nE + e];
      const v3 up = mk(sn[0], sn[1], sn[2]);
      const v3 ur = sub(up, u0);
      const v3f uf = mkf((float)ur.x, (float)ur.y, (float)ur.z);
      axpyf(sx, uf, rb[12 + i]);
      axpyf(se, uf, rb[24 + i]);
      axpyf(sxx, uf, rb[36 + i]);
      axpyf(see, uf, rb[48 + i]);
      axpyf(sxe, uf, rb[60 + i]);
      if (i == 0 || i == 1 || i == N) {
        const v3 x = mk(sn[3], sn[4], sn[5]);
        if (i == 0) xa = x;
        else if (i == 1) xb = x;
        else xc = x;
      }
    }
    du1 = addf(sclf(sx, R00), sclf(se, R10));
    du2 = addf(sclf(sx, R01), sclf(se, R11));
    du11 = addf(addf(sclf(sxx, R00 * R00), sclf(see, R10 * R10)), sclf(sxe, 2 * R00 * R10));
    du22 = addf(addf(sclf(sxx, R01 * R01), sclf(see, R11 * R11)), sclf(sxe, 2 * R01 * R11));
    du12 = addf(addf(sclf(sxx, R00 * R01), sclf(see, R10 * R11)), sclf(sxe, R01 * R10 + R00 * R11));
    det = xa.x * (xb.y * xc.z - xc.y * xb.z) - xb.x * (xa.y * xc.z - xc.y * xa.z) +
          xc.x * (xa.y * xb.z - xb.y * xa.z);
  }
  float* s_q = reinterpret_cast<float*>(s_buf);
  v3f ka3 = mkf(0, 0, 0), kG1 = ka3, kG2 = ka3;
  float kJ = 0, ks3 = 0, ks4 = 0, ks5 = 0;
  if (valid) {
    const v3f A1 = ldrow3f(T, F_A1, nE, e), A2 = ldrow3f(T, F_A2, nE, e);
    const v3f A11 = ldrow3f(T, F_A11, nE, e), A22 = ldrow3f(T, F_A22, nE, e), A12 = ldrow3f(T, F_A12, nE, e);
    const v3f A3 = ldrow3f(T, F_A3, nE, e);
    const float J = T[(size_t)F_J * nE + e];
    const v3f Nn = sclf(A3, J);
    const v3f a1 = addf(A1, du1), a2 = addf(A2, du2);
    /* membrane strain from the differences, alpha_3 doubled */
    float alpha[3] = {dotf(A1, du1) + 0.5f * dotf(du1, du1), dotf(A2, du2) + 0.5f * dotf(du2, du2),
                      dotf(A1, du2) + dotf(du1, A2) + dotf(du1, du2)};
    const v3f dn = addf(addf(crsf(A1, du2), crsf(du1, A2)), crsf(du1, du2));
    const v3f nvec = addf(Nn, dn);
    const float jac = sqrtf(dotf(nvec, nvec));
    const float djac = (2 * dotf(Nn, dn) + dotf(dn, dn)) / (jac + J);
    const v3f d3 = subf(sclf(dn, 1 / jac), sclf(Nn, djac / (jac * J)));
    const v3f a3 = addf(A3, d3);
    const v3f a11 = addf(A11, du11), a22 = addf(A22, du22), a12 = addf(A12, du12);
    /* bending strain ff2_ref - ff2, beta_3 doubled */
    float beta[3] = {-(dotf(A11, d3) + dotf(du11, a3)), -(dotf(A22, d3) + dotf(du22, a3)),
                     -2 * (dotf(A12, d3) + dotf(du12, a3))};
    const float ff2[3] = {dotf(a11, a3), dotf(a22, a3), dotf(a12, a3)};
    const float* Hs = T + (size_t)F_H * nE + e;
    const float H00 = Hs[0], H01 = Hs[(size_t)nE], H02 = Hs[(size_t)2 * nE], H11 = Hs[(size_t)3 * nE],
                H12 = Hs[(size_t)4 * nE], H22 = Hs[(size_t)5 * nE];
    const float Hm[9] = {H00, H01, H02, H01, H11, H12, H02, H12, H22};
    const float K = (float)C.K, D = (float)C.D;
    float nn[3], mm[3];
    for (int r = 0; r < 3; ++r) {
      nn[r] = K * (Hm[r * 3] * alpha[0] + Hm[r * 3 + 1] * alpha[1] + Hm[r * 3 + 2] * alpha[2]);
      mm[r] = D * (Hm[r * 3] * beta[0] + Hm[r * 3 + 1] * beta[1] + Hm[r * 3 + 2] * beta[2]);
    }
    const v3f a2xa3 = crsf(a2, a3), a3xa1 = crsf(a3, a1);
    const float invjac = 1 / jac;
    const v3f p11 = sclf(addf(crsf(a11, a2), sclf(a2xa3, ff2[0])), invjac);
    const v3f p12 = sclf(addf(crsf(a22, a2), sclf(a2xa3, ff2[1])), invjac);
    const v3f p13 = sclf(addf(crsf(a12, a2), sclf(a2xa3, ff2[2])), invjac);
    const v3f p21 = sclf(addf(crsf(a1, a11), sclf(a3xa1, ff2[0])), invjac);
    const v3f p22 = sclf(addf(crsf(a1, a22), sclf(a3xa1, ff2[1])), invjac);
    const v3f p23 = sclf(addf(crsf(a1, a12), sclf(a3xa1, ff2[2])), invjac);
    const float JxW = T[(size_t)F_JXW * nE + e];
    const float m2x2 = 2 * mm[2];
    const v3f g1 = addf(addf(addf(addf(sclf(a1, nn[0]), sclf(a2, nn[2])), sclf(p11, mm[0])), sclf(p12, mm[1])),
                        sclf(p13, m2x2));
    const v3f g2 = addf(addf(addf(addf(sclf(a2, nn[1]), sclf(a1, nn[2])), sclf(p21, mm[0])), sclf(p22, mm[1])),
                        sclf(p23, m2x2));
    const v3f G[6] = {sclf(a3, JxW), sclf(g1, JxW), sclf(g2, JxW), sclf(a3, JxW * mm[0]), sclf(a3, JxW * mm[1]),
                      sclf(a3, JxW * m2x2)};
    ka3 = a3; kG1 = G[1]; kG2 = G[2];
    kJ = JxW; ks3 = JxW * mm[0]; ks4 = JxW * mm[1]; ks5 = JxW * m2x2;
    Et = 0.5 * (double)JxW * ((double)nn[0] * alpha[0] + (double)nn[1] * alpha[1] + (double)nn[2] * alpha[2]);
    Eb = 0.5 * (double)JxW * ((double)mm[0] * beta[0] + (double)mm[1] * beta[1] + (double)mm[2] * beta[2]);
    const double new_area = 0.5 * (double)jac;
    oarea = (mode == M_INIT || C.volume_at_prediction) ? new_area : P.area[e];
    P.area[e] = new_area;
    narea = new_area;
    /* curvatures (measure_properties, ShellSystem.C:2746-2751) from the
     * deformed first fundamental form ff1 = A.A + the metric changes */
    const float f0 = dotf(A1, A1) + 2 * alpha[0], f1 = dotf(A2, A2) + 2 * alpha[1], f2 = dotf(A1, A2) + alpha[2];
    const double wd = new_area / ((double)f0 * f1 - (double)f2 * f2);
    kgs = wd * ((double)ff2[0] * ff2[1] - (double)ff2[2] * ff2[2]);
    kmn = wd * 0.5 * ((double)ff2[0] * f1 + (double)ff2[1] * f0 - 2.0 * ff2[2] * f2);
    if (mode == M_FORCES) {
      double* em = P.emeas + 4 * (size_t)e;
      em[0] = Eb / new_area; em[1] = Et / new_area; em[2] = kgs / new_area; em[3] = kmn / new_area;
    }
  }
  /* block records in four passes of three patch slots (as k_elem) */
  const float* rb = s_rb + 72 * (N - 3);
  float* s_g0 = s_q + kQArea;
  v3f kVx = ka3, kVe = ka3;
  float cxx = 0, cee = 0, cxe = 0;
  __syncthreads();   /* the staged nodes are consumed; s_buf now takes q_i and JxW a3 */
  if (valid) {
    const v3f kG0 = sclf(ka3, kJ);
    s_g0[3 * t] = kG0.x; s_g0[3 * t + 1] = kG0.y; s_g0[3 * t + 2] = kG0.z;
    kVx = addf(sclf(kG1, R00), sclf(kG2, R01));
    kVe = addf(sclf(kG1, R10), sclf(kG2, R11));
    cxx = ks3 * (R00 * R00) + ks4 * (R01 * R01) + ks5 * (R00 * R01);
    cee = ks3 * (R10 * R10) + ks4 * (R11 * R11) + ks5 * (R10 * R11);
    cxe = ks3 * (2 * R00 * R10) + ks4 * (2 * R01 * R11) + ks5 * (R01 * R10 + R00 * R11);
  }
  v3f rq[2] = {mkf(0, 0, 0), mkf(0, 0, 0)}, rp[2] = {mkf(0, 0, 0), mkf(0, 0, 0)};
#pragma unroll 1
  for (int g = 0; g < 4; ++g) {
    if (valid) {
      float* qs = s_q + kQRec * t;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int i = 3 * g + k;
        const float c = rb[36 + i] * cxx + rb[48 + i] * cee + rb[60 + i] * cxe;
        const v3f q = subf(addf(sclf(kVx, rb[12 + i]), sclf(kVe, rb[24 + i])), sclf(ka3, c));
        qs[3 * k] = q.x; qs[3 * k + 1] = q.y; qs[3 * k + 2] = q.z;
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int j = rec_of(t, m);
      if (j < nbn) {
        const uint32_t cnt = s_ngc[j];
        int k0 = s_nst[j];
        for (int gg = 0; gg < g; ++gg) k0 += (cnt >> (8 * gg)) & 255;
        const int k1 = k0 + ((cnt >> (8 * g)) & 255);
        for (int k = k0; k < k1; ++k) {
          const int ent = s_ent[k], le = (ent >> 4) & 255, i = ent & 15;
          const float* qs = s_q + kQRec * le + 3 * (i - 3 * g);
          const float* g0 = s_g0 + 3 * le;
          const float phi = s_rb[72 * (ent >> 12) + i];
          rq[m] = addf(rq[m], mkf(qs[0], qs[1], qs[2]));
          rp[m] = addf(rp[m], sclf(mkf(g0[0], g0[1], g0[2]), phi));
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const int j = rec_of(t, m);
    if (j < nbn) {
      float* rec = P.brec32 + (size_t)kBRec * (b0 + j);
      rec[0] = rq[m].x; rec[1] = rq[m].y; rec[2] = rq[m].z; rec[3] = rp[m].x; rec[4] = rp[m].y; rec[5] = rp[m].z;
    }
  }
  double rr[kElemRec] = {Et, Eb, dumax, det, oarea, dev, kgs, kmn, narea};
#pragma unroll
  for (int c = 0; c < kElemRec; ++c) rr[c] = (c == 2 || c == 5) ? wave_reduce<true>(rr[c]) : wave_reduce<false>(rr[c]);
  if ((t & 63) == 0)
#pragma unroll
    for (int c = 0; c < kElemRec; ++c) s_w[c][t / 64] = rr[c];
  __syncthreads();
  if (t < kElemRec) {
    double acc = s_w[t][0];
    for (int w = 1; w < kElemBlock / 64; ++w)
      acc = (t == 2 || t == 5) ? fmax(acc, s_w[t][w]) : acc + s_w[t][w];
    P.epart[(size_t)blk * kElemRec + t] = acc;
  }
}

} // namespace shx
