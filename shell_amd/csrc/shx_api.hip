/*
 * shx_api.hip -- host side of libshellhip: the extern "C" entry points of
 * include/shellhip.h, device memory ownership and the launch sequence of one
 * explicit step attempt.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "shellhip.h"
#include "shx_device.h"
#include "shx_kernels.h"
#include "shx_math.h"
#include "shx_setup.h"

using namespace shx;

struct shx_ctx {
  int dev = 0;
  hipStream_t st = nullptr;
  Topology T;
  shx_params prm{};
  SConst C{};
  DPtrs P{};
  std::vector<void*> allocs;
  std::vector<double> rawb;
  int nblkE = 0, nblkN = 0, nblkC = 0;
  std::string err;
  bool prof = false;
  std::vector<hipEvent_t> evpool;
  std::vector<int> evclass;
  double tsum[5] = {0, 0, 0, 0, 0};
  int64_t tcnt[5] = {0, 0, 0, 0, 0};
};

namespace {

thread_local std::string g_create_err;

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      ctx->err = std::string(#x) + ": " + hipGetErrorString(e_);                           \
      return SHX_E_HIP;                                                                    \
    }                                                                                      \
  } while (0)

template <class T>
int dalloc(shx_ctx* ctx, T*& p, size_t n)
{
  void* q = nullptr;
  HIPCHK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
  HIPCHK(hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(T)));
  ctx->allocs.push_back(q);
  p = static_cast<T*>(q);
  return 0;
}

template <class T>
int upload(shx_ctx* ctx, T*& p, const std::vector<T>& h)
{
  if (int r = dalloc(ctx, p, h.size())) return r;
  if (!h.empty()) HIPCHK(hipMemcpy((void*)p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

void set_newmark_host(DShell& s, const SConst& C, double dt)
{
  dt = std::max(dt, C.dt_min);
  dt = std::min(dt, C.dt_max);
  s.dt = dt;
  const double b = C.beta, g = C.gamma;
  s.nm[0] = b * dt * dt;
  s.nm[1] = b != 0 ? g / (b * dt) : 0;
  s.nm[2] = b != 0 ? 1 / (b * dt) : 0;
  s.nm[3] = b != 0 ? 1 / (2 * b) - 1 : 0;
  s.nm[4] = b != 0 ? g / b - 1 : 0;
  s.nm[5] = b != 0 ? dt * (g / (2 * b) - 1) : 0;
  s.nm[6] = (1 - g) * dt;
  s.nm[7] = g * dt;
  s.nm[8] = dt * dt * (0.5 - b);
  s.nm[9] = dt * dt * std::fabs(b - 1. / 6);
}

int read_sh(shx_ctx* ctx, DShell& s)
{
  HIPCHK(hipMemcpyAsync(&s, ctx->P.sh, sizeof(DShell), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

int write_sh(shx_ctx* ctx, const DShell& s)
{
  HIPCHK(hipMemcpyAsync(ctx->P.sh, &s, sizeof(DShell), hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

/* profiling: an event pair around each launch of a kernel class */
void prof_begin(shx_ctx* ctx, int cls)
{
  if (!ctx->prof) return;
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  (void)hipEventRecord(a, ctx->st);
  ctx->evpool.push_back(a);
  ctx->evpool.push_back(b);
  ctx->evclass.push_back(cls);
}
void prof_end(shx_ctx* ctx)
{
  if (!ctx->prof) return;
  (void)hipEventRecord(ctx->evpool.back(), ctx->st);
}
void prof_collect(shx_ctx* ctx)
{
  if (ctx->evclass.empty()) return;
  (void)hipStreamSynchronize(ctx->st);
  for (size_t k = 0; k < ctx->evclass.size(); ++k) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->evpool[2 * k], ctx->evpool[2 * k + 1]) == hipSuccess) {
      ctx->tsum[ctx->evclass[k]] += ms;
      ctx->tcnt[ctx->evclass[k]] += 1;
    }
    (void)hipEventDestroy(ctx->evpool[2 * k]);
    (void)hipEventDestroy(ctx->evpool[2 * k + 1]);
  }
  ctx->evpool.clear();
  ctx->evclass.clear();
}

/* the broad phase: four kernels that return at once unless the device-side
 * trigger (or a forced rebuild) set DShell::rebuild */
int launch_broad(shx_ctx* ctx, int posmode)
{
  const DPtrs& P = ctx->P;
  const SConst& C = ctx->C;
  const int nb = (C.ne + 255) / 256;
  const int nbl = (int)(((long long)C.ne * kBpLanes + 255) / 256);
  prof_begin(ctx, 4);
  k_bp_box<<<nb, 256, 0, ctx->st>>>(P, C, posmode);
  k_bp_bin<<<nbl, 256, 0, ctx->st>>>(P, C);
  k_bp_pairs<<<nbl, 256, 0, ctx->st>>>(P, C, posmode);
  k_bp_emit<<<nb, 256, 0, ctx->st>>>(P, C);
  prof_end(ctx);
  HIPCHK(hipGetLastError());
  return 0;
}

/* one step attempt: element force (+ predictor trigger), broad phase,
 * narrow phase, node gather + Newmark (+ finalize), corrector trigger */
int enqueue_attempt(shx_ctx* ctx)
{
  const DPtrs& P = ctx->P;
  const SConst& C = ctx->C;
  hipStream_t st = ctx->st;
  prof_begin(ctx, 0);
  k_elem<<<ctx->nblkE, kElemBlock, 0, st>>>(P, C, M_STEP);
  prof_end(ctx);
  if (C.contact) {
    if (int r = launch_broad(ctx, M_STEP)) return r;
    prof_begin(ctx, 2);
    k_contact<<<ctx->nblkC, 256, 0, st>>>(P, C, M_STEP);
    prof_end(ctx);
  }
  prof_begin(ctx, 1);
  k_node<<<ctx->nblkN, kNodeBlock, 0, st>>>(P, C, M_STEP, 0.0, C.contact ? ctx->nblkC : 0);
  prof_end(ctx);
  if (C.contact)
    if (int r = launch_broad(ctx, M_FORCES)) return r;
  HIPCHK(hipGetLastError());
  return 0;
}

int check_err(shx_ctx* ctx, const DShell& s)
{
  if (s.err) {
    ctx->err = "device list capacity exceeded (bits " + std::to_string(s.err) + ")";
    return SHX_E_CAPACITY;
  }
  return 0;
}

void fill_report(const DShell& s, shx_report* rep)
{
  if (!rep) return;
  rep->t = s.t; rep->dt = s.dt;
  rep->residual = s.residual; rep->fnorm = s.fnorm; rep->fext_norm = s.fext_norm;
  rep->e_tens = s.e_tens; rep->e_bend = s.e_bend; rep->e_self = s.e_self; rep->e_fric = s.e_fric;
  rep->e_kin = s.e_kin; rep->volume = s.volume; rep->area = s.area; rep->du_max = s.du_max;
  rep->n_contacts = s.n_contacts; rep->n_pairs = s.n_pairs; rep->n_rebuilds = s.n_rebuilds;
  rep->problem = s.problem; rep->n_steps = s.n_steps; rep->n_rejected = s.n_rejected;
  rep->converged_step = s.converged_step; rep->reserved = 0;
  rep->residual_deform = s.residual_deform;
}

} // namespace

extern "C" {

const char* shx_last_error(const shx_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int shx_destroy(shx_ctx* ctx)
{
  if (!ctx) return SHX_OK;
  (void)hipSetDevice(ctx->dev);
  prof_collect(ctx);
  for (void* p : ctx->allocs) (void)hipFree(p);
  if (ctx->st) (void)hipStreamDestroy(ctx->st);
  delete ctx;
  return SHX_OK;
}

int shx_create(shx_ctx** out, int nv, int ne, const double* X, const int32_t* tri, const shx_params* prm,
               int device)
{
  if (!out || nv <= 0 || ne <= 0 || !X || !tri || !prm) { g_create_err = "invalid argument"; return SHX_E_ARG; }
  *out = nullptr;
  if (prm->thickness <= 0 || prm->E1 <= 0 || prm->E2 <= 0 || prm->G12 <= 0 || prm->rho <= 0 ||
      prm->nu < 0 || prm->nu >= 1 || prm->dt <= 0 || prm->newmark_beta < 0) {
    g_create_err = "invalid parameters (ShellSystem.C:326-332 asserts)";
    return SHX_E_ARG;
  }
  shx_ctx* ctx = new shx_ctx();
  ctx->dev = device;
  ctx->prm = *prm;
  auto fail = [&](int code) { g_create_err = ctx->err; shx_destroy(ctx); return code; };
  if (hipSetDevice(device) != hipSuccess) { ctx->err = "hipSetDevice failed"; return fail(SHX_E_HIP); }
  if (hipStreamCreate(&ctx->st) != hipSuccess) { ctx->err = "hipStreamCreate failed"; return fail(SHX_E_HIP); }
  std::string terr;
  if (!build_topology(nv, ne, tri, ctx->T, terr)) { ctx->err = terr; return fail(SHX_E_MESH); }
  const Topology& T = ctx->T;

  /* constants: update_derived (ShellSystem.C:421-487) */
  SConst& C = ctx->C;
  C.nv = nv; C.ne = ne;
  C.h = prm->thickness; C.rho = prm->rho;
  C.E1 = prm->E1; C.E2 = prm->E2; C.nu = prm->nu; C.G12 = prm->G12;
  C.nu2 = (prm->E1 * prm->nu) / prm->E2;
  const double nu21 = (prm->E2 * prm->nu) / prm->E1;
  const double one_nu2 = 1 - prm->nu * nu21;
  C.K = prm->memb_stiff_mult * prm->thickness / one_nu2;
  C.D = prm->bend_stiff_mult * std::pow(prm->thickness, 3) / (12 * one_nu2);
  const double Eavg = (prm->E1 + prm->E2) * 0.5;
  C.stiff_density = 0.5 * Eavg / (one_nu2 * prm->thickness);
  C.stiff_static_friction = prm->fric_stiff_mult * Eavg * prm->thickness / one_nu2;
  C.d2_friction = std::pow((1 + prm->static_friction_tol) * prm->thickness, 2);
  C.pressure_final = prm->pressure_final; C.pressure_ramp_time = prm->pressure_ramp_time;
  C.bulk_modulus = prm->bulk_modulus; C.volume_final_ratio = prm->volume_final_ratio;
  C.volume_ramp_time = prm->volume_ramp_time;
  C.contact_stiffness = prm->contact_stiffness; C.mu_s = prm->mu_s; C.mu_d = prm->mu_d;
  C.self_damping = prm->self_damping; C.self_friction_damping = prm->self_friction_damping;
  C.friction_stop_velocity = prm->friction_stop_velocity;
  C.cell_size_mult = prm->cell_size_mult; C.cell_margin = prm->cell_margin;
  C.beta = prm->newmark_beta; C.gamma = prm->newmark_gamma;
  C.dt_min = std::max(prm->dt_min, DBL_EPSILON);
  C.dt_max = prm->dt_max > 0 ? prm->dt_max : DBL_MAX;
  C.rel_loc_err_min = prm->rel_loc_err_min; C.rel_loc_err_max = prm->rel_loc_err_max;
  C.adapt_dt_every = prm->adapt_dt_every;
  C.contact = prm->contact ? 1 : 0;
  C.self_friction = (prm->mu_s != 0 || prm->mu_d != 0) ? 1 : 0;
  C.quirk_fulltime = prm->quirk_fulltime; C.quirk_volume_schedule = prm->quirk_volume_schedule;
  C.conv_tol = 0; C.load_end_time = std::max(prm->pressure_ramp_time, prm->volume_ramp_time);
  C.stop_on_converged = 0;
  C.volume_at_prediction = prm->volume_at_prediction ? 1 : 0;
  {
    unsigned m = (unsigned)std::ceil(std::sqrt((double)ne) / prm->cell_size_mult);
    C.n_cells_max = m * m * m;
  }

  /* Loop basis at the quadrature point, valences 3..6 */
  ctx->rawb.assign(4 * 72, 0.0);
  for (int val = 3; val <= 6; ++val) loop_basis_at_centroid(val, &ctx->rawb[(val - 3) * 72]);

  /* structure-of-arrays copies of the topology */
  std::vector<int32_t> pidx((size_t)12 * ne), triS((size_t)3 * ne);
  std::vector<int8_t> vcls(ne);
  for (int e = 0; e < ne; ++e) {
    for (int i = 0; i < 12; ++i) {
      const int n = T.patch[(size_t)e * 12 + i];
      pidx[(size_t)i * ne + e] = n >= 0 ? n : T.patch[(size_t)e * 12]; /* padded slot: basis is zero */
    }
    for (int k = 0; k < 3; ++k) triS[(size_t)k * ne + e] = T.tri[(size_t)3 * e + k];
    vcls[e] = (int8_t)T.valence[e];
  }
  std::vector<double> Xh(X, X + (size_t)3 * nv);
  DPtrs& P = ctx->P;
  ctx->nblkE = (ne + kElemBlock / kElemLanes - 1) / (kElemBlock / kElemLanes);
  ctx->nblkN = (nv + kNodeBlock / kNodeLanes - 1) / (kNodeBlock / kNodeLanes);
  ctx->nblkC = 120;
  int r = 0;
  double* dX;
  int32_t *dp, *dt3, *dis, *di;
  int8_t* dv;
  double* drb;
  if ((r = upload(ctx, dX, Xh)) || (r = upload(ctx, dp, pidx)) || (r = upload(ctx, dv, vcls)) ||
      (r = upload(ctx, dt3, triS)) || (r = upload(ctx, drb, ctx->rawb)) || (r = upload(ctx, dis, T.inc_start)) ||
      (r = upload(ctx, di, T.inc)))
    return fail(r);
  P.X = dX; P.pidx = dp; P.vcls = dv; P.tri = dt3; P.rawb = drb; P.inc_start = dis; P.inc = di;
  if ((r = dalloc(ctx, P.tab, (size_t)T_ROWS * ne)) || (r = dalloc(ctx, P.minv, nv)) ||
      (r = dalloc(ctx, P.cdamp, nv)) || (r = dalloc(ctx, P.qe, (size_t)36 * ne)) || (r = dalloc(ctx, P.g0, (size_t)3 * ne)) || (r = dalloc(ctx, P.area, ne)) ||
      (r = dalloc(ctx, P.ctr, 8)) ||
      (r = dalloc(ctx, P.fc, (size_t)3 * nv)) || (r = dalloc(ctx, P.fout, (size_t)3 * nv)) ||
      (r = dalloc(ctx, P.fext_out, (size_t)3 * nv)) || (r = dalloc(ctx, P.epart, (size_t)5 * ctx->nblkE)) ||
      (r = dalloc(ctx, P.npart, (size_t)kNodeRec * ctx->nblkN)) || (r = dalloc(ctx, P.cpart, (size_t)3 * ctx->nblkC)) ||
      (r = dalloc(ctx, P.sh, 1)))
    return fail(r);
  for (int b = 0; b < 2; ++b)
    if ((r = dalloc(ctx, P.U[b], (size_t)3 * nv)) || (r = dalloc(ctx, P.V[b], (size_t)3 * nv)) ||
        (r = dalloc(ctx, P.A[b], (size_t)3 * nv)))
      return fail(r);

  if (C.contact) {
    std::vector<int32_t> cs, ci;
    build_close_sets(T, X, (1 + prm->neighbor_margin) * prm->thickness, cs, ci);
    int32_t *dcs, *dci;
    if ((r = upload(ctx, dcs, cs)) || (r = upload(ctx, dci, ci))) return fail(r);
    P.close_start = dcs; P.close_ids = dci;
    /* hashed cell grid: about four entries per bucket for the expected
     * number of (cell, element) entries, (1/cell_size_mult + 2)^3 each */
    const double kc = std::pow(1.0 / prm->cell_size_mult + 2.0, 3.0);
    const double want = std::max(4096.0, (double)ne * kc / 4.0);
    unsigned bits = 12;
    while ((double)(1u << bits) < want && bits < 24) ++bits;
    P.hb_bits = bits;
    P.pair_cap = 32 * ne + 4096;
    const int ntiles = (ne + 255) / 256;
    if ((r = dalloc(ctx, P.ebox, (size_t)6 * ne)) || (r = dalloc(ctx, P.bpart, (size_t)7 * ntiles)) ||
        (r = dalloc(ctx, P.ecell, (size_t)6 * ne)) || (r = dalloc(ctx, P.bcount, (size_t)1 << bits)) ||
        (r = dalloc(ctx, P.bent, ((size_t)1 << bits) * kBucketCap)) ||
        (r = dalloc(ctx, P.found_cnt, (size_t)2 * (ne + 1))) ||
        (r = dalloc(ctx, P.found_e2, (size_t)kRowCap * ne)) || (r = dalloc(ctx, P.found_d2, (size_t)kRowCap * ne)))
      return fail(r);
    for (int b = 0; b < 2; ++b)
      if ((r = dalloc(ctx, P.row_start[b], (size_t)ne + 1)) || (r = dalloc(ctx, P.pe1[b], (size_t)P.pair_cap)) ||
          (r = dalloc(ctx, P.pe2[b], (size_t)P.pair_cap)) || (r = dalloc(ctx, P.pd2[b], (size_t)P.pair_cap)) ||
          (r = dalloc(ctx, P.pstick[b], (size_t)P.pair_cap)) || (r = dalloc(ctx, P.pso[b], (size_t)3 * P.pair_cap)))
        return fail(r);
  }

  k_precompute<<<ctx->nblkE, 256, 0, ctx->st>>>(P, C);
  k_mass<<<ctx->nblkN, 256, 0, ctx->st>>>(P, C, prm->damping_alpha, prm->damping_beta, prm->damping_gamma);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->st) != hipSuccess) {
    ctx->err = "setup kernels failed";
    return fail(SHX_E_HIP);
  }
  DShell s;
  std::memset(&s, 0, sizeof(s));
  set_newmark_host(s, C, prm->dt);
  s.fresh = 1;
  s.converged_step = -1;
  if ((r = write_sh(ctx, s))) return fail(r);
  *out = ctx;
  return SHX_OK;
}

int shx_set_state(shx_ctx* ctx, const double* u, const double* v, const double* a, double t)
{
  if (!ctx) return SHX_E_ARG;
  DShell s;
  if (int r = read_sh(ctx, s)) return r;
  const size_t bytes = sizeof(double) * 3 * ctx->C.nv;
  if (u) HIPCHK(hipMemcpy(ctx->P.U[s.par], u, bytes, hipMemcpyHostToDevice));
  if (v) HIPCHK(hipMemcpy(ctx->P.V[s.par], v, bytes, hipMemcpyHostToDevice));
  if (a) HIPCHK(hipMemcpy(ctx->P.A[s.par], a, bytes, hipMemcpyHostToDevice));
  s.t = t;
  return write_sh(ctx, s);
}

int shx_get_state(shx_ctx* ctx, double* u, double* v, double* a, double* t)
{
  if (!ctx) return SHX_E_ARG;
  DShell s;
  if (int r = read_sh(ctx, s)) return r;
  const size_t bytes = sizeof(double) * 3 * ctx->C.nv;
  if (u) HIPCHK(hipMemcpy(u, ctx->P.U[s.par], bytes, hipMemcpyDeviceToHost));
  if (v) HIPCHK(hipMemcpy(v, ctx->P.V[s.par], bytes, hipMemcpyDeviceToHost));
  if (a) HIPCHK(hipMemcpy(a, ctx->P.A[s.par], bytes, hipMemcpyDeviceToHost));
  if (t) *t = s.t;
  return SHX_OK;
}

int shx_forces(shx_ctx* ctx, int precision, const double* u_pred, const double* v, double t, double* f,
               double* f_ext, double* energies, double* elem_area)
{
  if (!ctx || !u_pred) return SHX_E_ARG;
  if (precision != SHX_FP64) { ctx->err = "precision not available"; return SHX_E_ARG; }
  DShell s;
  if (int r = read_sh(ctx, s)) return r;
  s.done = 0;
  if (int r = write_sh(ctx, s)) return r;
  const size_t nv = ctx->C.nv, ne = ctx->C.ne, bytes = sizeof(double) * 3 * nv;
  HIPCHK(hipMemcpy(ctx->P.U[s.par], u_pred, bytes, hipMemcpyHostToDevice));
  if (v) HIPCHK(hipMemcpy(ctx->P.V[s.par], v, bytes, hipMemcpyHostToDevice));
  const DPtrs& P = ctx->P;
  const SConst& C = ctx->C;
  k_elem<<<ctx->nblkE, kElemBlock, 0, ctx->st>>>(P, C, M_FORCES);
  if (C.contact) k_contact<<<ctx->nblkC, 256, 0, ctx->st>>>(P, C, M_FORCES);
  k_node<<<ctx->nblkN, kNodeBlock, 0, ctx->st>>>(P, C, M_FORCES, t, C.contact ? ctx->nblkC : 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->st));
  if (f) HIPCHK(hipMemcpy(f, P.fout, bytes, hipMemcpyDeviceToHost));
  if (f_ext) HIPCHK(hipMemcpy(f_ext, P.fext_out, bytes, hipMemcpyDeviceToHost));
  if (elem_area) HIPCHK(hipMemcpy(elem_area, P.area, sizeof(double) * ne, hipMemcpyDeviceToHost));
  if (int r = read_sh(ctx, s)) return r;
  if (energies) { energies[0] = s.e_tens; energies[1] = s.e_bend; energies[2] = s.e_self; energies[3] = s.e_fric; }
  return check_err(ctx, s);
}

int shx_contact_rebuild(shx_ctx* ctx, int force)
{
  if (!ctx) return SHX_E_ARG;
  if (!ctx->C.contact) return 0;
  DShell s;
  if (int r = read_sh(ctx, s)) return r;
  s.rebuild = force ? 1 : (s.du_sum >= ctx->C.cell_margin * (ctx->C.cell_size_mult * s.avg_elem_size));
  s.done = 0;
  if (int r = write_sh(ctx, s)) return r;
  if (int r = launch_broad(ctx, M_FORCES)) return r;
  if (int r = read_sh(ctx, s)) return r;
  prof_collect(ctx);
  if (int r = check_err(ctx, s)) return r;
  return s.n_pairs;
}

int shx_init_dynamics(shx_ctx* ctx, int precision)
{
  if (!ctx) return SHX_E_ARG;
  if (precision != SHX_FP64) { ctx->err = "precision not available"; return SHX_E_ARG; }
  DShell s;
  if (int r = read_sh(ctx, s)) return r;
  s.done = 0; s.n_steps = 0; s.n_rejected = 0; s.converged_step = -1; s.fresh = 1; s.attempts = 0;
  if (int r = write_sh(ctx, s)) return r;
  const DPtrs& P = ctx->P;
  const SConst& C = ctx->C;
  k_elem<<<ctx->nblkE, kElemBlock, 0, ctx->st>>>(P, C, M_INIT);
  HIPCHK(hipGetLastError());
  if (C.contact) {
    if (int r = read_sh(ctx, s)) return r;
    s.rebuild = 1;
    if (int r = write_sh(ctx, s)) return r;
    if (int r = launch_broad(ctx, M_INIT)) return r;
    k_contact<<<ctx->nblkC, 256, 0, ctx->st>>>(P, C, M_INIT);
  }
  if (int r = read_sh(ctx, s)) return r;
  k_node<<<ctx->nblkN, kNodeBlock, 0, ctx->st>>>(P, C, M_INIT, s.t + s.dt, C.contact ? ctx->nblkC : 0);
  HIPCHK(hipGetLastError());
  if (int r = read_sh(ctx, s)) return r;
  prof_collect(ctx);
  return check_err(ctx, s);
}

int shx_step(shx_ctx* ctx, int precision, int n_steps, double tol, int stop_on_converged, shx_report* rep)
{
  if (!ctx || n_steps < 0) return SHX_E_ARG;
  if (precision != SHX_FP64) { ctx->err = "precision not available"; return SHX_E_ARG; }
  DShell s;
  if (int r = read_sh(ctx, s)) return r;
  if (n_steps == 0) { fill_report(s, rep); return SHX_OK; }
  s.steps_target = s.n_steps + n_steps;
  s.done = 0;
  if (int r = write_sh(ctx, s)) return r;
  ctx->C.conv_tol = tol;
  ctx->C.stop_on_converged = stop_on_converged ? 1 : 0;
  const long long max_attempts = 64LL * n_steps + 4096;
  long long enq = 0;
  while (true) {
    const int remaining = s.steps_target - s.n_steps;
    const int chunk = std::max(1, std::min(remaining, 256));
    for (int k = 0; k < chunk; ++k)
      if (int r = enqueue_attempt(ctx)) return r;
    enq += chunk;
    if (int r = read_sh(ctx, s)) return r;
    prof_collect(ctx);
    if (s.done || s.err) break;
    if (enq > max_attempts) { ctx->err = "step attempts exceeded the bound"; return SHX_E_STATE; }
  }
  fill_report(s, rep);
  return check_err(ctx, s);
}

int shx_pairs(shx_ctx* ctx, uint32_t* e1, uint32_t* e2, int cap)
{
  if (!ctx) return SHX_E_ARG;
  if (!ctx->C.contact) return 0;
  DShell s;
  if (int r = read_sh(ctx, s)) return r;
  const int n = std::min(s.n_pairs, cap);
  if (n > 0 && e1) HIPCHK(hipMemcpy(e1, ctx->P.pe1[s.pair_buf], sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  if (n > 0 && e2) HIPCHK(hipMemcpy(e2, ctx->P.pe2[s.pair_buf], sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  return s.n_pairs;
}

int shx_elem_tables(shx_ctx* ctx, double* out)
{
  if (!ctx || !out) return SHX_E_ARG;
  const int ne = ctx->C.ne;
  std::vector<double> tab((size_t)T_ROWS * ne);
  HIPCHK(hipMemcpy(tab.data(), ctx->P.tab, sizeof(double) * tab.size(), hipMemcpyDeviceToHost));
  for (int e = 0; e < ne; ++e) {
    double* o = out + (size_t)101 * e;
    auto g = [&](int row) { return tab[(size_t)row * ne + e]; };
    const double R00 = g(T_R00), R01 = g(T_R01), R10 = g(T_R10), R11 = g(T_R11);
    const double* rb = &ctx->rawb[(ctx->T.valence[e] - 3) * 72];
    for (int i = 0; i < 12; ++i) {
      const double dx = rb[12 + i], de = rb[24 + i], dxx = rb[36 + i], dee = rb[48 + i], dxe = rb[60 + i];
      o[i] = rb[i];
      o[12 + i] = dx * R00 + de * R10;
      o[24 + i] = dx * R01 + de * R11;
      o[36 + i] = dxx * R00 * R00 + dee * R10 * R10 + 2 * dxe * R00 * R10;
      o[48 + i] = dxx * R01 * R01 + dee * R11 * R11 + 2 * dxe * R01 * R11;
      o[60 + i] = dxx * R00 * R01 + dee * R10 * R11 + dxe * R01 * R10 + dxe * R00 * R11;
    }
    for (int k = 0; k < 21; ++k) o[72 + k] = g(T_A1 + k);
    o[93] = g(T_JXW);
    o[94] = R00; o[95] = R01; o[96] = R10; o[97] = R11;
    for (int k = 0; k < 3; ++k) o[98 + k] = g(T_XYZ + k);
  }
  return SHX_OK;
}

int shx_mass(shx_ctx* ctx, double* Minv, double* Cd)
{
  if (!ctx) return SHX_E_ARG;
  const int nv = ctx->C.nv;
  std::vector<double> m(nv), c(nv);
  HIPCHK(hipMemcpy(m.data(), ctx->P.minv, sizeof(double) * nv, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(c.data(), ctx->P.cdamp, sizeof(double) * nv, hipMemcpyDeviceToHost));
  for (int n = 0; n < nv; ++n)
    for (int q = 0; q < 3; ++q) {
      if (Minv) Minv[3 * n + q] = m[n];
      if (Cd) Cd[3 * n + q] = c[n];
    }
  return SHX_OK;
}

int shx_patches(shx_ctx* ctx, int32_t* idx, int32_t* tri)
{
  if (!ctx) return SHX_E_ARG;
  if (idx) std::copy(ctx->T.patch.begin(), ctx->T.patch.end(), idx);
  if (tri) std::copy(ctx->T.tri.begin(), ctx->T.tri.end(), tri);
  return SHX_OK;
}

int shx_cfl_omega_max(shx_ctx* ctx, int iters, double* omega)
{
  if (!ctx || !omega || iters <= 0) return SHX_E_ARG;
  const int nv = ctx->C.nv, n3 = 3 * nv;
  DShell s;
  if (int r = read_sh(ctx, s)) return r;
  std::vector<double> u0(n3), x(n3), fp(n3), fm(n3), y(n3), up(n3), minv(n3);
  HIPCHK(hipMemcpy(u0.data(), ctx->P.U[s.par], sizeof(double) * n3, hipMemcpyDeviceToHost));
  if (int r = shx_mass(ctx, minv.data(), nullptr)) return r;
  SConst Ci = ctx->C;
  Ci.pressure_final = 0; Ci.bulk_modulus = 0; Ci.contact = 0;
  uint64_t st = 1;
  for (int i = 0; i < n3; ++i) {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    x[i] = (double)(st >> 11) / 9007199254740992.0 - 0.5;
  }
  const double eps = 1e-7 * ctx->C.h;
  double lam = 0;
  s.done = 0;
  if (int r = write_sh(ctx, s)) return r;
  auto fint = [&](const std::vector<double>& u, std::vector<double>& f) -> int {
    HIPCHK(hipMemcpy(ctx->P.U[s.par], u.data(), sizeof(double) * n3, hipMemcpyHostToDevice));
    k_elem<<<ctx->nblkE, kElemBlock, 0, ctx->st>>>(ctx->P, Ci, M_FORCES);
    k_node<<<ctx->nblkN, kNodeBlock, 0, ctx->st>>>(ctx->P, Ci, M_FORCES, 0.0, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(f.data(), ctx->P.fout, sizeof(double) * n3, hipMemcpyDeviceToHost));
    return 0;
  };
  for (int it = 0; it < iters; ++it) {
    double nx = 0;
    for (double q : x) nx += q * q;
    nx = std::sqrt(nx);
    for (double& q : x) q /= nx;
    for (int i = 0; i < n3; ++i) up[i] = eps * x[i];
    if (int r = fint(up, fp)) return r;
    for (int i = 0; i < n3; ++i) up[i] = -eps * x[i];
    if (int r = fint(up, fm)) return r;
    double num = 0;
    for (int i = 0; i < n3; ++i) {
      y[i] = minv[i] * (-(fp[i] - fm[i]) / (2 * eps));
      num += x[i] * y[i];
    }
    lam = num;
    x = y;
  }
  HIPCHK(hipMemcpy(ctx->P.U[s.par], u0.data(), sizeof(double) * n3, hipMemcpyHostToDevice));
  *omega = std::sqrt(std::fabs(lam));
  return SHX_OK;
}

int shx_set_profiling(shx_ctx* ctx, int enable)
{
  if (!ctx) return SHX_E_ARG;
  prof_collect(ctx);
  ctx->prof = enable != 0;
  for (int k = 0; k < 5; ++k) { ctx->tsum[k] = 0; ctx->tcnt[k] = 0; }
  return SHX_OK;
}

int shx_kernel_times(shx_ctx* ctx, double* avg_ms, int64_t* counts)
{
  if (!ctx) return SHX_E_ARG;
  prof_collect(ctx);
  for (int k = 0; k < 5; ++k) {
    if (avg_ms) avg_ms[k] = ctx->tcnt[k] ? ctx->tsum[k] / ctx->tcnt[k] : 0.0;
    if (counts) counts[k] = ctx->tcnt[k];
  }
  return SHX_OK;
}

void shx_tri_distance(const double* p, double* out)
{
  v3 A[3], B[3];
  for (int k = 0; k < 3; ++k) {
    A[k] = mk(p[3 * k], p[3 * k + 1], p[3 * k + 2]);
    B[k] = mk(p[9 + 3 * k], p[9 + 3 * k + 1], p[9 + 3 * k + 2]);
  }
  tri_hit h;
  tri_tri(A, B, h);
  out[0] = h.d2; out[1] = h.dc.x; out[2] = h.dc.y; out[3] = h.dc.z;
  out[4] = h.v0; out[5] = h.v1; out[6] = h.w0; out[7] = h.w1;
}

int shx_raw_elem_basis(int valence, double* out)
{
  double b[6 * kMaxPatch];
  if (!loop_basis_at_centroid(valence, b)) return SHX_E_ARG;
  const int P = valence + 6;
  for (int k = 0; k < 6; ++k)
    for (int j = 0; j < P; ++j) out[k * P + j] = b[k * kMaxPatch + j];
  return SHX_OK;
}

} // extern "C"
