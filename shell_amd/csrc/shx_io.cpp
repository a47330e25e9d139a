/*
 * shx_io.cpp -- the reference's on-disk formats on this path, written
 * against the public C-ABI: the state file of ShellSystem::save / load
 * (ShellSystem.C:2956-3153, with the driver's leading time step,
 * shell.C:380-387) and the .vtu file of ShellSystem::write_vtk
 * (ShellSystem.C:3156-3320).
 *
 * State file deviations, each deliberate:
 *  - the measurement block: the GPU keeps the dt-weighted sums of the
 *    samples and their count, not every sample; it is written as that many
 *    lines, each the mean sample, so that any reader of the reference's
 *    format forms the same averages;
 *  - the reference's load reads the Nc column twice (the Nc_cav column into
 *    avg_Nc, ShellSystem.C:3089-3090); this reader takes the columns as
 *    save writes them;
 *  - dynamic growth data (per element alpha, beta, G) follow the sticking
 *    contacts as in the reference; after a load, the growth state and the
 *    cavity springs are set after the initialisation, as load leaves them
 *    (the reference's initialisation does not evaluate forces on a loaded
 *    state).
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "shellhip.h"

namespace {
thread_local std::string g_io_err;
int io_fail(const std::string& m, int code = SHX_E_ARG)
{
  g_io_err = m;
  return code;
}
}

extern "C" {

const char* shx_io_last_error(void) { return g_io_err.c_str(); }

int shx_state_save(shx_ctx* ctx, int shell, const char* path, uint32_t time_step, double wall_time)
{
  if (!ctx || !path) return io_fail("invalid argument");
  int nv = 0, ne = 0;
  if (int r = shx_batch_mesh(ctx, shell, &nv, &ne, nullptr, nullptr)) return io_fail(shx_last_error(ctx), r);
  std::vector<double> u(3 * (size_t)nv), v(u.size()), a(u.size());
  double t = 0;
  if (int r = shx_batch_get_state(ctx, shell, u.data(), v.data(), a.data(), &t)) return io_fail(shx_last_error(ctx), r);
  std::vector<shx_report> reps(shx_batch_size(ctx));
  if (int r = shx_batch_reports(ctx, reps.data())) return io_fail(shx_last_error(ctx), r);
  double sums[14];
  int n = 0;
  if (int r = shx_batch_averages(ctx, shell, sums, &n, 0)) return io_fail(shx_last_error(ctx), r);
  std::vector<double> cso(3 * (size_t)nv);
  if (int r = shx_batch_cavity_springs(ctx, shell, cso.data(), 0)) return io_fail(shx_last_error(ctx), r);
  const int np = shx_batch_pair_state(ctx, shell, nullptr, nullptr, nullptr, nullptr, nullptr, 0);
  if (np < 0) return io_fail(shx_last_error(ctx), np);
  std::vector<uint32_t> e1(np > 0 ? np : 1), e2(e1.size());
  std::vector<int32_t> st(e1.size());
  std::vector<double> so(3 * e1.size());
  if (np > 0) shx_batch_pair_state(ctx, shell, e1.data(), e2.data(), nullptr, st.data(), so.data(), np);
  std::ofstream f(path);
  if (!f) return io_fail(std::string("cannot write ") + path);
  f << std::setprecision(std::numeric_limits<double>::digits10 + 2);
  f << time_step << "\n";
  f << reps[shell].dt << "\n" << t << "\n" << wall_time << "\n";
  for (int i = 0; i < nv; ++i) {
    for (int c = 0; c < 3; ++c) f << u[3 * i + c] << "\t" << v[3 * i + c] << "\t" << a[3 * i + c] << "\t";
    f << cso[3 * i] << "\t" << cso[3 * i + 1] << "\t" << cso[3 * i + 2] << "\n";
  }
  /* columns dt, area, vol, pcav, Nc_cav, Nc, Ekin, Etens, Ebend, Ecav, Eself,
   * Efric, kGauss, kMean; each line the mean sample */
  f << n << "\n";
  for (int k = 0; k < n; ++k) {
    double m[14];
    for (int q = 0; q < 14; ++q) m[q] = sums[q] / n;
    f << m[0] << "\t" << m[1] << "\t" << m[2] << "\t" << m[11] << "\t" << m[12] << "\t" << m[3] << "\t" << m[4]
      << "\t" << m[5] << "\t" << m[6] << "\t" << m[13] << "\t" << m[7] << "\t" << m[8] << "\t" << m[9] << "\t"
      << m[10] << "\n";
  }
  int nstick = 0;
  for (int k = 0; k < np; ++k) nstick += st[k] != 0;
  f << nstick << "\n";
  for (int k = 0; k < np; ++k)
    if (st[k])
      f << e1[k] << "\t" << e2[k] << "\t" << so[3 * k] << "\t" << so[3 * k + 1] << "\t" << so[3 * k + 2] << "\n";
  shx_params prm;
  if (int r = shx_batch_params(ctx, shell, &prm)) return io_fail(shx_last_error(ctx), r);
  if (prm.metric_growth_rate != 0 || prm.curvature_growth_rate != 0) {
    std::vector<double> g(9 * (size_t)ne);
    if (int r = shx_batch_growth_state(ctx, shell, g.data(), 0)) return io_fail(shx_last_error(ctx), r);
    for (int e = 0; e < ne; ++e) {
      const double* q = &g[9 * (size_t)e];
      f << q[0] << "\t" << q[1] << "\t" << q[2] << "\t" << q[3] << "\t" << q[4] << "\t" << q[5] << "\t" << q[6]
        << "\t" << q[7] << "\t" << q[8] << "\n";
    }
  }
  if (!f) return io_fail(std::string("write failed: ") + path);
  return SHX_OK;
}

int shx_state_load(shx_ctx* ctx, int shell, const char* path, int precision, uint32_t* time_step, double* wall_time)
{
  if (!ctx || !path) return io_fail("invalid argument");
  int nv = 0, ne = 0;
  if (int r = shx_batch_mesh(ctx, shell, &nv, &ne, nullptr, nullptr)) return io_fail(shx_last_error(ctx), r);
  shx_params prm;
  if (int r = shx_batch_params(ctx, shell, &prm)) return io_fail(shx_last_error(ctx), r);
  std::ifstream f(path);
  if (!f) return io_fail(std::string("cannot read ") + path);
  uint32_t step = 0;
  double dt = 0, t = 0, wall = 0;
  f >> step >> dt >> t >> wall;
  std::vector<double> u(3 * (size_t)nv), v(u.size()), a(u.size()), cso(u.size());
  for (int i = 0; i < nv; ++i) {
    for (int c = 0; c < 3; ++c) f >> u[3 * i + c] >> v[3 * i + c] >> a[3 * i + c];
    f >> cso[3 * i] >> cso[3 * i + 1] >> cso[3 * i + 2];
  }
  int n = 0;
  f >> n;
  if (!f || n < 0) return io_fail(std::string("malformed state file ") + path, SHX_E_ARG);
  double sums[14] = {};
  for (int k = 0; k < n; ++k) {
    double c[14];
    for (double& x : c) f >> x;
    const double smp[14] = {c[0], c[1], c[2], c[5], c[6], c[7], c[8], c[10], c[11], c[12], c[13], c[3], c[4], c[9]};
    for (int q = 0; q < 14; ++q) sums[q] += smp[q];
  }
  int nc = 0;
  f >> nc;
  if (!f || nc < 0) return io_fail(std::string("malformed state file ") + path, SHX_E_ARG);
  std::vector<uint32_t> e1(nc > 0 ? nc : 1), e2(e1.size());
  std::vector<double> so(3 * e1.size());
  for (int k = 0; k < nc; ++k) f >> e1[k] >> e2[k] >> so[3 * k] >> so[3 * k + 1] >> so[3 * k + 2];
  if (!f) return io_fail(std::string("malformed state file ") + path, SHX_E_ARG);
  const bool dyn = prm.metric_growth_rate != 0 || prm.curvature_growth_rate != 0;
  std::vector<double> grow(dyn ? 9 * (size_t)ne : 0);
  for (double& x : grow) f >> x;
  if (!f) return io_fail(std::string("malformed state file ") + path + " (growth data)", SHX_E_ARG);
  /* load (ShellSystem.C:3050-3153), then apply_initial_conditions without
   * the initial acceleration (:855-905): the loaded a is kept; V0 becomes
   * the loaded state's enclosed volume, as in the reference */
  if (prm.adapt_dt_every > 0)
    if (int r = shx_batch_set_dt(ctx, shell, dt)) return io_fail(shx_last_error(ctx), r);
  if (int r = shx_batch_set_state(ctx, shell, u.data(), v.data(), a.data(), t)) return io_fail(shx_last_error(ctx), r);
  if (prm.contact && nc > 0)
    if (int r = shx_batch_load_pairs(ctx, shell, nc, e1.data(), e2.data(), so.data()))
      return io_fail(shx_last_error(ctx), r);
  if (int r = shx_batch_init_shells(ctx, precision, 1, &shell, nullptr)) return io_fail(shx_last_error(ctx), r);
  if (int r = shx_batch_set_state(ctx, shell, u.data(), v.data(), a.data(), t)) return io_fail(shx_last_error(ctx), r);
  /* the loaded cavity springs, as load leaves them (the initialisation's
   * force evaluation does not run on a loaded state, ShellSystem.C:887-896) */
  if (prm.cavity_contact)
    if (int r = shx_batch_cavity_springs(ctx, shell, cso.data(), 1)) return io_fail(shx_last_error(ctx), r);
  if (dyn)
    if (int r = shx_batch_growth_state(ctx, shell, grow.data(), 1)) return io_fail(shx_last_error(ctx), r);
  if (int r = shx_batch_averages(ctx, shell, sums, &n, 1)) return io_fail(shx_last_error(ctx), r);
  if (time_step) *time_step = step;
  if (wall_time) *wall_time = wall;
  return SHX_OK;
}

int shx_write_vtu(shx_ctx* ctx, int shell, const char* path)
{
  if (!ctx || !path) return io_fail("invalid argument");
  int nv = 0, ne = 0;
  if (int r = shx_batch_mesh(ctx, shell, &nv, &ne, nullptr, nullptr)) return io_fail(shx_last_error(ctx), r);
  std::vector<double> X(3 * (size_t)nv), u(3 * (size_t)nv), nf(5 * (size_t)nv), ef(4 * (size_t)ne);
  std::vector<int32_t> tri(3 * (size_t)ne);
  shx_batch_mesh(ctx, shell, nullptr, nullptr, X.data(), tri.data());
  if (int r = shx_batch_get_state(ctx, shell, u.data(), nullptr, nullptr, nullptr)) return io_fail(shx_last_error(ctx), r);
  if (int r = shx_batch_fields(ctx, shell, nf.data(), ef.data())) return io_fail(shx_last_error(ctx), r);
  /* the real nodes: an open mesh's ghost layer follows them and is not
   * written (write_vtk, ShellSystem.C:3189) */
  int nw = 0;
  for (const int32_t q : tri) nw = std::max(nw, q + 1);
  std::ofstream file(path);
  if (!file) return io_fail(std::string("cannot write ") + path);
  unsigned char count;                         /* a line break after every 256 values */
  auto array = [&](const char* type, const char* name, const char* extra) {
    file << "        <DataArray type=\"" << type << "\" Name=\"" << name << "\"" << extra << " format=\"ascii\">\n";
    count = 0;
  };
  auto put = [&](auto x) {
    file << x << " ";
    if (++count == 0) file << "\n";
  };
  auto end = [&]() { file << "\n" << "        </DataArray>\n"; };
  file << "<?xml version=\"1.0\"?>\n";
  file << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
  file << "  <UnstructuredGrid>\n";
  file << "    <Piece NumberOfPoints=\"" << nw << "\" NumberOfCells=\"" << ne << "\">\n";
  file << "      <PointData>\n";
  const char* var[3] = {"ux", "uy", "uz"};
  for (int c = 0; c < 3; ++c) {
    array("Float64", var[c], "");
    for (int i = 0; i < nw; ++i) put(u[3 * i + c]);
    end();
  }
  const char* nname[5] = {"ecav", "efric", "ekin", "eself", "mass"};
  for (int k = 0; k < 5; ++k) {
    array("Float64", nname[k], "");
    for (int i = 0; i < nw; ++i) put(nf[5 * i + k]);
    end();
  }
  file << "      </PointData>\n";
  file << "      <CellData>\n";
  const char* ename[4] = {"ebend", "etens", "kGauss", "kMean"};
  for (int k = 0; k < 4; ++k) {
    array("Float64", ename[k], "");
    for (int e = 0; e < ne; ++e) put(ef[4 * e + k]);
    end();
  }
  file << "      </CellData>\n";
  file << "      <Points>\n";
  array("Float64", "Points", " NumberOfComponents=\"3\"");
  for (size_t i = 0; i < 3 * (size_t)nw; ++i) put(X[i]);
  end();
  file << "      </Points>\n";
  file << "      <Cells>\n";
  array("Int64", "connectivity", "");
  for (size_t i = 0; i < tri.size(); ++i) put(tri[i]);
  end();
  array("Int64", "offsets", "");
  for (int i = 3; i <= 3 * ne; i += 3) put(i);
  end();
  array("UInt8", "types", "");
  for (int e = 0; e < ne; ++e) put("5");
  end();
  file << "      </Cells>\n";
  file << "    </Piece>\n";
  file << "  </UnstructuredGrid>\n";
  file << "</VTKFile>\n";
  if (!file) return io_fail(std::string("write failed: ") + path);
  return SHX_OK;
}

} // extern "C"
