/*
 * shx_run.cpp -- shellhip_run, the command-line driver of the explicit
 * shell path on one GPU: the reference's shell.C (shell.C:44-399) as a C++
 * program over the C-ABI of libshellhip (include/shellhip.h).
 *
 *   shellhip_run [-c configfile] [-l loadfile] -m meshfile.off [-d datadir] [-key value ...]
 *
 * It reads the config and its command-line overrides, prepares the mesh
 * (OFF file, midpoint refinements, projection, scale, rotation,
 * translation), applies the initial distortion (drand48 seeded with
 * random_seed, ShellSystem.C:863-869) or loads a state file, and steps the
 * shell with the reference's output cadence: a property line every
 * data_time, a .vtu file every vtk_time (data/shell%010u.vtu), a state
 * file every state_time wall seconds and at the end
 * (data/shell_state%010u.dat, linked as shell_state_latest.dat), until the
 * stopping criterion, time_steps, the wall-time limit, a signal or a
 * problem of severity STOP_SIM ends the run.  Exit codes as shell.C:23:
 * 0 success, 1 wall-time limit, 2 signal, 3 error.
 *
 * Differences from shell.C: a mesh file is required (the reference's
 * fallback, a Delaunay square with boundary ghosts, comes from libMesh's
 * mesh generators); randomness (libMesh's random node
 * distortion) is refused; the hard-coded change of vtk_time after t > 2
 * (shell.C:333) and the "Volume growth red" debug line of
 * print_properties are not reproduced.  Steps between outputs run on the
 * GPU in one call (their number follows from the time step when dt is
 * constant; one step per call under adaptive dt).
 */
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "shellhip.h"

namespace {

enum ExitCode { UNFINISHED = -1, SUCCESS = 0, TIME_LIMIT_REACHED, SIGNAL_CAUGHT, ERROR };
volatile std::sig_atomic_t g_signal = 0;

void catch_signal(int sig) { g_signal = sig; }

std::string now_string(const char* fmt)
{
  char b[64];
  std::time_t t = std::time(nullptr);
  std::strftime(b, sizeof(b), fmt, std::localtime(&t));
  return b;
}

bool read_file(const std::string& path, std::string& out)
{
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream s;
  s << f.rdbuf();
  out = s.str();
  return true;
}

bool copy_file(const std::string& from, const std::string& to_dir)
{
  std::string text;
  if (!read_file(from, text)) return false;
  const size_t slash = from.find_last_of('/');
  std::ofstream f(to_dir + "/" + (slash == std::string::npos ? from : from.substr(slash + 1)), std::ios::binary);
  f << text;
  return (bool)f;
}

void mkdir_p(const std::string& dir)
{
  std::string cur;
  std::stringstream ss(dir);
  std::string part;
  if (!dir.empty() && dir[0] == '/') cur = "/";
  while (std::getline(ss, part, '/')) {
    if (part.empty()) continue;
    cur += part + "/";
    mkdir(cur.c_str(), 0777);
  }
}

double seconds_since(const timespec& t0)
{
  timespec t1;
  clock_gettime(CLOCK_MONOTONIC, &t1);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

void print_header()
{
  /* ShellSystem::print_header (ShellSystem.h:454-464) */
  std::cout << "  timestamp    step       dt          time        wall_time "
            << "area        volume      density     pressure    Nc_cav      Nc_self     "
            << "E_kin       E_tens      E_bend      E_cav       E_self      E_fric      kGauss      kMean      \n"
            << "------------------------------------------------------------"
            << "------------------------------------------------------------------------"
            << "-----------------------------------------------------------------------------------------------"
            << std::endl;
}

/* ShellSystem::print_properties (ShellSystem.C:2849-2898) */
void print_properties(shx_ctx* ctx, unsigned time_step, double secs_running)
{
  double p[16];
  if (shx_batch_properties(ctx, 0, p) != SHX_OK) return;
  if (secs_running < 1) secs_running = std::floor(secs_running * 1e7) * 1e-7;
  char line[512];
  std::snprintf(line, sizeof(line),
                "%12s %10d %11.6g %11.6g %9.8g %11.6g %11.6g %11.6g %11.6g %11.6g %11.6g %11.6g %11.6g %11.6g "
                "%11.6g %11.6g %11.6g % 11.5g % 11.5g",
                now_string("%y%m%d%H%M%S").c_str(), (int)time_step, p[0], p[1], secs_running, p[2], p[3], p[4],
                p[5], p[6], p[7], p[8], p[9], p[10], p[11], p[12], p[13], p[14], p[15]);
  std::cout << "# " << line << std::endl;
}

const char* problem_name(int p)
{
  static const char* names[] = {"CLEAR: Problem resolved",
                                "WARNING: No change in displacement",
                                "WARNING: A physical solution is not guaranteed, because the increment is too large",
                                "ERROR: An unknown problem occurred",
                                "ERROR: The number of linked cells seems to have exploded",
                                "ERROR: The timestep is too small",
                                "ERROR: Problem in shell-shell contact",
                                "ERROR: The solution is not finite"};
  return (p >= 0 && p <= 7) ? names[p] : "ERROR: An unknown problem occurred";
}

} // namespace

int main(int argc, char** argv)
{
  const std::time_t start_time = std::time(nullptr);
  timespec start_mono;
  clock_gettime(CLOCK_MONOTONIC, &start_mono);
  std::string cfgname = "config.cfg", data_dir = "data", load_filename, mesh_filename;
  std::vector<const char*> overrides;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      std::cout << "Usage: " << argv[0] << " [-c configfile] [-l loadfile] [-m meshfile] [-d datadir] [options]\n";
      return SUCCESS;
    }
    const bool has_value = i + 1 < argc;
    if (a == "-c" && has_value) cfgname = argv[++i];
    else if (a == "-d" && has_value) data_dir = argv[++i];
    else if (a == "-l" && has_value) load_filename = argv[++i];
    else if (a == "-m" && has_value) mesh_filename = argv[++i];
    else overrides.push_back(argv[i]);
  }
  std::cout << "Invocation at " << now_string("%Y-%m-%d %H:%M:%S") << ":";
  for (int i = 0; i < argc; ++i) std::cout << " " << argv[i];
  std::cout << std::endl;
  std::signal(SIGHUP, SIG_IGN);
  std::signal(SIGQUIT, catch_signal);
  std::signal(SIGTERM, catch_signal);

  std::string cfg_text;
  if (!read_file(cfgname, cfg_text)) {
    std::cout << "ERROR: Can't open config file \"" << cfgname << "\" for reading!\n";
    return ERROR;
  }
  mkdir_p(data_dir);
  copy_file(cfgname, data_dir);
  shx_config cfg;
  if (shx_config_parse(cfg_text.c_str(), (int)overrides.size(), overrides.data(), &cfg) != SHX_OK) {
    std::cout << "ERROR: " << cfg.message << std::endl;
    return ERROR;
  }
  if (cfg.message[0]) std::cout << "WARNING: " << cfg.message << std::endl;

  unsigned time_step = 0;
  bool loaded = false;
  if (!load_filename.empty()) {
    std::ifstream lf(load_filename);
    if (!lf) {
      std::cout << "WARNING: Can't open load file \"" << load_filename << "\" for reading! Starting new simulation.\n";
    } else {
      lf >> time_step;
      loaded = true;
      if (cfg.time_steps && time_step >= (unsigned)cfg.time_steps) {
        std::cout << "ERROR: Loaded state is already finished!\n";
        return SUCCESS;
      }
    }
  }

  /* the mesh (shell.C:146-230) */
  if (mesh_filename.empty()) {
    std::cout << "ERROR: a triangle mesh is required (-m meshfile.off)\n";
    return ERROR;
  }
  std::string mesh_text;
  if (!read_file(mesh_filename, mesh_text)) {
    std::cout << "ERROR: Can't open mesh file \"" << mesh_filename << "\" for reading!\n";
    return ERROR;
  }
  copy_file(mesh_filename, data_dir);
  int nv = 0, ne = 0;
  if (shx_mesh_read_off(mesh_text.c_str(), &nv, &ne, nullptr, nullptr) != SHX_OK) {
    std::cout << "ERROR: \"" << mesh_filename << "\" is not an OFF triangle mesh\n";
    return ERROR;
  }
  std::vector<double> X(3 * (size_t)nv);
  std::vector<int32_t> tri(3 * (size_t)ne);
  shx_mesh_read_off(mesh_text.c_str(), &nv, &ne, X.data(), tri.data());
  if (cfg.refinements > 0) {
    int nv2 = 0, ne2 = 0;
    shx_mesh_refine(nv, ne, X.data(), tri.data(), cfg.refinements, &nv2, &ne2, nullptr, nullptr);
    std::vector<double> X2(3 * (size_t)nv2);
    std::vector<int32_t> T2(3 * (size_t)ne2);
    shx_mesh_refine(nv, ne, X.data(), tri.data(), cfg.refinements, &nv2, &ne2, X2.data(), T2.data());
    X.swap(X2);
    tri.swap(T2);
    nv = nv2;
    ne = ne2;
  }
  if (cfg.randomness != 0) {
    std::cout << "ERROR: randomness (libMesh's random node distortion) is not supported\n";
    return ERROR;
  }
  shx_mesh_transform(nv, X.data(), cfg.project, cfg.scale, cfg.rotate, cfg.translate);
  if (cfg.add_ghosts) {
    /* prepare_subdiv_mesh with ghosts (shell.C:228-230): an open mesh gets
     * its ghost layer, a closed one is unchanged */
    int nv2 = 0, ne2 = 0;
    if (shx_mesh_add_ghosts(nv, ne, X.data(), tri.data(), &nv2, &ne2, nullptr, nullptr) != SHX_OK) {
      std::cout << "ERROR: the mesh cannot take a ghost layer (a boundary vertex with more than 3 elements, or a "
                   "boundary touching itself)\n";
      return ERROR;
    }
    if (ne2 > ne) {
      std::vector<double> X2(3 * (size_t)nv2);
      std::vector<int32_t> T2(3 * (size_t)ne2);
      shx_mesh_add_ghosts(nv, ne, X.data(), tri.data(), &nv2, &ne2, X2.data(), T2.data());
      cfg.prm.ghost_elems = ne2 - ne;
      X.swap(X2);
      tri.swap(T2);
      nv = nv2;
      ne = ne2;
    }
  }

  shx_ctx* ctx = nullptr;
  if (shx_create(&ctx, nv, ne, X.data(), tri.data(), &cfg.prm, 0) != SHX_OK) {
    std::cout << "ERROR: " << shx_last_error(nullptr) << std::endl;
    return ERROR;
  }
  {
    double hmin = 1e300, hmax = 0, havg = 0;
    for (int e = 0; e < ne; ++e) {
      double lo = 1e300, hi = 0;
      for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b) {
          double d2 = 0;
          for (int c = 0; c < 3; ++c) {
            const double d = X[3 * tri[3 * e + a] + c] - X[3 * tri[3 * e + b] + c];
            d2 += d * d;
          }
          lo = std::min(lo, std::sqrt(d2));
          hi = std::max(hi, std::sqrt(d2));
        }
      hmin = std::min(hmin, lo);
      hmax = std::max(hmax, hi);
      havg += (lo + hi) / 2;
    }
    havg /= ne;
    std::cout << "\nnodes   : " << nv << "\nelements: " << ne << "\nDOFs    : " << 3 * nv << "\n"
              << "ELEMENT SIZE:\nmin: " << hmin << "\navg: " << havg << "\nmax: " << hmax << "\n" << std::endl;
  }
  std::cout.precision(17);

  const int precision = SHX_FP64;
  double wall_time = 0;
  if (loaded) {
    if (shx_state_load(ctx, 0, load_filename.c_str(), precision, &time_step, &wall_time) != SHX_OK) {
      std::cout << "ERROR: " << shx_io_last_error() << std::endl;
      shx_destroy(ctx);
      return ERROR;
    }
  } else {
    /* apply_initial_conditions: u += distortion (2 drand48() - 1) thickness
     * per degree of freedom (ShellSystem.C:863-869; srand48(random_seed),
     * ShellSystem.C:342-343) */
    srand48(cfg.random_seed);
    std::vector<double> u(3 * (size_t)nv, 0.0), z(u.size(), 0.0);
    for (int n = 0; n < nv; ++n)
      for (int c = 0; c < 3; ++c) u[3 * n + c] += cfg.distortion * (2 * drand48() - 1) * cfg.prm.thickness;
    if (shx_set_state(ctx, u.data(), z.data(), z.data(), 0.0) != SHX_OK || shx_init_dynamics(ctx, precision) != SHX_OK) {
      std::cout << "ERROR: " << shx_last_error(ctx) << std::endl;
      shx_destroy(ctx);
      return ERROR;
    }
  }
  auto running = [&]() { return wall_time + seconds_since(start_mono); };
  char path[1024];
  const std::string vtk_fmt = data_dir + "/shell%010u.vtu";
  if (cfg.data_time > 0) print_header();
  shx_report rep;
  shx_batch_reports(ctx, &rep);
  if (!loaded) {
    if (cfg.data_time > 0) {
      shx_batch_measure(ctx);
      print_properties(ctx, time_step, running());
    }
    if (cfg.vtk_time > 0) {
      std::snprintf(path, sizeof(path), vtk_fmt.c_str(), time_step);
      shx_write_vtu(ctx, 0, path);
    }
  }
  double t = rep.t;
  unsigned last_data = cfg.data_time > 0 ? (unsigned)(t / cfg.data_time) : 0;
  unsigned last_vtk = cfg.vtk_time > 0 ? (unsigned)(t / cfg.vtk_time) : 0;
  unsigned last_state = 0;
  int exit_code = UNFINISHED, last_problem = 0;
  const std::string crit = cfg.stopping_criterion;
  auto criterion = [&](double tt) {
    double v = 0;
    return !crit.empty() && shx_expr_eval(crit.c_str(), tt, &v) == SHX_OK && v != 0;
  };
  while (true) {
    if (criterion(t)) {
      exit_code = SUCCESS;
    } else if (cfg.walltime_limit > 0 && std::difftime(std::time(nullptr), start_time) >= cfg.walltime_limit) {
      exit_code = TIME_LIMIT_REACHED;
    } else if (g_signal) {
      std::cout << "Received signal " << (int)g_signal << ": " << strsignal(g_signal) << " ("
                << now_string("%Y-%m-%d %H:%M:%S") << ")" << std::endl;
      exit_code = SIGNAL_CAUGHT;
    } else {
      /* steps up to the next output, stop or step limit; the time the GPU
       * forms by repeated addition is formed the same way here */
      int k = 1;
      if (cfg.prm.adapt_dt_every == 0) {
        double tt = t;
        const double dt = rep.dt;
        for (k = 1; k < 20000; ++k) {
          tt += dt;
          if (cfg.data_time > 0 && (unsigned)(tt / cfg.data_time) > last_data) break;
          if (cfg.vtk_time > 0 && (unsigned)(tt / cfg.vtk_time) > last_vtk) break;
          if (cfg.time_steps && time_step + k >= (unsigned)cfg.time_steps) break;
          if (criterion(tt)) break;
        }
      }
      const int before = rep.n_steps;
      const int rc = shx_step(ctx, precision, k, 0.0, 0, &rep);
      if (rc != SHX_OK && rc != SHX_E_CAPACITY) {
        std::cout << "ERROR: " << shx_last_error(ctx) << std::endl;
        exit_code = ERROR;
      } else {
        time_step += (unsigned)(rep.n_steps - before);
        t = rep.t;
        if (rep.problem != last_problem) {
          std::cout << problem_name(rep.problem) << " in time step " << time_step << "!" << std::endl;
          last_problem = rep.problem;
        }
        const bool stop = rep.problem >= SHX_STOP_SIM || (cfg.prm.stop_on_problem && rep.problem > 0) ||
                          rc == SHX_E_CAPACITY;
        const unsigned dc = cfg.data_time > 0 ? (unsigned)(t / cfg.data_time) : 0;
        if (dc > last_data) {
          last_data = dc;
          print_properties(ctx, time_step, running());
        }
        const unsigned vc = cfg.vtk_time > 0 ? (unsigned)(t / cfg.vtk_time) : 0;
        if (vc > last_vtk) {
          last_vtk = vc;
          std::snprintf(path, sizeof(path), vtk_fmt.c_str(), time_step);
          shx_write_vtu(ctx, 0, path);
        }
        if (cfg.time_steps && time_step >= (unsigned)cfg.time_steps) exit_code = SUCCESS;
        if (stop) exit_code = ERROR;
      }
    }
    const unsigned sc = cfg.state_time > 0 ? (unsigned)(running() / cfg.state_time) : 0;
    if (exit_code > UNFINISHED || sc > last_state) {
      last_state = sc;
      char name[64];
      std::snprintf(name, sizeof(name), "shell_state%010u.dat", time_step);
      const std::string full = data_dir + "/" + name;
      if (shx_state_save(ctx, 0, full.c_str(), time_step, running()) == SHX_OK) {
        const std::string latest = data_dir + "/shell_state_latest.dat";
        unlink(latest.c_str());
        if (symlink(name, latest.c_str()) != 0) copy_file(full, data_dir);
      } else {
        std::cout << "ERROR: " << shx_io_last_error() << std::endl;
      }
      if (exit_code > UNFINISHED) break;
    }
  }
  shx_destroy(ctx);
  return exit_code;
}
