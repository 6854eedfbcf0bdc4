// congfd: host driver of the MI355X RHS / SSP-RK3 path.
//
// Stands in for the reference's zoneMain + BlockSolver::stepsLoopCFL
// (zone_FD/zoneMain.cpp:7-364, zone_FD/src/src/blockSolver.cpp:170-233): it
// reads the same 16-line info.txt, builds the case, advances with the CFL time
// step clamped to outputDt, prints and appends the per-step
// "time = ... dt=...  rhoLinf = ..." line to "<case> - <scheme> - <grid> - rholinf.txt",
// writes the solution at every output time, and writes timeInfo.txt.  The
// compute path is the C ABI of libcongfd (include/congfd.h).
//
// Solution files: CGNS with the reference's file names (Info::filename) and
// tree (cgnsio.cpp:3-101: vertex coordinates, cell-centred Density, XVelocity,
// [YVelocity, [ZVelocity,]] Pressure), written by cgns_writer.cpp without
// libcgns; --format tecplot writes the same variables as a Tecplot ASCII
// point file instead.  --initial-only writes the initial solution and stops
// before any device call (it runs without a GPU).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "cgns_writer.hpp"
#include "info.hpp"

using namespace cfdhost;

static void usage() {
  std::fprintf(stderr,
               "usage: congfd [info.txt] [--flux hllc|roe] [--recon char|component] [--inter N]\n"
               "              [--case N] [--dim N] [--imax I J K] [--zone x0 x1 y0 y1 z0 z1]\n"
               "              [--endstep N] [--outputdt T] [--cfl C] [--true-periodic] [--device N]\n"
               "              [--outdir DIR] [--no-output] [--fp32] [--gravity]\n"
               "              [--format cgns|tecplot] [--initial-only]\n");
}

#define CHECK(call)                                                                      \
  do {                                                                                   \
    int rc_ = (call);                                                                    \
    if (rc_ != 0) {                                                                      \
      std::fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, cfd_last_error());        \
      return 1;                                                                          \
    }                                                                                    \
  } while (0)

static std::string write_cgns_file(const std::string& path, const Info& info, const cfd_config& cfg,
                                   const std::vector<double>& cons) {
  const int dim = info.dim, nv = dim + 2;
  CgnsSolution s;
  s.dim = dim;
  s.h = info.geth();
  for (int d = 0; d < dim; d++) {
    s.vertices[d] = cfg.n[d] + 1;
    s.lo[d] = info.calZone[2 * d];
  }
  const long N = (long)cfg.n[0] * cfg.n[1] * cfg.n[2];
  s.fields.assign(nv, std::vector<double>(N));
  for (long c = 0; c < N; c++) {
    // Equation::consToPrimEuler1D/2D (zone_FD/src/src/equation.cpp:36-92), the
    // primitive state the reference writes; 3D extends the 2D expression
    const double* u = &cons[c * nv];
    const double r = u[0];
    double vel[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) vel[d] = u[1 + d] / r;
    const double E = u[nv - 1] / r;
    double e;
    if (dim == 1) e = -vel[0] * vel[0] / 2 + E;
    else if (dim == 2) e = -((vel[0] * vel[0] + vel[1] * vel[1]) / 2) + E;
    else e = -((vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]) / 2) + E;
    s.fields[0][c] = r;
    for (int d = 0; d < dim; d++) s.fields[1 + d][c] = vel[d];
    s.fields[nv - 1][c] = r * ((1.4 - 1) * e);
  }
  return write_cgns(path, s);
}

static void write_tecplot(const std::string& path, const Info& info, const cfd_config& cfg,
                          const std::vector<double>& cons) {
  const int dim = info.dim, nv = dim + 2;
  const double h = info.geth(), cmin = info.calZone[0];
  std::FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return;
  const char* coord[3] = {"CoordinateX", "CoordinateY", "CoordinateZ"};
  const char* vel[3] = {"XVelocity", "YVelocity", "ZVelocity"};
  std::fprintf(f, "VARIABLES =");
  for (int d = 0; d < dim; d++) std::fprintf(f, " \"%s\"", coord[d]);
  std::fprintf(f, " \"Density\"");
  for (int d = 0; d < dim; d++) std::fprintf(f, " \"%s\"", vel[d]);
  std::fprintf(f, " \"Pressure\"\nZONE T=\"cells t=%.4f\", I=%d", info.t, cfg.n[0]);
  if (dim > 1) std::fprintf(f, ", J=%d", cfg.n[1]);
  if (dim > 2) std::fprintf(f, ", K=%d", cfg.n[2]);
  std::fprintf(f, ", DATAPACKING=POINT\n");
  const long N = (long)cfg.n[0] * cfg.n[1] * cfg.n[2];
  for (long c = 0; c < N; c++) {
    long idx[3] = {c % cfg.n[0], (c / cfg.n[0]) % cfg.n[1], c / ((long)cfg.n[0] * cfg.n[1])};
    for (int d = 0; d < dim; d++) std::fprintf(f, "%.10e ", cmin + (idx[d] + 0.5) * h);
    const double* u = &cons[c * nv];
    double r = u[0], ke = 0;
    std::fprintf(f, "%.16e ", r);
    for (int d = 0; d < dim; d++) {
      double vv = u[1 + d] / r;
      ke += vv * vv;
      std::fprintf(f, "%.16e ", vv);
    }
    std::fprintf(f, "%.16e\n", (1.4 - 1) * (u[nv - 1] - 0.5 * r * ke));
  }
  std::fclose(f);
}

int main(int argc, char** argv) {
  Info info;
  std::string infoPath = "info.txt", outdir = ".";
  bool output = true, initialOnly = false, cgns = true;
  std::vector<std::string> args(argv + 1, argv + argc);
  // positional info file first, then overrides
  size_t a = 0;
  if (a < args.size() && args[a].rfind("--", 0) != 0) infoPath = args[a++];
  if (readInfoTxt(infoPath, info)) std::printf("file mode initialization finished\n");
  else std::printf("file mode initialization failed\n");
  for (; a < args.size(); a++) {
    const std::string& s = args[a];
    auto next = [&]() -> const char* {
      if (a + 1 >= args.size()) { usage(); std::exit(2); }
      return args[++a].c_str();
    };
    if (s == "--flux") { std::string v = next(); info.fluxMethod = v == "roe" ? CFD_FLUX_ROE : CFD_FLUX_HLLC; }
    else if (s == "--recon") { std::string v = next(); info.recon = v == "component" ? CFD_RECON_COMPONENT : CFD_RECON_CHAR; }
    else if (s == "--inter") info.interMethod = std::atoi(next());
    else if (s == "--case") info.nCase = std::atoi(next());
    else if (s == "--dim") info.dim = std::atoi(next());
    else if (s == "--imax") { for (int d = 0; d < 3; d++) info.iMax[d] = std::atoi(next()); }
    else if (s == "--zone") { for (int d = 0; d < 6; d++) info.calZone[d] = std::atof(next()); }
    else if (s == "--endstep") info.endStep = std::atoi(next());
    else if (s == "--outputdt") info.outputDt = std::atof(next());
    else if (s == "--cfl") info.CFL = std::atof(next());
    else if (s == "--true-periodic") info.truePeriodic = true;
    else if (s == "--device") info.device = std::atoi(next());
    else if (s == "--outdir") outdir = next();
    else if (s == "--no-output") output = false;
    else if (s == "--fp32") info.precision = CFD_PRECISION_FP32;
    else if (s == "--gravity") info.gravity = true;
    else if (s == "--format") { std::string v = next(); cgns = v != "tecplot"; }
    else if (s == "--initial-only") initialOnly = true;
    else { usage(); return 2; }
  }

  cfd_config cfg;
  std::memset(&cfg, 0, sizeof(cfg));
  cfg.dim = info.dim;
  cfg.inter_method = info.interMethod;
  cfg.flux_method = info.fluxMethod;
  cfg.diff_method = info.diffMethod;
  cfg.recon = info.recon;
  cfg.device = info.device;
  cfg.precision = info.precision;
  std::vector<double> cons;
  std::string why = setupCase(info, cons, cfg);
  if (!why.empty()) {
    std::fprintf(stderr, "congfd: %s\n", why.c_str());
    return 2;
  }
  auto base = [&]() { return outdir + "/" + info.filename(); };
  auto write_solution = [&]() -> int {
    if (cgns) {
      std::string err = write_cgns_file(base() + ".cgns", info, cfg, cons);
      if (!err.empty()) {
        std::fprintf(stderr, "congfd: %s\n", err.c_str());
        return 1;
      }
    } else {
      write_tecplot(base() + ".dat", info, cfg, cons);
    }
    return 0;
  };
  if (initialOnly) return write_solution();

  cfd_handle* h = nullptr;
  CHECK(cfd_create(&cfg, &h));
  CHECK(cfd_set_state(h, cons.data()));
  CHECK(cfd_set_time(h, info.t));
  std::printf("initialized %s, %s, %s, %d cells\n", info.caseName().c_str(), info.schemeName().c_str(),
              info.gridInfo().c_str(), cfg.n[0] * cfg.n[1] * cfg.n[2]);

  std::string logName = outdir + "/" + info.caseName() + " - " + info.schemeName() + " - " + info.gridInfo() + " - rholinf.txt";
  auto dump = [&]() -> int {
    if (!output) return 0;
    CHECK(cfd_get_state(h, cons.data()));
    return write_solution();
  };

  // BlockSolver::stepsLoopCFL (blockSolver.cpp:170-233): the dt clamp, the
  // RK3 steps, the clock and the per-step log values run on the device up to
  // the next output time (cfd_steps_to_output); the host prints the log lines
  // and writes the solution.
  const double eps = 1e-15;
  long timesteps = 0;
  long long solverMs = 0;
  const int maxSteps = 256;
  std::vector<double> steplog(3 * maxSteps);
  auto t0 = std::chrono::high_resolution_clock::now();
  while (info.step < info.endStep) {
    if (info.outputT < eps && dump()) return 1;
    int n = 0, reached = 0;
    auto s0 = std::chrono::high_resolution_clock::now();
    CHECK(cfd_steps_to_output(h, info.CFL, info.outputDt, &info.outputT, maxSteps, steplog.data(), &n, &reached));
    auto s1 = std::chrono::high_resolution_clock::now();
    solverMs += std::chrono::duration_cast<std::chrono::milliseconds>(s1 - s0).count();
    std::ofstream lf;
    if (output) lf.open(logName, std::ios::app);
    for (int i = 0; i < n; i++) {
      info.t = steplog[3 * i];
      char line[160];
      std::snprintf(line, sizeof(line), "time = %.4f dt=%.10f  rhoLinf = %.4f  \n", info.t, steplog[3 * i + 1],
                    steplog[3 * i + 2]);
      std::fputs(line, stdout);
      if (output) lf << line;
    }
    timesteps += n;
    if (reached) info.step++;
  }
  if (dump()) return 1;
  auto t1 = std::chrono::high_resolution_clock::now();
  long long total = std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
  std::printf("totaltime= %lld   Finish\ntimesteps= %ld   Finish\nsolvertime= %lld\n", total, timesteps, solverMs);
  if (output) {
    std::ofstream ti(outdir + "/timeInfo.txt");
    ti << info.interMethod << "\n"
       << "totaltime= " << total << "   Finish\n"
       << "timesteps= " << timesteps << "   Finish\n"
       << "solvertime= " << solverMs << "\n";
  }
  cfd_destroy(h);
  return 0;
}
