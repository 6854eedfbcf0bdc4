// Initial conditions and boundary wiring of the driver's cases.
// 1D / 2D presets follow Initializer::solInit (zone_FD/src/src/initializer.cpp:9-560)
// and Initializer::initBnds (:743-952); cell centres follow
// Initializer::initUniformBlock (:573-711), whose cell-centre loop offsets every
// direction by calZone[0] (the x minimum) and uses the x spacing.
#include <cmath>
#include <cstdint>
#include <vector>

#include "info.hpp"

namespace cfdhost {

static const double G = 1.4;

static void put(std::vector<double>& c, std::initializer_list<double> v) {
  for (double x : v) c.push_back(x);
}

static void set_bc(cfd_config& cfg, int ax, int lo, int hi) {
  cfg.bc[2 * ax] = lo;
  cfg.bc[2 * ax + 1] = hi;
}

// splitmix64 (Steele, Lea, Flood 2014): deterministic phases for the 3D case
static uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Real field on an n^3 periodic grid: sum of unit-amplitude cosines with
// random phases over 1 <= |k| <= kmax (evaluated by three separable passes),
// scaled to max |f| = 1.
static std::vector<double> random_phase(int n, int kmax, uint64_t& seed) {
  const int K = 2 * kmax + 1;
  const double two_pi = 2 * M_PI;
  std::vector<double> cr((size_t)K * K * K), ci((size_t)K * K * K);
  for (int a = 0; a < K; a++)
    for (int b = 0; b < K; b++)
      for (int c = 0; c < K; c++) {
        int kz = a - kmax, ky = b - kmax, kx = c - kmax;
        double kk = std::sqrt((double)(kx * kx + ky * ky + kz * kz));
        double ph = (splitmix64(seed) >> 11) * (1.0 / 9007199254740992.0) * two_pi;
        double amp = (kk >= 1 && kk <= kmax) ? 1.0 : 0.0;
        size_t i = ((size_t)a * K + b) * K + c;
        cr[i] = amp * std::cos(ph);
        ci[i] = amp * std::sin(ph);
      }
  std::vector<double> er((size_t)K * n), ei((size_t)K * n);
  for (int k = 0; k < K; k++)
    for (int x = 0; x < n; x++) {
      double ang = two_pi * (k - kmax) * x / n;
      er[(size_t)k * n + x] = std::cos(ang);
      ei[(size_t)k * n + x] = std::sin(ang);
    }
  // pass x: g[a][b][x]
  std::vector<double> gr((size_t)K * K * n, 0.0), gi((size_t)K * K * n, 0.0);
  for (int ab = 0; ab < K * K; ab++)
    for (int c = 0; c < K; c++) {
      double r = cr[(size_t)ab * K + c], im = ci[(size_t)ab * K + c];
      if (r == 0 && im == 0) continue;
      for (int x = 0; x < n; x++) {
        double e1 = er[(size_t)c * n + x], e2 = ei[(size_t)c * n + x];
        gr[(size_t)ab * n + x] += r * e1 - im * e2;
        gi[(size_t)ab * n + x] += r * e2 + im * e1;
      }
    }
  // pass y: h[a][y][x]
  std::vector<double> hr((size_t)K * n * n, 0.0), hi((size_t)K * n * n, 0.0);
  for (int a = 0; a < K; a++)
    for (int b = 0; b < K; b++)
      for (int y = 0; y < n; y++) {
        double e1 = er[(size_t)b * n + y], e2 = ei[(size_t)b * n + y];
        const double* pr = &gr[((size_t)a * K + b) * n];
        const double* pi = &gi[((size_t)a * K + b) * n];
        double* qr = &hr[((size_t)a * n + y) * n];
        double* qi = &hi[((size_t)a * n + y) * n];
        for (int x = 0; x < n; x++) {
          qr[x] += pr[x] * e1 - pi[x] * e2;
          qi[x] += pr[x] * e2 + pi[x] * e1;
        }
      }
  // pass z: real part
  std::vector<double> f((size_t)n * n * n, 0.0);
  for (int a = 0; a < K; a++)
    for (int z = 0; z < n; z++) {
      double e1 = er[(size_t)a * n + z], e2 = ei[(size_t)a * n + z];
      const double* pr = &hr[(size_t)a * n * n];
      const double* pi = &hi[(size_t)a * n * n];
      double* o = &f[(size_t)z * n * n];
      for (size_t yx = 0; yx < (size_t)n * n; yx++) o[yx] += pr[yx] * e1 - pi[yx] * e2;
    }
  double m = 0;
  for (double v : f) m = std::fmax(m, std::fabs(v));
  if (m > 0)
    for (double& v : f) v /= m;
  return f;
}

std::string setupCase(const Info& info, std::vector<double>& cons, cfd_config& cfg) {
  auto ic = info.icMax();
  const double h = info.geth();
  const double cmin = info.calZone[0];
  cons.clear();
  for (int d = 0; d < 3; d++) {
    cfg.n[d] = d < info.dim ? ic[d] : 1;
    cfg.h[d] = h;
  }
  for (int f = 0; f < 6; f++) {
    cfg.bc[f] = CFD_BC_OUTFLOW;
    for (int v = 0; v < 5; v++) cfg.bc_prim[f][v] = 0.0;
  }
  auto xc = [&](int i) { return cmin + (i + 0.5) * h; };
  cfg.origin = cmin;
  for (int d = 0; d < 3; d++) cfg.gravity[d] = 0.0;

  if (info.dim == 1) {
    const int n = ic[0];
    for (int i = 0; i < n; i++) {
      double x = xc(i);
      switch (info.nCase) {
      case 0:  // Sod (initializer.cpp:79-103)
        if (x < 0) put(cons, {1, 0, 1.0 / (G - 1) * 1});
        else put(cons, {0.125, 0, 1.0 / (G - 1) * 0.1});
        break;
      case 1:  // Shu-Osher (:105-133)
        if (x <= 0.5) put(cons, {3.857143, 3.857143 * 2.629369, 1.0 / (G - 1.0) * (10.0 + 1.0 / 3.0) + 3.857143 * 2.629369 * 2.629369 / 2});
        else put(cons, {1.0 + 0.2 * std::sin(5.0 * x), 0.0, 1.0 / (G - 1.0) * 1.0});
        break;
      case 2: {  // Lax (:135-159)
        double r = x <= 0 ? 0.445 : 0.5, u = x <= 0 ? 0.698 : 0, p = x <= 0 ? 3.528 : 0.571;
        put(cons, {r, r * u, 1.0 / (G - 1) * p + r * u * u / 2});
        break;
      }
      case 3: {  // Sedov (:161-188)
        double E = std::fabs(x) < 1e-10 ? 3200000.0 / (xc(i) - xc(i - 1)) : 1e-12;
        put(cons, {1, 0, 1 * E});
        break;
      }
      case 4: {  // Woodward-Colella (:190-218)
        double p = x < 0.1 ? 1000 : (x >= 0.9 ? 100 : 0.01);
        put(cons, {1, 0, 1.0 / (G - 1) * p});
        break;
      }
      case 5: {  // double rarefaction (:220-242)
        double u = x < 0 ? -2 : 2;
        put(cons, {1, u, 1.0 / (G - 1) * 0.4 + u * u / 2});
        break;
      }
      default: return "1D case " + std::to_string(info.nCase) + " is not defined";
      }
    }
    int b = info.nCase == 4 ? CFD_BC_SYMMETRY1D : CFD_BC_OUTFLOW;
    set_bc(cfg, 0, b, b);
    return "";
  }

  if (info.dim == 2) {
    const int nx = ic[0], ny = ic[1];
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        double x = xc(i), y = xc(j);
        double r = 1, u = 0, v = 0, p = 1;
        switch (info.nCase) {
        case 0:  // 2D Riemann (:247-288); states written as in the reference
          if (x > 0.3) {
            if (y > 0.3) { put(cons, {1.5, 0, 0, 1.0 / (G - 1) * 1.5}); }
            else { put(cons, {0.5323, 0, 0.5323 * 1.206, 1.0 / (G - 1) * 0.3 + 1.206 * 1.206 / 2 * 0.5323}); }
          } else {
            if (y > 0.3) { put(cons, {0.5323, 0.5323 * 1.206, 0, 1.0 / (G - 1) * 0.3 + 1.206 * 1.206 / 2 * 0.5323}); }
            else { put(cons, {0.138, 0.138 * 1.206, 0.138 * 1.206, 1.0 / (G - 1) * 0.029 + 1.206 * 1.206 * 0.138}); }
          }
          continue;
        case 1:  // 2D Riemann vortex (:289-352)
          if (x > 0.0) { if (y > 0.0) { r = 1; u = 0.75; v = -0.5; } else { r = 3; u = -0.75; v = -0.5; } }
          else { if (y > 0) { r = 2; u = 0.75; v = 0.5; } else { r = 1; u = -0.75; v = 0.5; } }
          p = 1.0;
          break;
        case 2:  // implosion (:354-390)
          if (std::fabs(x) + std::fabs(y) < 0.15) { r = 0.125; p = 0.14; } else { r = 1.0; p = 1.0; }
          break;
        case 5:  // 2D Riemann another (:460-507)
          if (x > 0.0) { if (y > 0.0) { r = 0.5313; p = 0.4; } else { r = 1; v = 0.7276; p = 1; } }
          else { if (y > 0) { r = 1; u = 0.7276; p = 1; } else { r = 0.8; p = 1; } }
          break;
        case 6:  // KHI (:509-541)
          if (std::fabs(y) <= 0.25) { r = 2.0; u = -0.5; } else { r = 1.0; u = 0.5; }
          v = 0.01 * std::sin(2 * M_PI * x);
          p = 2.5;
          break;
        case 10: {  // isentropic vortex on a periodic box (not a reference preset)
          const double L = info.calZone[1] - info.calZone[0], eps = 5.0;
          double xs = std::fmod(x - info.calZone[0] - L / 2 + 10 * L, L) - L / 2;
          double ys = std::fmod(y - info.calZone[0] - L / 2 + 10 * L, L) - L / 2;
          double f = std::exp(0.5 * (1.0 - (xs * xs + ys * ys)));
          u = 1.0 - eps / (2 * M_PI) * f * ys;
          v = 1.0 + eps / (2 * M_PI) * f * xs;
          double T = 1.0 - (G - 1) * eps * eps / (8 * G * M_PI * M_PI) * f * f;
          r = std::pow(T, 1.0 / (G - 1));
          p = std::pow(r, G);
          break;
        }
        case 3:  // Rayleigh-Taylor (:396-433): heavy fluid below y = 0.5
          if (y <= 0.5) {
            r = 2.0;
            p = 2.0 * y + 1.0;
          } else {
            r = 1.0;
            p = y + 3.0 / 2.0;
          }
          v = -0.025 * std::sqrt(G * p / r) * std::cos(8.0 * M_PI * x);
          break;
        case 4:  // double Mach reflection (:436-468): oblique Mach-10 shock from x = 1/6
          if (y >= std::sqrt(3) * (x - 1.0 / 6.0)) {
            r = 8.0;
            u = 8.25 * std::cos(M_PI / 6);
            v = -8.25 * std::sin(M_PI / 6);
            p = 116.5;
          } else {
            r = 1.4;
            p = 1.0;
          }
          break;
        default: return "2D case " + std::to_string(info.nCase) + " is not defined";
        }
        put(cons, {r, r * u, r * v, 1.0 / (G - 1) * p + r * (u * u + v * v) / 2});
      }
    if (info.nCase == 2) {
      set_bc(cfg, 0, CFD_BC_SYMMETRYX, CFD_BC_SYMMETRYX);
      set_bc(cfg, 1, CFD_BC_SYMMETRYY, CFD_BC_SYMMETRYY);
    } else if (info.nCase == 3) {
      // initBnds (:796-799, 831-862): symmetric sides, Dirichlet bottom and top
      set_bc(cfg, 0, CFD_BC_SYMMETRYX, CFD_BC_SYMMETRYX);
      set_bc(cfg, 1, CFD_BC_DIRICHLET, CFD_BC_DIRICHLET);
      const double lo[4] = {2.0, 0, 0, 1.0}, hi[4] = {1.0, 0, 0, 2.5};
      for (int v = 0; v < 4; v++) {
        cfg.bc_prim[2][v] = lo[v];
        cfg.bc_prim[3][v] = hi[v];
      }
      // SourceTerm GRAVITY, g = (0, 1); the reference leaves it switched off
      // (zoneMain.cpp:32 is commented out) -- driver flag --gravity
      if (info.gravity) cfg.gravity[1] = 1.0;
    } else if (info.nCase == 4) {
      // initDoubleMachBnds (:880-934): post-shock Dirichlet inflow, outflow,
      // the wall (Dirichlet for x < 1/6, reflecting beyond), DoubleMachUp on top
      const double post[4] = {8.0, 8.25 * std::cos(M_PI / 6), -8.25 * std::sin(M_PI / 6), 116.5};
      set_bc(cfg, 0, CFD_BC_DIRICHLET, CFD_BC_OUTFLOW);
      set_bc(cfg, 1, CFD_BC_DMR_WALL, CFD_BC_DMR_TOP);
      for (int v = 0; v < 4; v++) {
        cfg.bc_prim[0][v] = post[v];
        cfg.bc_prim[2][v] = post[v];
        cfg.bc_prim[3][v] = post[v];
      }
    } else if (info.nCase == 6) {
      // PERIODIC1D as the reference wires it for 2D Euler (a mirror copy), or a true periodic halo
      int b = info.truePeriodic ? CFD_BC_PERIODIC : CFD_BC_MIRROR_COPY;
      set_bc(cfg, 0, b, b);
      set_bc(cfg, 1, b, b);
    } else if (info.nCase == 10) {
      set_bc(cfg, 0, CFD_BC_PERIODIC, CFD_BC_PERIODIC);
      set_bc(cfg, 1, CFD_BC_PERIODIC, CFD_BC_PERIODIC);
    }
    return "";
  }

  if (info.dim == 3) {
    if (info.nCase != 0) return "3D case " + std::to_string(info.nCase) + " is not defined";
    const int n = ic[0];
    if (ic[1] != n || ic[2] != n) return "3D case 0 needs a cubic grid";
    uint64_t seed = 20261203ull;
    const double amp = 0.05, mach = 0.3;
    auto fr = random_phase(n, 8, seed), fu = random_phase(n, 8, seed), fv = random_phase(n, 8, seed),
         fw = random_phase(n, 8, seed);
    const double p = 1.0 / (G * mach * mach);
    cons.reserve((size_t)n * n * n * 5);
    for (size_t i = 0; i < (size_t)n * n * n; i++) {
      double r = 1.0 + amp * fr[i], u = amp * fu[i], v = amp * fv[i], w = amp * fw[i];
      put(cons, {r, r * u, r * v, r * w, p / (G - 1) + 0.5 * r * (u * u + v * v + w * w)});
    }
    for (int d = 0; d < 3; d++) set_bc(cfg, d, CFD_BC_PERIODIC, CFD_BC_PERIODIC);
    return "";
  }
  return "dim must be 1, 2 or 3";
}

}  // namespace cfdhost
