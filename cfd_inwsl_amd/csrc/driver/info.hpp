// Host-side run configuration of the CongFD driver: the fields of the
// reference's Info (zone_FD/src/include/info.hpp:4-66) used on the RHS/RK3
// path, the 16-line info.txt reader (zone_FD/zoneMain.cpp:254-306) and the
// output naming of Info::filename (zone_FD/src/src/info.cpp:229-257).
#pragma once
#include <array>
#include <string>
#include <vector>

#include "../../../include/congfd.h"

namespace cfdhost {

struct Info {
  // reference defaults (info.hpp:9-42; zoneMain.cpp:24-29 sets HLLC / MND6)
  int interMethod = 2;  // WCNS5
  int fluxMethod = CFD_FLUX_HLLC;
  int diffMethod = CFD_DIFF_MND6;
  int recon = CFD_RECON_CHAR;
  int nCase = 2;
  double t = 0;
  int step = 0;
  int endStep = 25000;
  int dim = 1;
  double outputDt = 0.01;
  double outputT = 0;
  double CFL = 0.5;
  std::array<int, 3> iMax{201, 201, 2};               // vertices per direction
  std::array<double, 6> calZone{0, 0.3, 0, 0.3, 0, 2};
  int threads = 1;                                    // info.txt line 16 (OpenMP threads: unused on GPU)
  bool truePeriodic = false;                          // use a true periodic halo where the reference mirrors
  int device = 0;
  int precision = CFD_PRECISION_FP64;
  bool gravity = false;                               // RTI: SourceTerm GRAVITY (off in the reference)

  std::array<int, 3> icMax() const;   // cells per direction (info.cpp:115-128)
  double geth() const;                // one spacing for every direction (info.cpp:310-314, initializer.cpp:594-597)
  std::string caseName() const;
  std::string schemeName() const;
  std::string gridInfo() const;
  std::string filename() const;       // "<case> - <scheme> - <grid> - t=<t>.<ext>"
};

// Reads the reference's 16-line info.txt; returns false if the file is absent.
bool readInfoTxt(const std::string& path, Info& info);

// Conservative initial state (AoS [cell][nv], x fastest) and boundary wiring
// of case info.nCase (Initializer::solInit / initBnds, zone_FD/src/src/initializer.cpp).
// Returns an empty string on success, else the reason the case is unsupported.
std::string setupCase(const Info& info, std::vector<double>& cons, cfd_config& cfg);

}  // namespace cfdhost
