#include "info.hpp"

#include <cstdio>
#include <fstream>
#include <map>

namespace cfdhost {

std::array<int, 3> Info::icMax() const {
  std::array<int, 3> r;
  for (int i = 0; i < 3; i++) r[i] = iMax[i] < 2 ? 1 : iMax[i] - 1;
  return r;
}

double Info::geth() const { return (calZone[1] - calZone[0]) / (iMax[0] - 1); }

// exampleStr1D / exampleStr2D (zone_FD/src/src/info.cpp:201-217); 3D names are new.
std::string Info::caseName() const {
  static const std::map<int, std::string> d1 = {{0, "Sod"}, {1, "ShuOsher"}, {2, "Lax"}, {3, "sedov"},
                                                {4, "Woodward_Colella"}, {5, "Double_sparse_wave"}};
  static const std::map<int, std::string> d2 = {{0, "2D_Riemann"}, {1, "2D_Riemann_vortex"}, {2, "implosion"},
                                                {3, "RTI"}, {4, "Double_Mach"}, {5, "2D_Riemann_another"},
                                                {6, "KHI"}, {10, "isentropic_vortex"}};
  static const std::map<int, std::string> d3 = {{0, "periodic_random_phase"}};
  const auto& m = dim == 1 ? d1 : (dim == 2 ? d2 : d3);
  auto it = m.find(nCase);
  return it == m.end() ? std::string("unknown") : it->second;
}

// disStr (zone_FD/src/src/info.cpp:155-198): the name each InterMethod value
// gives the output files.
std::string Info::schemeName() const {
  static const std::map<int, std::string> s = {
      {0, "FIRSTORDER"}, {1, "MUSCL"}, {2, "WCNS5"}, {3, "WCNSZ5"}, {4, "WCNS5Char"}, {5, "WCNSZ5Char"},
      {6, "WCNS5CONG"}, {7, "TCNSCongA"}, {8, "WCNS5CONGZ"}, {9, "WCNS5CONGZCT4"}, {10, "WCNS5CONGZCT7"},
      {11, "TCNS5"}, {12, "TCNS5CT4"}, {13, "TCNS5CT7"}, {14, "LINEAR5"}, {15, "MUCSLIN5"}, {16, "INTERMAX"},
      {17, "whfTCNSN"}, {18, "whfTCNSNA"}, {19, "whfTCNSNAS"}, {20, "whfTCNSNS"}, {21, "whfTCNSNLAD"},
      {22, "congTCNS5CT5"}, {23, "congTCNS5CT10"}, {24, "whfAITCNSNS"}, {25, "whfAITCNSNA"},
      {26, "whfAITCNSNAS_1"}, {27, "whfAITCNSNAS_2"}, {28, "whfAITCNSNLADS"}, {29, "whfAITCNSNAZS"},
      {30, "whfAITCNSNmyASF002_1"}, {31, "whfAITCNSNmyASF002_2"}, {32, "whfAITCNSNmyASF002_ai1"},
      {33, "whfzycTCNSNmyASF002_1"}, {34, "whfCOMPARE"}, {35, "TENO"}, {36, "TENO-A"}, {37, "TENO-S"},
      {38, "TENO-AS-myF203_NoS"}, {39, "TENO-AS-myF202_NoS"}};
  auto it = s.find(interMethod);
  return it == s.end() ? std::to_string(interMethod) : it->second;
}

std::string Info::gridInfo() const {
  if (dim == 1) return std::to_string(iMax[0]);
  if (dim == 2) return std::to_string(iMax[0]) + "x" + std::to_string(iMax[1]);
  return std::to_string(iMax[0]) + "x" + std::to_string(iMax[1]) + "x" + std::to_string(iMax[2]);
}

std::string Info::filename() const {
  char tb[64];
  std::snprintf(tb, sizeof(tb), "t=%.4f", t);
  return caseName() + " - " + schemeName() + " - " + gridInfo() + " - " + tb;
}

bool readInfoTxt(const std::string& path, Info& info) {
  std::ifstream f(path);
  if (!f.is_open()) return false;
  int n;
  double x;
  f >> n;
  if (n < 16) info.interMethod = n;  // zoneMain.cpp:261-263: only values below INTERMAX
  f >> n;
  info.endStep = n;
  f >> x;
  info.outputDt = x;
  f >> x;
  info.CFL = x;
  f >> n;
  info.nCase = n;
  for (int i = 0; i < 6; i++) f >> info.calZone[i];
  for (int i = 0; i < 3; i++) f >> info.iMax[i];
  f >> info.dim;
  f >> info.threads;
  return true;
}

}  // namespace cfdhost
