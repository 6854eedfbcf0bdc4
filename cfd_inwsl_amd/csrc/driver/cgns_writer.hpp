// CGNS solution files written without libcgns / libhdf5.
//
// The reference writes each output with libcgns (zone_FD/src/src/cgnsio.cpp:3-101):
// base "Base" (cell and physical dimension = dim), structured zone "Zone  1"
// with vertex sizes iMax, cell sizes iMax-1 and zero boundary sizes, vertex
// coordinates CoordinateX/Y/Z (RealDouble), and a cell-centred FlowSolution
// "flowSolution" holding Density, XVelocity, [YVelocity, [ZVelocity,]] Pressure
// (RealDouble).  This writer produces the same CGNS tree in the HDF5 form
// libcgns uses (every node an HDF5 group with name/label/type/flags
// attributes, data in a " data" dataset), as a self-contained HDF5 file:
// version-2 superblock, version-2 object headers with compact link storage and
// link creation order, compact or contiguous datasets, lookup3 checksums.
#pragma once
#include <array>
#include <string>
#include <vector>

namespace cfdhost {

struct CgnsSolution {
  int dim = 1;
  std::array<int, 3> vertices{1, 1, 1};   // iMax per direction
  std::array<double, 3> lo{0, 0, 0};      // first vertex coordinate per direction
  double h = 1;                           // vertex spacing (one for every direction)
  // cell-centred fields, x fastest: Density, XVelocity, [YVelocity, [ZVelocity,]] Pressure
  std::vector<std::vector<double>> fields;
};

// Writes the file; returns an empty string on success, else the reason.
std::string write_cgns(const std::string& path, const CgnsSolution& sol);

// Bob Jenkins' lookup3 hashlittle (HDF5's metadata checksum); exposed for tests.
unsigned lookup3(const unsigned char* data, size_t n, unsigned initval = 0);

}  // namespace cfdhost
