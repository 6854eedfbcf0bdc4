// Launch planning of the sweep kernels: which tile shape and segment length
// to use along each direction.  The default is an occupancy/lane-efficiency
// model; cfd_autotune refines it by timing the candidates on the device.
#pragma once
#include "sweep.hpp"

namespace cfd {

struct PlanCandidate {
  int tile;
  int seg;
  double cost;  // modelled lane slots (lower is better)
};

inline long plan_npencil(const int* n, int ax) {
  return ax == 0 ? (long)n[1] * n[2] : (ax == 1 ? (long)n[0] * n[2] : (long)n[0] * n[1]);
}

inline long plan_pencil_blocks(const int* n, int ax, int tile) {
  const int cp = tile_cp(tile);
  return (plan_npencil(n, ax) + cp - 1) / cp;
}

// Lane slots spent by the interface phase (the dominant one): every segment of
// length L runs ceil((L + 3) / CA) chunks of CA x CP lanes.  The kernel is
// latency bound, so a launch with fewer workgroups than the chip holds at
// once (256 CUs x 4 resident workgroups) is charged in proportion to the
// missing ones.
inline double plan_cost(const int* n, int ax, int tile, int seg) {
  const int ca = tile_ca(tile);
  const long nb = plan_pencil_blocks(n, ax, tile);
  const int len = n[ax];
  long chunks = 0;
  for (int c0 = 0; c0 < len; c0 += seg) {
    int l = len - c0 < seg ? len - c0 : seg;
    chunks += (l + 3 + ca - 1) / ca;
  }
  const long nseg = (len + seg - 1) / seg;
  const double blocks = (double)nb * nseg;
  double cost = (double)nb * chunks * 256.0;
  const double want = 1024.0;
  if (blocks < want) cost *= want / blocks;
  return cost;
}

// Candidate (tile, segment) pairs, cheapest first.
inline int plan_candidates(const int* n, int ax, PlanCandidate* out, int maxc) {
  const int tiles[2] = {ax == 0 ? 0 : 1, ax == 0 ? 1 : 2};
  PlanCandidate all[256];
  int na = 0;
  const int len = n[ax];
  for (int t = 0; t < 2; t++) {
    const int ca = tile_ca(tiles[t]);
    // whole pencil, and lengths that fill the last chunk exactly
    if (na < 256) all[na++] = {tiles[t], len, plan_cost(n, ax, tiles[t], len)};
    for (int k = 1; k * ca - 3 < len && na < 256; k++) {
      int seg = k * ca - 3;
      if (seg < 8) continue;
      all[na++] = {tiles[t], seg, plan_cost(n, ax, tiles[t], seg)};
    }
  }
  // selection sort by cost (small lists)
  for (int i = 0; i < na; i++)
    for (int j = i + 1; j < na; j++)
      if (all[j].cost < all[i].cost || (all[j].cost == all[i].cost && all[j].seg > all[i].seg)) {
        PlanCandidate t = all[i];
        all[i] = all[j];
        all[j] = t;
      }
  int m = na < maxc ? na : maxc;
  for (int i = 0; i < m; i++) out[i] = all[i];
  return m;
}

inline void plan_default(const int* n, int ax, int* tile, int* seg) {
  PlanCandidate c[1];
  if (plan_candidates(n, ax, c, 1) == 1) {
    *tile = c[0].tile;
    *seg = c[0].seg;
  }
}

}  // namespace cfd
