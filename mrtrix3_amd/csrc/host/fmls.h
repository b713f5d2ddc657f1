// Host-side set-up of FOD segmentation (fod2fixel's FMLS, src/dwi/fmls.cpp): the direction set,
// its adjacency, the integration weights, the nearest-direction lookup grid and the precomputed
// Legendre table the peak search interpolates.  Everything here is computed once per process
// and uploaded to the GPU; nothing is on the per-voxel path.
#pragma once

#include <vector>

namespace mrcsd {
namespace fmls {

constexpr int kMaxAdj = 8;   // neighbours per direction, -1 padded (5 or 6 on the 1281 set)

struct DirectionSet {
  int n = 0;
  std::vector<double> unit;      // [n][3]
  std::vector<int> adj;          // [n][kMaxAdj], ascending, -1 padded
  std::vector<double> weights;   // [n] integration weights
  // FastLookupSet grid (directions/set.cpp:369-467)
  int n_az = 0, n_el = 0;
  double az_step = 0, el_step = 0, az_begin = 0, el_begin = 0;
  std::vector<int> cell_start;   // [n_az n_el + 1]
  std::vector<int> cell_dirs;    // candidate directions of each cell, ascending
};

// The 1281-direction tessellation fod2fixel segments on (cmd/fod2fixel.cpp:318), with all of the
// above (built on first use, then shared).
const DirectionSet& tessellation_1281();

// Builds the tables for directions given as [azimuth elevation] rows.
DirectionSet make_direction_set(const std::vector<double>& az_el);

// Neighbours: the edges of the convex hull of the directions and their antipodes
// (Set::initialise_adjacency, directions/set.cpp:70-313).
std::vector<int> hull_adjacency(const std::vector<double>& unit, int n);

// IntegrationWeights (fmls.cpp:109-141): least-squares weights that integrate every even SH
// basis function up to LforN(n) + 2 exactly where they can (the constant to 4 pi).
std::vector<double> integration_weights(const std::vector<double>& unit, int n);

// Nearest direction through the lookup grid (FastLookupSet::select_direction, set.cpp:336-355).
int select_direction(const DirectionSet& s, const double* p);

// PrecomputedAL<double> (core/math/SH.h:418-434): [ndir][(lmax/2+1)^2], row k at elevation
// k pi / (ndir - 1).
std::vector<double> legendre_table(int lmax, int ndir);

}  // namespace fmls
}  // namespace mrcsd
