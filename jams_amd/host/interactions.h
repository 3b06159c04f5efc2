// interactions.h -- exchange interaction templates and the per-site
// neighbour list (core/interactions.cc, containers/interaction_list.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.h"
#include "config.h"
#include "lattice.h"

namespace jams {

struct InteractionData {
  int basis_site_i = -1, basis_site_j = -1;
  std::string type_i, type_j;
  Vec3 r = {0, 0, 0};       // interaction vector, cartesian, lattice constants
  Mat3 J{};                 // tensor in input energy units
  Vec3i T = {0, 0, 0};      // lattice translation vector
};

enum class CoordinateFormat { kCartesian, kFractional };

// Neighbour list sorted by (i, j); pair p carries values[v[p]]
// (jams::InteractionList<Mat3,2>).
struct NeighbourList {
  std::vector<int32_t> i, j, v;
  std::vector<Mat3> values;
  int64_t size() const { return (int64_t)i.size(); }
  int64_t num_interactions(int64_t site) const;
};

struct NeighbourOptions {
  CoordinateFormat coords = CoordinateFormat::kCartesian;
  bool use_symops = true;
  double energy_cutoff = 0.0;
  double radius_cutoff = 100.0;
  double distance_tolerance = kLatticeTolerance;
  bool check_no_zero = true;
  bool check_identical_count = true;
  bool check_identical_total = true;
};

// Template rows from an `interactions` list (core/interactions.cc:254-289)
// or an exc_file (:207-251).
std::vector<InteractionData> interactions_from_setting(const Setting &s, bool &kkr);
std::vector<InteractionData> interactions_from_file(const std::string &path, bool &kkr);

// post_process_interactions (:292-347) + neighbour_list_from_interactions
// (:350-395) + neighbour_list_checks (:446-510).
NeighbourList generate_neighbour_list(const Lattice &lat, std::vector<InteractionData> rows,
                                      bool kkr, const NeighbourOptions &opt);

}  // namespace jams
