// exchange_pairs.h -- distance-based exchange pair producers: the
// exchange-neartree and exchange-functional modules
// (hamiltonian/exchange_neartree.cc, hamiltonian/exchange_functional.cc).
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "common.h"
#include "config.h"
#include "interactions.h"
#include "lattice.h"

namespace jams {

// one inserted pair: row site i, column site j, tensor J (energy units of the
// engine)
struct PairValue {
  int32_t i, j;
  Mat3 J;
};

// Candidate pairs (i, image of j) of the periodic supercell within a search
// radius: a uniform cell list over the sites and the periodic images that can
// lie within the radius of the supercell.
class PairSearch {
 public:
  struct Candidate {
    int32_t j;
    Vec3 rij;   // r_j(image) - r_i
    double d;   // norm(r_i - r_j(image))
  };
  PairSearch(const Lattice &lat, double r_search);
  // every image within one bin (>= r_search) of site i, in a fixed order;
  // callers apply their own distance test
  void candidates(int64_t i, std::vector<Candidate> &out) const;

 private:
  struct Image {
    Vec3 p;
    int32_t j;
  };
  const Lattice &lat_;
  std::vector<Vec3> pos_;
  std::vector<Image> img_;
  std::vector<int64_t> order_, start_;
  double bin_ = 1.0;
  double org_[3] = {0, 0, 0};
  int64_t nb_[3] = {1, 1, 1};
  int64_t bin_coord(double x, int d) const;
  int64_t bin_of(const Vec3 &p) const;
};

// the supercell edges L_d a_d (lattice constants)
std::array<Vec3, 3> supercell_vectors(const Lattice &lat);
// Lattice::max_interaction_radius (core/lattice.cc:1009-1011, 1155-1201)
double max_interaction_radius(const Lattice &lat);
// distance_units of a Hamiltonian setting, as the factor to lattice constants
double distance_unit(const Setting &s, const Lattice &lat);

using Functional = std::function<double(const Vec3 &)>;
// ExchangeFunctionalHamiltonian::functional_from_params
Functional make_functional(const std::string &name, const std::vector<double> &params,
                           double energy_conv, double distance_conv);

// The pairs each module inserts (before finalize); counter is the reference's
// "total interactions".
std::vector<PairValue> neartree_pairs(const Setting &s, const Lattice &lat, double energy_conv,
                                      int64_t &counter);
std::vector<PairValue> functional_pairs(const Setting &s, const Lattice &lat, double energy_conv,
                                        int64_t &counter);

// symmetry_check = "force_symmetric"
void force_symmetric(std::vector<PairValue> &pairs);
// sorted (i, j) neighbour list with de-duplicated tensors
NeighbourList pairs_to_list(std::vector<PairValue> pairs);

}  // namespace jams
