// lattice.h -- materials, unit cell and supercell of a JAMS config
// (core/lattice.cc, containers/material.h of the reference).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.h"
#include "config.h"

namespace jams {

struct Material {
  std::string name;
  double moment = 0.0;   // meV/T (input in Bohr magnetons)
  double gyro = 0.0;     // rad/(ps T) (input as a fraction of gamma_e)
  double alpha = 0.01;
  Vec3 spin = {0.0, 0.0, 1.0};
  bool randomize = false;
};

struct BasisSite {
  int material = 0;
  Vec3 frac = {0, 0, 0};
};

class Lattice {
 public:
  // Reads materials, unitcell and lattice groups (core/lattice.cc:221-255)
  // and generates the supercell (:577-749).  `seed` drives spin = "random".
  void init_from_config(const Config &cfg, bool gilbert_prefactor, uint64_t seed);

  int64_t num_sites() const { return (int64_t)site_material_.size(); }
  int num_basis_sites() const { return (int)basis_.size(); }
  int num_materials() const { return (int)materials_.size(); }
  const Material &material(int i) const { return materials_[i]; }
  int material_index(const std::string &name) const;  // -1 if absent
  const BasisSite &basis_site(int m) const { return basis_[m]; }
  int size(int d) const { return size_[d]; }
  bool periodic(int d) const { return periodic_[d]; }
  double parameter() const { return parameter_; }
  const Mat3 &cell() const { return cell_; }             // columns are a1, a2, a3
  const Mat3 &inverse_cell() const { return inv_cell_; }
  Vec3 fractional_to_cartesian(const Vec3 &f) const { return cell_ * f; }
  Vec3 cartesian_to_fractional(const Vec3 &c) const { return inv_cell_ * c; }

  // ((i*Ly + j)*Lz + k)*nb + m  (core/lattice.cc:622-649)
  int64_t site_index(int i, int j, int k, int m) const {
    return (((int64_t)i * size_[1] + j) * size_[2] + k) * (int64_t)basis_.size() + m;
  }
  // core/lattice.cc:987-1007: wrap periodic, reject outside open boundaries
  bool apply_boundary_conditions(int &a, int &b, int &c) const;
  int site_material(int64_t i) const { return site_material_[i]; }
  int site_basis(int64_t i) const { return (int)(i % (int64_t)basis_.size()); }
  Vec3 site_position(int64_t i) const;  // cartesian, lattice constants

  // per-spin arrays of globals:: (core/lattice.cc:688-729)
  std::vector<double> mus, gyro, alpha, spins;

  // crystal symmetry (replaces spglib): rotations in fractional coordinates
  // of the zero-translation space-group operations
  bool symops_enabled() const { return symops_; }
  const std::vector<Mat3> &site_point_group(int m) const { return point_groups_[m]; }
  std::vector<Vec3> generate_symmetric_points(int m, const Vec3 &r_cart, double tol) const;

 private:
  std::vector<Material> materials_;
  std::vector<BasisSite> basis_;
  Mat3 cell_{}, inv_cell_{};
  double parameter_ = 0.0;
  Vec3i size_{1, 1, 1};
  bool periodic_[3] = {true, true, true};
  bool symops_ = true;
  std::vector<int> site_material_;
  std::vector<std::vector<Mat3>> point_groups_;

  void read_materials(const Setting &s);
  void read_unitcell(const Setting &s);
  void find_symmetry();
};

// Spins from a TSV file of N x 3 doubles (helpers/load.h:22-48), or from
// the dataset /spins of an HDF5 file named *.h5 (helpers/load.h:13-19, :53-57).
std::vector<double> load_spins_tsv(const std::string &path, int64_t n);

// Marsaglia point on the unit sphere (helpers/random.h:12-26) from a seeded
// counter-based stream: the reference draws from std::random_device, which
// cannot be reproduced, so "random" spins here are a function of sim.seed.
void random_sphere_spins(uint64_t seed, int64_t n, double *out);

}  // namespace jams
