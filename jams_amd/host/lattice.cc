// lattice.cc -- see lattice.h.
#include "lattice.h"

#include "h5.h"

#include <algorithm>
#include <cctype>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

namespace jams {

std::string lowercase(std::string s) {
  for (auto &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

std::string capitalize(std::string s) {
  for (auto &c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

namespace {

// core/lattice.cc:50-64: one +1 shift for negative values, 1 -> 0
Vec3 normalise_fractional(Vec3 r, double eps) {
  for (int n = 0; n < 3; ++n) {
    if (r[n] < 0.0) r[n] = r[n] + 1.0;
    if (approximately_equal(r[n], 1.0, eps)) r[n] = 0.0;
  }
  return r;
}

bool fractional_is_normalised(const Vec3 &r, double eps) {
  for (int n = 0; n < 3; ++n)
    if (r[n] < 0.0 || r[n] > 1.0 || approximately_equal(r[n], 1.0, eps)) return false;
  return true;
}

Vec3 read_vec3(const Setting &s) {
  if (!s.is_aggregate() || s.length() != 3) throw ConfigError("expected a 3-vector at " + s.path());
  return {s[0].as_double(), s[1].as_double(), s[2].as_double()};
}

// Philox4x32-10 (Salmon et al. 2011) for set-up randomness on the host.
void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

double setup_uniform(uint64_t seed, uint64_t stream, uint64_t index) {
  uint32_t c[4] = {(uint32_t)index, (uint32_t)(index >> 32), (uint32_t)stream, 0xFFFFFFFFu};
  philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t k = ((uint64_t)c[0] << 21) | (uint64_t)(c[1] >> 11);
  return (double)k * 0x1.0p-53;
}

}  // namespace

void random_sphere_spins(uint64_t seed, int64_t n, double *out) {
  uint64_t idx = 0;
  for (int64_t i = 0; i < n; ++i) {
    double v1, v2, q;
    do {
      v1 = std::fma(2.0, setup_uniform(seed, 0, idx++), -1.0);
      v2 = std::fma(2.0, setup_uniform(seed, 0, idx++), -1.0);
      q = (v1 * v1) + (v2 * v2);
    } while (q > 1.0);
    const double ss = 2.0 * std::sqrt(1.0 - q);
    out[3 * i + 0] = v1 * ss;
    out[3 * i + 1] = v2 * ss;
    out[3 * i + 2] = std::fma(-2.0, q, 1.0);
  }
}

std::vector<double> load_spins_tsv(const std::string &path, int64_t n) {
  const std::string ext = path.size() > 3 ? path.substr(path.size() - 3) : "";
  if (ext == ".h5") {
    // helpers/load.h:13-19, :53-57: the dataset /spins of an HDF5 file
    std::vector<uint64_t> dims;
    std::vector<double> s = h5::read_doubles(path, "/spins", &dims);
    if ((int64_t)s.size() != 3 * n)
      throw RuntimeError("loading array from file: '" + path + "' expected size: " +
                         std::to_string(3 * n) + " actual size: " + std::to_string(s.size()));
    return s;
  }
  std::ifstream f(path);
  if (!f.is_open()) throw RuntimeError("failed to open file: " + path);
  std::vector<double> s;
  s.reserve((size_t)(3 * n));
  for (std::string line; std::getline(f, line);) {
    size_t p = line.find_first_not_of(" \t\r");
    if (p == std::string::npos || line[p] == '#') continue;
    std::istringstream is(line);
    double v;
    while (is >> v) s.push_back(v);
  }
  if ((int64_t)s.size() != 3 * n)
    throw RuntimeError("loading array from file: '" + path + "' expected size: " +
                       std::to_string(3 * n) + " actual size: " + std::to_string(s.size()));
  return s;
}

int Lattice::material_index(const std::string &name) const {
  for (int i = 0; i < (int)materials_.size(); ++i)
    if (materials_[i].name == name) return i;
  return -1;
}

void Lattice::read_materials(const Setting &s) {
  // containers/material.h:31-55
  for (int i = 0; i < s.length(); ++i) {
    const Setting &m = s[i];
    Material mat;
    mat.name = m["name"].as_string();
    if (material_index(mat.name) >= 0)
      throw RuntimeError("the material " + mat.name + " is specified twice in the configuration");
    mat.moment = m["moment"].as_double() * kBohrMagnetonIU;
    mat.gyro = m.get<double>("gyro", 1.0) * kGyromagneticRatioIU;
    mat.alpha = m.get<double>("alpha", 0.01);
    if (m.exists("spin")) {
      const Setting &sp = m["spin"];
      if (sp.is_string()) {
        if (lowercase(sp.as_string()) == "random") mat.randomize = true;
        else throw ConfigError("spin setting string must be \"random\"");
      } else if (sp.is_aggregate() && sp.length() == 3) {
        mat.spin = read_vec3(sp);
      } else if (sp.is_aggregate() && sp.length() == 2) {
        const double th = sp[0].as_double() * kPi / 180.0, ph = sp[1].as_double() * kPi / 180.0;
        mat.spin = {std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)};
      } else {
        throw ConfigError("spin setting array is not length 2 or 3");
      }
    }
    materials_.push_back(mat);
  }
}

void Lattice::read_unitcell(const Setting &s) {
  // core/lattice.cc:356-392: basis rows form the unit-cell matrix whose
  // columns are the lattice vectors
  const Setting &b = s["basis"];
  if (b.length() != 3) throw ConfigError("unitcell.basis must have 3 rows");
  for (int r = 0; r < 3; ++r) {
    Vec3 row = read_vec3(b[r]);
    for (int c = 0; c < 3; ++c) cell_[r][c] = row[c];
  }
  inv_cell_ = inverse(cell_);
  parameter_ = s["parameter"].as_double();
  if (parameter_ <= 0.0) throw RuntimeError("lattice parameter must be positive");
  const std::string fmt = capitalize(s.get<std::string>("coordinate_format", "FRACTIONAL"));
  if (fmt != "FRACTIONAL" && fmt != "CARTESIAN")
    throw RuntimeError("Unknown coordinate format for atom positions in unit cell");
  const Setting &pos = s["positions"];
  if (!pos.is_list()) throw RuntimeError("unitcell.positions from a file is not supported by this build");
  for (int i = 0; i < pos.length(); ++i) {
    BasisSite site;
    const std::string mname = pos[i][0].as_string();
    site.material = material_index(mname);
    if (site.material < 0)
      throw RuntimeError("material " + mname + " in the motif is not defined in the configuration");
    site.frac = read_vec3(pos[i][1]);
    if (fmt == "CARTESIAN") site.frac = cartesian_to_fractional(site.frac);
    site.frac = normalise_fractional(site.frac, kLatticeTolerance);
    if (!fractional_is_normalised(site.frac, kLatticeTolerance))
      throw RuntimeError("atom position " + std::to_string(i) + " is not a valid fractional coordinate");
    basis_.push_back(site);
  }
  if (basis_.empty()) throw RuntimeError("unitcell has no positions");
}

// Space-group operations without translation: integer rotations W (entries
// in {-1, 0, 1}, |det| = 1) that preserve the metric M^T M and map every
// basis site onto a site of the same material modulo a lattice vector.  The
// reference asks spglib for all operations and keeps those whose translation
// is zero (core/lattice.cc:1127-1153); this enumerates exactly that subset.
void Lattice::find_symmetry() {
  const Mat3 G = transpose(cell_) * cell_;
  double gmax = 0.0;
  for (auto &r : G)
    for (double v : r) gmax = std::max(gmax, std::abs(v));
  std::vector<Mat3> ops;
  int w[9];
  for (int code = 0; code < 19683; ++code) {
    int x = code;
    for (int q = 0; q < 9; ++q) {
      w[q] = x % 3 - 1;
      x /= 3;
    }
    Mat3 W{};
    for (int q = 0; q < 9; ++q) W[q / 3][q % 3] = w[q];
    const double det = W[0][0] * (W[1][1] * W[2][2] - W[1][2] * W[2][1]) -
                       W[0][1] * (W[1][0] * W[2][2] - W[1][2] * W[2][0]) +
                       W[0][2] * (W[1][0] * W[2][1] - W[1][1] * W[2][0]);
    if (std::abs(det) != 1.0) continue;
    const Mat3 GW = transpose(W) * G * W;
    bool metric = true;
    for (int i = 0; i < 3 && metric; ++i)
      for (int j = 0; j < 3 && metric; ++j)
        if (std::abs(GW[i][j] - G[i][j]) > kLatticeTolerance * gmax) metric = false;
    if (!metric) continue;
    bool maps = true;
    for (const auto &a : basis_) {
      const Vec3 p = W * a.frac;
      bool found = false;
      for (const auto &b : basis_) {
        if (b.material != a.material) continue;
        Vec3 d = p - b.frac;
        for (auto &v : d) v -= std::nearbyint(v);
        if (norm(cell_ * d) <= kLatticeTolerance) {
          found = true;
          break;
        }
      }
      if (!found) {
        maps = false;
        break;
      }
    }
    if (maps) ops.push_back(W);
  }
  point_groups_.assign(basis_.size(), {});
  for (size_t m = 0; m < basis_.size(); ++m) {
    const Vec3 p = basis_[m].frac;
    for (const auto &W : ops) {
      const Vec3 q = normalise_fractional(W * p, kLatticeTolerance);
      if (approximately_equal(p, q, kLatticeTolerance)) point_groups_[m].push_back(W);
    }
  }
}

std::vector<Vec3> Lattice::generate_symmetric_points(int m, const Vec3 &r_cart, double tol) const {
  // core/lattice.cc:1015-1035
  const Vec3 r_frac = cartesian_to_fractional(r_cart);
  std::vector<Vec3> pts{r_cart};
  for (const auto &W : point_groups_[m]) {
    const Vec3 r = fractional_to_cartesian(W * r_frac);
    bool exists = false;
    for (const auto &q : pts)
      if (approximately_equal(q, r, tol)) {
        exists = true;
        break;
      }
    if (!exists) pts.push_back(r);
  }
  return pts;
}

bool Lattice::apply_boundary_conditions(int &a, int &b, int &c) const {
  int *v[3] = {&a, &b, &c};
  for (int d = 0; d < 3; ++d) {
    if (!periodic_[d] && (*v[d] < 0 || *v[d] >= size_[d])) return false;
    *v[d] = (*v[d] + size_[d]) % size_[d];
  }
  return true;
}

Vec3 Lattice::site_position(int64_t i) const {
  const int64_t nb = (int64_t)basis_.size();
  const int m = (int)(i % nb);
  int64_t cell = i / nb;
  const int k = (int)(cell % size_[2]);
  cell /= size_[2];
  const int j = (int)(cell % size_[1]);
  const int ii = (int)(cell / size_[1]);
  const Vec3 f = {basis_[m].frac[0] + ii, basis_[m].frac[1] + j, basis_[m].frac[2] + k};
  return cell_ * f;
}

void Lattice::init_from_config(const Config &cfg, bool gilbert_prefactor, uint64_t seed) {
  const Setting &uc = cfg.lookup("unitcell");
  const Setting &lat = cfg.lookup("lattice");
  symops_ = uc.get<bool>("symops", true);
  read_materials(cfg.lookup("materials"));
  read_unitcell(uc);
  // core/lattice.cc:394-410
  const Setting &sz = lat["size"];
  for (int d = 0; d < 3; ++d) size_[d] = sz[d].as_int();
  if (lat.exists("periodic"))
    for (int d = 0; d < 3; ++d) periodic_[d] = lat["periodic"][d].as_bool();
  if (lat.exists("impurities")) throw RuntimeError("lattice.impurities is not supported by this build");
  if (lat.exists("global_rotation") || lat.exists("orientation_axis"))
    throw RuntimeError("lattice rotations are not supported by this build");
  for (int d = 0; d < 3; ++d)
    if (size_[d] <= 0) throw RuntimeError("lattice size must be positive");
  if (symops_) find_symmetry();

  // generate_supercell (core/lattice.cc:577-749)
  const int64_t n = (int64_t)size_[0] * size_[1] * size_[2] * (int64_t)basis_.size();
  if (n == 0) throw RuntimeError("the number of computed lattice sites was zero, check input");
  site_material_.resize((size_t)n);
  mus.resize((size_t)n);
  gyro.resize((size_t)n);
  alpha.resize((size_t)n);
  spins.resize((size_t)(3 * n));
  const bool normalise = lat.get<bool>("normalise_spins", true);
  std::vector<double> random_draws;
  bool any_random = false;
  for (auto &m : materials_) any_random |= m.randomize;
  if (any_random) {
    random_draws.resize((size_t)(3 * n));
    random_sphere_spins(seed, n, random_draws.data());
  }
  for (int64_t i = 0; i < n; ++i) {
    const int mi = basis_[i % (int64_t)basis_.size()].material;
    const Material &m = materials_[mi];
    site_material_[i] = mi;
    mus[i] = m.moment;
    alpha[i] = m.alpha;
    gyro[i] = gilbert_prefactor ? m.gyro / (1.0 + m.alpha * m.alpha) : m.gyro;
    Vec3 s = m.spin;
    if (m.randomize) s = {random_draws[3 * i], random_draws[3 * i + 1], random_draws[3 * i + 2]};
    if (m.moment == 0.0) s = {0.0, 0.0, 0.0};
    if (normalise) {
      const double len = norm(s);
      if (!(std::abs(len) <= 2.220446049250313e-16)) s = {s[0] / len, s[1] / len, s[2] / len};
    }
    for (int q = 0; q < 3; ++q) spins[3 * i + q] = s[q];
  }
  if (lat.exists("spins")) {
    const std::string fn = lat["spins"].as_string();
    spins = load_spins_tsv(fn, n);
  }
}

}  // namespace jams
