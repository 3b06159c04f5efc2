// exchange_pairs.cc -- the neighbour lists of the exchange-neartree and
// exchange-functional modules (hamiltonian/exchange_neartree.cc:14-155,
// hamiltonian/exchange_functional.cc:97-459), found by distance in the
// periodic supercell.
//
// The reference searches a near tree of the supercell's sites plus a halo of
// periodic images (lattice/interaction_neartree.cc:9-110).  Here the same
// images go into a uniform cell list; which pairs are accepted depends only
// on each candidate's distance, computed with the same expression, so the
// accepted set and the values do not depend on the search structure.
#include "exchange_pairs.h"

#include <algorithm>
#include <cmath>
#include <iostream>
#include <map>
#include <sstream>
#include <unordered_map>

namespace jams {

namespace {

// helpers/maths.h:38-40 (the other comparisons are in common.h)
bool less_than_approx_equal(double a, double b, double eps) {
  return (a - b) < (std::max(std::abs(a), std::abs(b)) * eps);
}

Vec3 column(const Mat3 &m, int d) { return {m[0][d], m[1][d], m[2][d]}; }

// distance between the two faces of the parallelepiped (a, b, c) spanned by
// a and b (maths/parallelepiped.cc:18-24)
double height(const Vec3 &a, const Vec3 &b, const Vec3 &c) {
  return std::abs(dot(cross(a, b), c)) / norm(cross(a, b));
}

}  // namespace

std::array<Vec3, 3> supercell_vectors(const Lattice &lat) {
  std::array<Vec3, 3> A;
  for (int d = 0; d < 3; ++d) A[d] = (double)lat.size(d) * column(lat.cell(), d);
  return A;
}

double max_interaction_radius(const Lattice &lat) {
  // jams::maximum_interaction_length (core/lattice.cc:1155-1201)
  const auto A = supercell_vectors(lat);
  const bool p0 = lat.periodic(0), p1 = lat.periodic(1), p2 = lat.periodic(2);
  const auto pg_in = [](const Vec3 &a, const Vec3 &b) {
    const double area = norm(cross(a, b));
    return 0.5 * std::min(area / norm(b), area / norm(a));
  };
  if (p0 && p1 && p2)
    return 0.5 * std::min({height(A[1], A[2], A[0]), height(A[2], A[0], A[1]), height(A[0], A[1], A[2])});
  if (p0 && p1) return pg_in(A[0], A[1]);
  if (p0 && p2) return pg_in(A[0], A[2]);
  if (p1 && p2) return pg_in(A[1], A[2]);
  if (p0) return 0.5 * norm(A[0]);
  if (p1) return 0.5 * norm(A[1]);
  if (p2) return 0.5 * norm(A[2]);
  const Vec3 &a = A[0], &b = A[1], &c = A[2];
  return std::max({norm(a + b + c), norm(-a + b + c), norm(a - b + c), norm(a + b - c)});
}

PairSearch::PairSearch(const Lattice &lat, double r_search_in) : lat_(lat) {
  // a little beyond the radius: the callers' tests accept distances up to a
  // relative tolerance above it
  const double r_search = r_search_in * (1.0 + 1e-3) + 1e-9;
  const auto A = supercell_vectors(lat);
  const int64_t n = lat.num_sites();
  pos_.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) pos_[(size_t)i] = lat.site_position(i);
  // image counts per periodic direction (interaction_neartree.cc:18-20)
  int nimg[3];
  const double h[3] = {height(A[1], A[2], A[0]), height(A[2], A[0], A[1]), height(A[0], A[1], A[2])};
  for (int d = 0; d < 3; ++d) nimg[d] = lat.periodic(d) ? (int)std::ceil(r_search / h[d]) : 0;
  // images whose fractional coordinates lie within r_search of the
  // supercell (the only ones that can be within r_search of a site)
  Mat3 Am;
  for (int r = 0; r < 3; ++r)
    for (int d = 0; d < 3; ++d) Am[r][d] = A[d][r];
  const Mat3 Ainv = inverse(Am);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t j = 0; j < n; ++j)
    for (int a = -nimg[0]; a <= nimg[0]; ++a)
      for (int b = -nimg[1]; b <= nimg[1]; ++b)
        for (int c = -nimg[2]; c <= nimg[2]; ++c) {
          // r + (h a + k b + l c), as interaction_neartree.cc:67 adds them
          const Vec3 p = pos_[(size_t)j] + (((double)a * A[0] + (double)b * A[1]) + (double)c * A[2]);
          if (a || b || c) {
            const Vec3 f = Ainv * p;
            bool keep = true;
            for (int d = 0; d < 3 && keep; ++d) {
              const double m = r_search / h[d] * (1.0 + 1e-9) + 1e-9;
              keep = f[d] >= -m && f[d] <= 1.0 + m;
            }
            if (!keep) continue;
          }
          img_.push_back({p, (int32_t)j});
          for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], p[d]);
            hi[d] = std::max(hi[d], p[d]);
          }
        }
  // uniform bins of edge >= r_search
  bin_ = std::max(r_search, 1e-6);
  for (int d = 0; d < 3; ++d) {
    org_[d] = lo[d];
    nb_[d] = std::max<int64_t>(1, (int64_t)std::floor((hi[d] - lo[d]) / bin_) + 1);
  }
  const int64_t nbins = nb_[0] * nb_[1] * nb_[2];
  start_.assign((size_t)nbins + 1, 0);
  std::vector<int64_t> key(img_.size());
  for (size_t q = 0; q < img_.size(); ++q) {
    key[q] = bin_of(img_[q].p);
    start_[(size_t)key[q] + 1]++;
  }
  for (int64_t b = 0; b < nbins; ++b) start_[(size_t)b + 1] += start_[(size_t)b];
  order_.resize(img_.size());
  std::vector<int64_t> fill(start_.begin(), start_.end() - 1);
  for (size_t q = 0; q < img_.size(); ++q) order_[(size_t)fill[(size_t)key[q]]++] = (int64_t)q;
}

int64_t PairSearch::bin_coord(double x, int d) const {
  const int64_t c = (int64_t)std::floor((x - org_[d]) / bin_);
  return std::min<int64_t>(std::max<int64_t>(c, 0), nb_[d] - 1);
}

int64_t PairSearch::bin_of(const Vec3 &p) const {
  return (bin_coord(p[0], 0) * nb_[1] + bin_coord(p[1], 1)) * nb_[2] + bin_coord(p[2], 2);
}

void PairSearch::candidates(int64_t i, std::vector<Candidate> &out) const {
  out.clear();
  const Vec3 &r = pos_[(size_t)i];
  int64_t c[3];
  for (int d = 0; d < 3; ++d) c[d] = bin_coord(r[d], d);
  for (int64_t x = std::max<int64_t>(c[0] - 1, 0); x <= std::min(c[0] + 1, nb_[0] - 1); ++x)
    for (int64_t y = std::max<int64_t>(c[1] - 1, 0); y <= std::min(c[1] + 1, nb_[1] - 1); ++y)
      for (int64_t z = std::max<int64_t>(c[2] - 1, 0); z <= std::min(c[2] + 1, nb_[2] - 1); ++z) {
        const int64_t b = (x * nb_[1] + y) * nb_[2] + z;
        for (int64_t q = start_[(size_t)b]; q < start_[(size_t)b + 1]; ++q) {
          const Image &im = img_[(size_t)order_[(size_t)q]];
          // the tree's distance, norm(origin - item)
          out.push_back({im.j, im.p - r, norm(r - im.p)});
        }
      }
  // a fixed order (it does not affect the result)
  std::sort(out.begin(), out.end(), [](const Candidate &a, const Candidate &b) {
    return a.j != b.j ? a.j < b.j : a.d < b.d;
  });
}

double distance_unit(const Setting &s, const Lattice &lat) {
  // core/hamiltonian.cc:139-155, distances converted to lattice constants.
  // ("A" is not accepted: the reference maps it to 1e-10 / (a 1e10), which
  // is not angstroms; "angstroms" is.)
  const std::string u = s.get<std::string>("distance_units", "lattice_constants");
  const double a = lat.parameter();
  if (u == "lattice_constants") return 1.0;
  if (u == "m" || u == "meters") return 1.0 / a;
  if (u == "nm" || u == "nanometers") return 1e-9 / a;
  if (u == "angstroms") return 1e-10 / a;
  throw RuntimeError("distance units: " + u + " is not known (lattice_constants, m, meters, nm, "
                     "nanometers, angstroms)");
}

NeighbourList pairs_to_list(std::vector<PairValue> pairs) {
  std::sort(pairs.begin(), pairs.end(), [](const PairValue &a, const PairValue &b) {
    return a.i != b.i ? a.i < b.i : a.j < b.j;
  });
  NeighbourList nl;
  std::map<Mat3, int32_t> index;
  for (const auto &p : pairs) {
    auto it = index.find(p.J);
    if (it == index.end()) {
      it = index.emplace(p.J, (int32_t)nl.values.size()).first;
      nl.values.push_back(p.J);
    }
    nl.i.push_back(p.i);
    nl.j.push_back(p.j);
    nl.v.push_back(it->second);
  }
  return nl;
}

// SparseMatrix::Builder::symmetrise (containers/sparse_matrix_builder.h:
// 364-425) on the scalar entries of diagonal tensors J I: a missing (j, i)
// entry is added with J_ij; two different values both become (J_ij+J_ji)/2.
void force_symmetric(std::vector<PairValue> &pairs) {
  std::sort(pairs.begin(), pairs.end(), [](const PairValue &a, const PairValue &b) {
    return a.i != b.i ? a.i < b.i : a.j < b.j;
  });
  std::vector<PairValue> extra;
  const auto find = [&](int32_t i, int32_t j) -> PairValue * {
    auto it = std::lower_bound(pairs.begin(), pairs.end(), PairValue{i, j, {}},
                               [](const PairValue &a, const PairValue &b) {
                                 return a.i != b.i ? a.i < b.i : a.j < b.j;
                               });
    return (it != pairs.end() && it->i == i && it->j == j) ? &*it : nullptr;
  };
  for (size_t n = 0; n < pairs.size(); ++n) {
    PairValue &ij = pairs[n];
    PairValue *ji = find(ij.j, ij.i);
    if (!ji) {
      Mat3 T{};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) T[a][b] = ij.J[b][a];
      extra.push_back({ij.j, ij.i, T});
      continue;
    }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double &x = ij.J[a][b], &y = ji->J[b][a];
        if (x != y) {
          const double s = 0.5 * (x + y);
          x = s;
          y = s;
        }
      }
  }
  pairs.insert(pairs.end(), extra.begin(), extra.end());
}

// ---- exchange-neartree -----------------------------------------------------

std::vector<PairValue> neartree_pairs(const Setting &s, const Lattice &lat, double energy_conv,
                                      int64_t &counter) {
  const double dconv = distance_unit(s, lat);
  const double energy_cutoff = s.get<double>("energy_cutoff", 1e-26) * energy_conv;
  const double width = s.get<double>("shell_width", 1e-3) * dconv;
  std::cout << "  energy_cutoff " << energy_cutoff << "\n  shell_width " << width / dconv << "\n";
  for (int i = 0; i < lat.num_basis_sites(); ++i)
    for (int j = i + 1; j < lat.num_basis_sites(); ++j)
      if (norm(lat.basis_site(i).frac - lat.basis_site(j).frac) < width)
        throw RuntimeError("Atoms " + std::to_string(i) + " and " + std::to_string(j) +
                           " in the unit cell are closer together than the shell_width");
  if (!s.exists("interactions")) throw ConfigError("no 'interactions' setting in ExchangeNeartree hamiltonian");
  struct Shell { int tb; double radius, J; };
  std::vector<std::vector<Shell>> list((size_t)lat.num_materials());
  double max_radius = 0.0;
  const Setting &is = s["interactions"];
  for (int k = 0; k < is.length(); ++k) {
    const std::string A = is[k][0].as_string(), B = is[k][1].as_string();
    const int ta = lat.material_index(A), tb = lat.material_index(B);
    if (ta < 0 || tb < 0)
      throw RuntimeError("exchange neartree interaction " + std::to_string(k) + ": material " +
                         (ta < 0 ? A : B) + " does not exist in the config");
    const double radius = is[k][2].as_double() * dconv;
    const double J = is[k][3].as_double() * energy_conv;
    max_radius = std::max(max_radius, radius);
    list[(size_t)ta].push_back({tb, radius, J});
    if (ta != tb) list[(size_t)tb].push_back({ta, radius, J});
  }
  // the tree holds sites up to max_radius + width; its tolerance is width/10
  PairSearch search(lat, max_radius + width);
  const double eps = width / 10.0;
  std::vector<PairValue> pairs;
  std::vector<int64_t> seen((size_t)lat.num_sites(), -1);
  std::vector<PairSearch::Candidate> cand;
  counter = 0;
  for (int64_t i = 0; i < lat.num_sites(); ++i) {
    search.candidates(i, cand);
    const int ti = lat.site_material(i);
    for (const Shell &sh : list[(size_t)ti]) {
      // InteractionNearTree::shell: annulus (radius - w/2, radius + w/2]
      const double inner = sh.radius - 0.5 * width, outer = sh.radius + 0.5 * width;
      for (const auto &c : cand) {
        if (definately_greater_than(c.d, outer, eps) || !definately_greater_than(c.d, inner, eps))
          continue;
        if (c.j == i) continue;
        if (lat.site_material(c.j) != sh.tb) continue;
        if (seen[(size_t)c.j] == i)
          throw RuntimeError("multiple interactions between spins " + std::to_string(i) + " and " +
                             std::to_string(c.j));
        seen[(size_t)c.j] = i;
        if (std::abs(sh.J) > energy_cutoff) {
          Mat3 T{};
          for (int a = 0; a < 3; ++a) T[a][a] = sh.J;
          pairs.push_back({(int32_t)i, c.j, T});
          ++counter;
        }
      }
    }
  }
  return pairs;
}

// ---- exchange-functional ---------------------------------------------------

namespace {

int expected_parameters(const std::string &f) {
  if (f == "rkky" || f == "exponential" || f == "gaussian" || f == "kaneyoshi") return 3;
  if (f == "gaussian_multi") return 9;
  if (f == "c3z") return 14;
  if (f == "step") return 2;
  return -1;
}

Mat3 rotation_z(double t) {
  Mat3 R{};
  R[0][0] = std::cos(t); R[0][1] = -std::sin(t);
  R[1][0] = std::sin(t); R[1][1] = std::cos(t);
  R[2][2] = 1.0;
  return R;
}

double pow2(double x) { return x * x; }

double f_gaussian(double r, double J0, double r0, double sigma) {
  return J0 * std::exp(-pow2(r - r0) / (2 * pow2(sigma)));
}

}  // namespace

Functional make_functional(const std::string &name, const std::vector<double> &p, double econv,
                           double dconv) {
  const int want = expected_parameters(name);
  if (want < 0) throw RuntimeError("unknown exchange functional: " + name);
  if ((int)p.size() != want)
    throw RuntimeError("the '" + name + "' exchange functional takes " + std::to_string(want) +
                       " parameters (" + std::to_string(p.size()) + " given)");
  const double tol = kLatticeTolerance;
  const auto nonzero = [&](int k, const char *what) {
    if (approximately_zero(p[(size_t)k], tol))
      throw RuntimeError("exchange functional '" + name + "' requires non-zero parameter '" + what + "'");
  };
  const auto positive = [&](int k, const char *what) {
    if (!definately_greater_than(p[(size_t)k], 0.0, tol))
      throw RuntimeError("exchange functional '" + name + "' requires positive parameter '" + what + "'");
  };
  std::cout << "  exchange functional: " << name << "\n";
  // hamiltonian/exchange_functional.cc:265-411
  if (name == "rkky") {
    nonzero(2, "k_F");
    const double J0 = econv * p[0], r0 = dconv * p[1], kF = p[2];
    return [=](const Vec3 &rij) {
      const double r = norm(rij);
      const double kr = 2 * kF * (r - r0);
      if (approximately_zero(kr, kLatticeTolerance))
        throw RuntimeError("rkky exchange functional: the form is singular where 2 k_F (r - r0) = 0");
      return -J0 * (kr * std::cos(kr) - std::sin(kr)) / (kr * kr * kr * kr);
    };
  }
  if (name == "exponential") {
    nonzero(2, "sigma");
    const double J0 = econv * p[0], r0 = dconv * p[1], sg = dconv * p[2];
    return [=](const Vec3 &rij) { return J0 * std::exp(-(norm(rij) - r0) / sg); };
  }
  if (name == "gaussian") {
    nonzero(2, "sigma");
    const double J0 = econv * p[0], r0 = dconv * p[1], sg = dconv * p[2];
    return [=](const Vec3 &rij) { return f_gaussian(norm(rij), J0, r0, sg); };
  }
  if (name == "gaussian_multi") {
    nonzero(2, "sigma0");
    nonzero(5, "sigma1");
    nonzero(8, "sigma2");
    double q[9];
    for (int k = 0; k < 9; ++k) q[k] = (k % 3 == 0 ? econv : dconv) * p[(size_t)k];
    return [=](const Vec3 &rij) {
      const double r = norm(rij);
      return f_gaussian(r, q[0], q[1], q[2]) + f_gaussian(r, q[3], q[4], q[5]) +
             f_gaussian(r, q[6], q[7], q[8]);
    };
  }
  if (name == "kaneyoshi") {
    nonzero(2, "sigma");
    const double J0 = econv * p[0], r0 = dconv * p[1], sg = dconv * p[2];
    return [=](const Vec3 &rij) {
      const double r = norm(rij);
      return J0 * pow2(r - r0) * std::exp(-pow2(r - r0) / (2 * pow2(sg)));
    };
  }
  if (name == "c3z") {
    positive(10, "l0");
    positive(11, "l1s");
    positive(12, "l1c");
    const Vec3 qs1 = (1.0 / dconv) * Vec3{p[0], p[1], p[2]};
    const Vec3 qc1 = (1.0 / dconv) * Vec3{p[3], p[4], p[5]};
    const double J0 = econv * p[6], J1s = econv * p[7], J1c = econv * p[8];
    const double d0 = dconv * p[9], l0 = dconv * p[10], l1s = dconv * p[11], l1c = dconv * p[12],
                 rs = dconv * p[13];
    const Mat3 R1 = rotation_z(2 * kPi / 3.0), R2 = rotation_z(4 * kPi / 3.0);
    const Vec3 qs[3] = {qs1, R1 * qs1, R2 * qs1}, qc[3] = {qc1, R1 * qc1, R2 * qc1};
    return [=](const Vec3 &rij) {
      const double r = norm(rij);
      const Vec3 rp = {rij[0], rij[1], 0.0};
      const double t0 = J0 * std::exp(-std::abs(r - d0) / l0);
      double ss = 0.0, cs = 0.0;
      for (const auto &q : qs) ss += std::sin(dot(q, rp));
      const double ts = J1s * std::exp(-std::abs(r - rs) / l1s) * ss;
      for (const auto &q : qc) cs += std::cos(dot(q, rp));
      const double tc = J1c * std::exp(-std::abs(r - rs) / l1c) * cs;
      return (t0 + ts + tc);
    };
  }
  // step
  const double J0 = econv * p[0], rc = dconv * p[1];
  return [=](const Vec3 &rij) {
    return less_than_approx_equal(norm(rij), rc, kLatticeTolerance) ? J0 : 0.0;
  };
}

std::vector<PairValue> functional_pairs(const Setting &s, const Lattice &lat, double energy_conv,
                                        int64_t &counter) {
  const double dconv = distance_unit(s, lat);
  if (!s.exists("interactions")) throw ConfigError("no 'interactions' setting in ExchangeFunctional hamiltonian");
  struct Entry { double rc; Functional f; };
  std::map<std::pair<int, int>, Entry> table;
  double max_cutoff = 0.0;
  const double rmax = max_interaction_radius(lat);
  const Setting &is = s["interactions"];
  for (int k = 0; k < is.length(); ++k) {
    const Setting &e = is[k];
    if (e.length() < 4) throw ConfigError("interaction requires at least 4 elements");
    const std::string A = e[0].as_string(), B = e[1].as_string(), fname = e[2].as_string();
    const double rc = dconv * e[3].as_double();
    const int ta = lat.material_index(A), tb = lat.material_index(B);
    if (ta < 0 || tb < 0)
      throw ConfigError("material " + (ta < 0 ? A : B) + " does not exist in config");
    if (definately_less_than(rc, 0.0, kLatticeTolerance)) throw ConfigError("cutoff radius cannot be negative");
    if (table.count({ta, tb}))
      throw RuntimeError("Interaction between types \"" + A + "\" and \"" + B + "\" is defined more than once.");
    if (rc > rmax) {
      std::ostringstream m;
      m << "cutoff radius " << rc << " is larger than the maximum cutoff radius " << rmax;
      throw RuntimeError(m.str());
    }
    max_cutoff = std::max(max_cutoff, rc);
    std::vector<double> params;
    for (int q = 4; q < e.length(); ++q) {
      if (e[q].is_list() || e[q].is_array()) {
        for (int r = 0; r < e[q].length(); ++r) params.push_back(e[q][r].as_double());
      } else {
        params.push_back(e[q].as_double());
      }
    }
    table[{ta, tb}] = {rc, make_functional(fname, params, energy_conv, dconv)};
  }
  PairSearch search(lat, max_cutoff);
  const double eps = kLatticeTolerance;
  std::vector<PairValue> pairs;
  std::vector<int64_t> seen((size_t)lat.num_sites(), -1);
  std::vector<PairSearch::Candidate> cand;
  counter = 0;
  for (int64_t i = 0; i < lat.num_sites(); ++i) {
    search.candidates(i, cand);
    const int ti = lat.site_material(i);
    for (const auto &c : cand) {
      // InteractionNearTree::neighbours: annulus (eps, max_cutoff]
      if (definately_greater_than(c.d, max_cutoff, eps) || !definately_greater_than(c.d, eps, eps))
        continue;
      if (c.j == i) continue;
      auto it = table.find({ti, lat.site_material(c.j)});
      if (it == table.end()) continue;
      if (!less_than_approx_equal(norm(c.rij), it->second.rc, eps)) continue;
      if (seen[(size_t)c.j] == i)
        throw RuntimeError("multiple interactions between spins " + std::to_string(i) + " and " +
                           std::to_string(c.j));
      seen[(size_t)c.j] = i;
      const double J = it->second.f(c.rij);
      ++counter;
      if (J == 0.0) continue;  // insert_interaction_scalar skips zeros
      Mat3 T{};
      for (int a = 0; a < 3; ++a) T[a][a] = J;
      pairs.push_back({(int32_t)i, c.j, T});
    }
  }
  return pairs;
}

}  // namespace jams
