// interactions.cc -- see interactions.h.
#include "interactions.h"

#include <algorithm>
#include <cfloat>
#include <fstream>
#include <sstream>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace jams {

int64_t NeighbourList::num_interactions(int64_t site) const {
  auto lo = std::lower_bound(i.begin(), i.end(), (int32_t)site);
  auto hi = std::upper_bound(lo, i.end(), (int32_t)site);
  return hi - lo;
}

namespace {

bool is_int_token(const std::string &s) {
  if (s.empty()) return false;
  size_t p = (s[0] == '-' || s[0] == '+') ? 1 : 0;
  if (p >= s.size()) return false;
  for (; p < s.size(); ++p)
    if (!std::isdigit((unsigned char)s[p])) return false;
  return true;
}

Mat3 scalar_tensor(double J) { return Mat3{{{J, 0, 0}, {0, J, 0}, {0, 0, J}}}; }

double max_abs(const Mat3 &a) {
  double m = 0.0;
  for (auto &r : a)
    for (double v : r) m = std::max(m, std::abs(v));
  return m;
}

// interactions.cc:58-75
Vec3 lattice_translation_vector(const Vec3 &r, double tol) {
  Vec3 T;
  for (int n = 0; n < 3; ++n) {
    const double ni = std::nearbyint(r[n]);
    T[n] = approximately_zero(r[n] - ni, tol) ? ni : std::floor(r[n]);
  }
  return T;
}

}  // namespace

std::vector<InteractionData> interactions_from_setting(const Setting &s, bool &kkr) {
  // discover_interaction_setting_format (:174-203)
  if (!s.is_list()) throw RuntimeError("exchange settings must be a list");
  if (s.length() == 0) return {};
  const Setting &r0 = s[0];
  if (r0[0].type() != r0[1].type()) throw RuntimeError("interaction type format is incorrect");
  kkr = r0[0].is_number();
  if (!r0[2].is_array()) throw RuntimeError("interaction vector format is incorrect");
  const bool scalar = r0[3].is_number();
  if (!scalar && !(r0[3].is_aggregate() && r0[3].length() == 9))
    throw RuntimeError("interaction energy format is incorrect");
  std::vector<InteractionData> out;
  for (int q = 0; q < s.length(); ++q) {
    const Setting &r = s[q];
    InteractionData d;
    if (kkr) {
      d.basis_site_i = r[0].as_int() - 1;
      d.basis_site_j = r[1].as_int() - 1;
    } else {
      d.type_i = r[0].as_string();
      d.type_j = r[1].as_string();
    }
    d.r = {r[2][0].as_double(), r[2][1].as_double(), r[2][2].as_double()};
    if (scalar) {
      d.J = scalar_tensor(r[3].as_double());
    } else {
      for (int x = 0; x < 9; ++x) d.J[x / 3][x % 3] = r[3][x].as_double();
    }
    out.push_back(d);
  }
  return out;
}

std::vector<InteractionData> interactions_from_file(const std::string &path, bool &kkr) {
  std::ifstream f(path);
  if (!f) throw RuntimeError("failed to open interaction file " + path);
  std::vector<InteractionData> out;
  bool discovered = false, tensor = false;
  int line_number = 0;
  for (std::string line; std::getline(f, line); ++line_number) {
    const size_t p = line.find_first_not_of(" \t\r");
    if (p == std::string::npos || line[p] == '#') continue;
    std::istringstream cols(line);
    std::vector<std::string> tok;
    for (std::string t; cols >> t;) tok.push_back(t);
    if (!discovered) {
      if (tok.size() == 6) tensor = false;
      else if (tok.size() == 14) tensor = true;
      else throw RuntimeError("interaction file has an incorrect number of columns");
      kkr = is_int_token(tok[0]) && is_int_token(tok[1]);
      discovered = true;
    }
    if (tok.size() != (tensor ? 14u : 6u))
      throw RuntimeError("failed to read line " + std::to_string(line_number) + " of interaction file");
    InteractionData d;
    try {
      if (kkr) {
        d.basis_site_i = std::stoi(tok[0]) - 1;
        d.basis_site_j = std::stoi(tok[1]) - 1;
      } else {
        d.type_i = tok[0];
        d.type_j = tok[1];
      }
      d.r = {std::stod(tok[2]), std::stod(tok[3]), std::stod(tok[4])};
      if (tensor) {
        for (int x = 0; x < 9; ++x) d.J[x / 3][x % 3] = std::stod(tok[5 + x]);
      } else {
        d.J = scalar_tensor(std::stod(tok[5]));
      }
    } catch (const std::exception &) {
      throw RuntimeError("failed to read line " + std::to_string(line_number) + " of interaction file");
    }
    out.push_back(d);
  }
  return out;
}

NeighbourList generate_neighbour_list(const Lattice &lat, std::vector<InteractionData> rows,
                                      bool kkr, const NeighbourOptions &opt) {
  const double tol = opt.distance_tolerance;
  auto mat_name = [&](int basis) { return lat.material(lat.basis_site(basis).material).name; };

  // fractional -> cartesian (:295-300)
  if (opt.coords == CoordinateFormat::kFractional)
    for (auto &d : rows) d.r = lat.fractional_to_cartesian(d.r);

  if (!kkr) {
    // complete_interaction_unitcell_positions (:98-123)
    std::vector<InteractionData> full;
    for (const auto &d : rows) {
      for (int bi = 0; bi < lat.num_basis_sites(); ++bi) {
        if (mat_name(bi) != d.type_i) continue;
        // find_unitcell_partner (:78-87)
        const Vec3 q = lat.cartesian_to_fractional(d.r) + lat.basis_site(bi).frac;
        const Vec3 off = q - lattice_translation_vector(q, tol);
        int bj = -1;
        for (int k = 0; k < lat.num_basis_sites(); ++k)
          if (approximately_equal(lat.basis_site(k).frac, off, tol)) {
            bj = k;
            break;
          }
        if (bj < 0 || mat_name(bj) != d.type_j) continue;
        InteractionData e = d;
        e.basis_site_i = bi;
        e.basis_site_j = bj;
        full.push_back(e);
      }
    }
    rows.swap(full);
  } else {
    // complete_interaction_typenames_names (:90-96)
    for (auto &d : rows) {
      if (d.basis_site_i < 0 || d.basis_site_i >= lat.num_basis_sites() || d.basis_site_j < 0 ||
          d.basis_site_j >= lat.num_basis_sites())
        throw RuntimeError("interaction refers to a unit-cell position that does not exist");
      d.type_i = mat_name(d.basis_site_i);
      d.type_j = mat_name(d.basis_site_j);
    }
  }

  if (opt.use_symops) {
    // apply_symops (:24-37): the tensor is copied unrotated to every point
    if (!lat.symops_enabled())
      throw RuntimeError("exchange symops requested but unitcell.symops is false");
    std::vector<InteractionData> sym;
    for (const auto &d : rows)
      for (const auto &p : lat.generate_symmetric_points(d.basis_site_i, d.r, tol)) {
        InteractionData e = d;
        e.r = p;
        sym.push_back(e);
      }
    rows.swap(sym);
  }
  if (opt.energy_cutoff > 0.0)
    rows.erase(std::remove_if(rows.begin(), rows.end(),
                              [&](const InteractionData &d) {
                                return definately_less_than(max_abs(d.J), opt.energy_cutoff, DBL_EPSILON);
                              }),
               rows.end());
  if (opt.radius_cutoff > 0.0)
    rows.erase(std::remove_if(rows.begin(), rows.end(),
                              [&](const InteractionData &d) {
                                return definately_greater_than(norm(d.r), opt.radius_cutoff, kLatticeTolerance);
                              }),
               rows.end());
  // lattice translation vectors (:331-344)
  for (auto &d : rows) {
    const Vec3 x = lat.cartesian_to_fractional(d.r) + lat.basis_site(d.basis_site_i).frac -
                   lat.basis_site(d.basis_site_j).frac;
    const Vec3 T = lattice_translation_vector(x, tol);
    for (int n = 0; n < 3; ++n) {
      if (!approximately_zero(T[n] - x[n], tol))
        throw RuntimeError("interaction vector is not a lattice translation between the two sites");
      d.T[n] = (int)T[n];
    }
  }

  // neighbour_list_from_interactions (:350-395), generated row by row: the
  // sites of a cell are visited in increasing index order, so sorting each
  // site's neighbours gives the (i, j)-sorted list of the reference's
  // VectorSet without a global sort.
  NeighbourList nl;
  std::vector<std::vector<int>> by_basis(lat.num_basis_sites());
  for (int q = 0; q < (int)rows.size(); ++q) by_basis[rows[q].basis_site_i].push_back(q);
  // distinct tensors: the template rows (values are per row)
  nl.values.reserve(rows.size());
  for (auto &d : rows) nl.values.push_back(d.J);
  const int64_t n = lat.num_sites();
  std::vector<int64_t> count((size_t)n + 1, 0);
  const int64_t ncell = (int64_t)lat.size(0) * lat.size(1) * lat.size(2);
  const int nb = lat.num_basis_sites();
  std::string error;
  auto neighbours_of = [&](int64_t cell, int m, std::vector<std::pair<int32_t, int32_t>> &out) -> bool {
    const int k = (int)(cell % lat.size(2));
    const int j = (int)((cell / lat.size(2)) % lat.size(1));
    const int i = (int)(cell / ((int64_t)lat.size(2) * lat.size(1)));
    out.clear();
    const int64_t local = lat.site_index(i, j, k, m);
    for (int q : by_basis[m]) {
      const InteractionData &d = rows[q];
      int a = i + d.T[0], b = j + d.T[1], c = k + d.T[2];
      if (!lat.apply_boundary_conditions(a, b, c)) continue;
      if (a < 0 || b < 0 || c < 0) {
#pragma omp critical(jams_nbr_error)
        error = "interaction reaches beyond one periodic image of the lattice";
        return false;
      }
      const int64_t nbr = lat.site_index(a, b, c, d.basis_site_j);
      if (lat.material(lat.site_material(local)).name != d.type_i ||
          lat.material(lat.site_material(nbr)).name != d.type_j)
        continue;
      out.emplace_back((int32_t)nbr, (int32_t)q);
    }
    std::sort(out.begin(), out.end());
    for (size_t x = 1; x < out.size(); ++x)
      if (out[x].first == out[x - 1].first) {
#pragma omp critical(jams_nbr_error)
        error = "Multiple interactions for sites " + std::to_string(local) + " and " +
                std::to_string(out[x].first);
        return false;
      }
    return true;
  };
  bool ok = true;
#pragma omp parallel reduction(&& : ok)
  {
    std::vector<std::pair<int32_t, int32_t>> buf;
#pragma omp for schedule(static)
    for (int64_t cell = 0; cell < ncell; ++cell)
      for (int m = 0; m < nb; ++m) {
        if (!neighbours_of(cell, m, buf)) {
          ok = false;
          continue;
        }
        count[cell * nb + m + 1] = (int64_t)buf.size();
      }
  }
  if (!ok) throw RuntimeError(error);
  for (int64_t s = 0; s < n; ++s) count[s + 1] += count[s];
  nl.i.resize((size_t)count[n]);
  nl.j.resize((size_t)count[n]);
  nl.v.resize((size_t)count[n]);
#pragma omp parallel
  {
    std::vector<std::pair<int32_t, int32_t>> buf;
#pragma omp for schedule(static)
    for (int64_t cell = 0; cell < ncell; ++cell)
      for (int m = 0; m < nb; ++m) {
        neighbours_of(cell, m, buf);
        const int64_t site = cell * nb + m;
        int64_t p = count[site];
        for (auto &e : buf) {
          nl.i[p] = (int32_t)site;
          nl.j[p] = e.first;
          nl.v[p] = e.second;
          ++p;
        }
      }
  }

  // neighbour_list_checks (:446-510)
  const bool all_periodic = lat.periodic(0) && lat.periodic(1) && lat.periodic(2);
  if (opt.check_no_zero)
    for (int64_t s = 0; s < n; ++s)
      if (count[s + 1] == count[s])
        throw RuntimeError("inconsistent neighbour list: some sites have no neighbours");
  if (opt.check_identical_count && all_periodic)
    for (int64_t s = 0; s < n; ++s)
      if (count[s + 1] - count[s] != count[s % nb + 1] - count[s % nb])
        throw RuntimeError(
            "inconsistent neighbour list: some sites have different numbers of neighbours for the "
            "same motif position");
  if (opt.check_identical_total && all_periodic) {
    auto diag_sum = [&](int64_t s) {
      Vec3 d = {0, 0, 0};
      for (int64_t p = count[s]; p < count[s + 1]; ++p)
        for (int q = 0; q < 3; ++q) d[q] += nl.values[nl.v[p]][q][q];
      return d;
    };
    std::vector<Vec3> ref(nb);
    for (int m = 0; m < nb; ++m) ref[m] = diag_sum(m);
    for (int64_t s = 0; s < n; ++s)
      if (!approximately_equal(diag_sum(s), ref[s % nb], 1e-6))
        throw RuntimeError("inconsistent neighbour list: J0");
  }
  return nl;
}

}  // namespace jams
