// plugins.cc -- see plugins.h.
#include "plugins.h"

#include "h5.h"

#include <algorithm>
#include <cstring>
#include <climits>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>

namespace jams {

void engine_check(jh_ctx *ctx, int rc) {
  if (rc != JH_OK) throw RuntimeError(std::string("engine: ") + jh_last_error(ctx));
}

namespace {

// core/units.h:15-26
double energy_unit(const std::string &name) {
  static const std::map<std::string, double> units = {
      {"joules", kJoule2meV}, {"J", kJoule2meV}, {"milli_electron_volts", 1.0}, {"meV", 1.0},
      {"milli_rydbergs", kmRyd2meV}, {"mRyd", kmRyd2meV}, {"rydbergs", kmRyd2meV * 1e3},
      {"Ryd", kmRyd2meV * 1e3}, {"Kelvin", kBoltzmannIU}, {"K", kBoltzmannIU}};
  auto it = units.find(name);
  if (it == units.end()) throw RuntimeError("energy units: " + name + " is not known");
  return it->second;
}

double max_abs(const Mat3 &a) {
  double m = 0.0;
  for (auto &r : a)
    for (double v : r) m = std::max(m, std::abs(v));
  return m;
}

Vec3 vec3(const Setting &s) {
  if (!s.is_aggregate() || s.length() != 3) throw ConfigError("expected a 3-vector at " + s.path());
  return {s[0].as_double(), s[1].as_double(), s[2].as_double()};
}

}  // namespace

// ---- Hamiltonian --------------------------------------------------------------

Hamiltonian::Hamiltonian(const Setting &s, const Lattice &lat) : lat_(lat) {
  // core/hamiltonian.cc:116-157
  std::string unit = s.get<std::string>("energy_units", "joules");
  if (s.exists("unit_name")) unit = s["unit_name"].as_string();
  energy_unit_conversion_ = energy_unit(unit);
  name_ = s["module"].as_string();
  field_.assign((size_t)(3 * lat.num_sites()), 0.0);
}

void Hamiltonian::calculate_fields(jh_ctx *ctx, double t) {
  engine_check(ctx, jh_compute_term_field(ctx, term_id_, t, field_.data()));
}

double Hamiltonian::calculate_total_energy(jh_ctx *ctx, double t) {
  double e = 0.0;
  engine_check(ctx, jh_term_energy(ctx, term_id_, t, &e));
  return e;
}

std::unique_ptr<Hamiltonian> Hamiltonian::create(const Setting &s, const Lattice &lat) {
  // core/hamiltonian.cc:80-113 (names of the terms on the hot path)
  if (s.exists("strategy"))
    throw RuntimeError("dipole hamiltonians now have specific names and 'strategy' has been removed");
  const std::string m = lowercase(s["module"].as_string());
  if (m == "exchange") return std::make_unique<ExchangeHamiltonian>(s, lat);
  if (m == "exchange-neartree") return std::make_unique<ExchangeNeartreeHamiltonian>(s, lat);
  if (m == "exchange-functional") return std::make_unique<ExchangeFunctionalHamiltonian>(s, lat);
  if (m == "zeeman") return std::make_unique<ZeemanHamiltonian>(s, lat, false);
  if (m == "applied-field") return std::make_unique<ZeemanHamiltonian>(s, lat, true);
  if (m == "uniaxial" || m == "uniaxial-anisotropy") return std::make_unique<UniaxialHamiltonian>(s, lat);
  if (m == "dipole-fft") return std::make_unique<DipoleFFTHamiltonian>(s, lat);
  throw RuntimeError("unknown hamiltonian " + s["module"].as_string() +
                     " (this build provides exchange, exchange-neartree, exchange-functional, "
                     "zeeman, applied-field, uniaxial, dipole-fft)");
  return nullptr;
}

// ---- exchange -------------------------------------------------------------------

ExchangeHamiltonian::ExchangeHamiltonian(const Setting &s, const Lattice &lat) : Hamiltonian(s, lat) {
  // hamiltonian/exchange.cc:12-172
  NeighbourOptions opt;
  opt.use_symops = s.get<bool>("symops", true);
  opt.energy_cutoff = s.get<double>("energy_cutoff", 0.0);
  opt.radius_cutoff = s.get<double>("radius_cutoff", 100.0);
  opt.distance_tolerance = s.get<double>("distance_tolerance", kLatticeTolerance);
  const double prefactor = s.get<double>("interaction_prefactor", 1.0);
  opt.check_no_zero = s.get<bool>("check_no_zero_motif_neighbour_count", true);
  opt.check_identical_count = s.get<bool>("check_identical_motif_neighbour_count", true);
  opt.check_identical_total = s.get<bool>("check_identical_motif_total_exchange", true);
  const bool check_symmetry = s.get<bool>("check_sparse_matrix_symmetry", true);
  const std::string fmt = capitalize(s.get<std::string>("coordinate_format", "CARTESIAN"));
  if (fmt == "CART" || fmt == "CARTESIAN") opt.coords = CoordinateFormat::kCartesian;
  else if (fmt == "FRAC" || fmt == "FRACTIONAL") opt.coords = CoordinateFormat::kFractional;
  else throw RuntimeError("Unknown coordinate format");
  // safety_check_distance_tolerance (core/interactions.cc:512-526)
  for (int i = 0; i < lat.num_basis_sites(); ++i)
    for (int j = i + 1; j < lat.num_basis_sites(); ++j)
      if (norm(lat.basis_site(i).frac - lat.basis_site(j).frac) < opt.distance_tolerance)
        throw RuntimeError("Atoms " + std::to_string(i) + " and " + std::to_string(j) +
                           " in the unit cell are closer together than the distance_tolerance");

  bool kkr = false;
  std::vector<InteractionData> rows;
  if (s.exists("exc_file")) rows = interactions_from_file(s["exc_file"].as_string(), kkr);
  else if (s.exists("interactions")) rows = interactions_from_setting(s["interactions"], kkr);
  else throw ConfigError("'exc_file' or 'interactions' settings are required");

  NeighbourList nl = generate_neighbour_list(lat, rows, kkr, opt);
  computed_ = nl.size();
  std::cout << "    computed interactions: " << computed_ << "\n";

  // insertion (exchange.cc:162-169): Jij = (prefactor * conversion) * J,
  // kept if max|Jij| > energy_cutoff * conversion
  const double scale = prefactor * energy_unit_conversion_;
  const double cutoff = opt.energy_cutoff * energy_unit_conversion_;
  std::vector<Mat3> vals(nl.values.size());
  std::vector<char> keep(nl.values.size());
  for (size_t v = 0; v < nl.values.size(); ++v) {
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) vals[v][a][b] = scale * nl.values[v][a][b];
    keep[v] = max_abs(vals[v]) > cutoff;
  }
  scaled_.values = vals;
  scaled_.i.reserve(nl.i.size());
  scaled_.j.reserve(nl.j.size());
  scaled_.v.reserve(nl.v.size());
  for (int64_t p = 0; p < nl.size(); ++p) {
    if (!keep[nl.v[p]]) continue;
    scaled_.i.push_back(nl.i[p]);
    scaled_.j.push_back(nl.j[p]);
    scaled_.v.push_back(nl.v[p]);
  }
  if (check_symmetry && !is_symmetric())
    throw RuntimeError("sparse matrix for " + name_ + " is not symmetric");
}

ExchangeHamiltonian::ExchangeHamiltonian(const Setting &s, const Lattice &lat, int)
    : Hamiltonian(s, lat) {}

ExchangeNeartreeHamiltonian::ExchangeNeartreeHamiltonian(const Setting &s, const Lattice &lat)
    : ExchangeHamiltonian(s, lat, 0) {
  // hamiltonian/exchange_neartree.cc:14-155; finalize(Symmetric)
  auto pairs = neartree_pairs(s, lat, energy_unit_conversion_, computed_);
  std::cout << "  total interactions " << computed_ << "\n  average interactions per spin "
            << (double)computed_ / (double)lat.num_sites() << "\n";
  scaled_ = pairs_to_list(std::move(pairs));
  if (!is_symmetric()) throw RuntimeError("sparse matrix for " + name_ + " is not symmetric");
}

ExchangeFunctionalHamiltonian::ExchangeFunctionalHamiltonian(const Setting &s, const Lattice &lat)
    : ExchangeHamiltonian(s, lat, 0) {
  // hamiltonian/exchange_functional.cc:97-258
  std::string check = "symmetric";
  if (s.exists("symmetry_check")) check = lowercase(s["symmetry_check"].as_string());
  if (check != "none" && check != "symmetric" && check != "force_symmetric")
    throw RuntimeError("invalid value for symmetry_check in ExchangeFunctionalHamiltonian");
  auto pairs = functional_pairs(s, lat, energy_unit_conversion_, computed_);
  std::cout << "  total interactions " << computed_ << "\n  average interactions per spin "
            << (double)computed_ / (double)lat.num_sites() << "\n";
  if (check == "force_symmetric") force_symmetric(pairs);
  scaled_ = pairs_to_list(std::move(pairs));
  if (check != "none" && !is_symmetric())
    throw RuntimeError("sparse matrix for " + name_ + (check == "force_symmetric" ? " failed to symmetrise"
                                                                                 : " is not symmetric"));
}

bool ExchangeHamiltonian::is_symmetric() const {
  // Builder::is_symmetric (sparse_matrix_builder.h:320-361) at block level:
  // A[3i+m][3j+n] == A[3j+n][3i+m] exactly, absent entries being zero
  const int64_t n = lat_.num_sites();
  std::vector<int64_t> ptr((size_t)n + 1, 0);
  for (auto i : scaled_.i) ptr[(size_t)i + 1]++;
  for (int64_t s = 0; s < n; ++s) ptr[s + 1] += ptr[s];
  for (int64_t p = 0; p < scaled_.size(); ++p) {
    const int32_t i = scaled_.i[p], j = scaled_.j[p];
    auto b = scaled_.j.begin() + ptr[j], e = scaled_.j.begin() + ptr[j + 1];
    auto it = std::lower_bound(b, e, i);
    const Mat3 &T = scaled_.values[scaled_.v[p]];
    if (it == e || *it != i) {
      if (max_abs(T) != 0.0) return false;
      continue;
    }
    const Mat3 &U = scaled_.values[scaled_.v[it - scaled_.j.begin()]];
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c)
        if (T[a][c] != U[c][a]) return false;
  }
  return true;
}

void ExchangeHamiltonian::build_csr(std::vector<int64_t> &rowptr, std::vector<int32_t> &col,
                                    std::vector<double> &val) const {
  // insert_interaction_tensor (sparse_interaction.cc:25-33) then the sorted,
  // merged CSR of build_csr (sparse_matrix_builder.h:187-248); the neighbour
  // list is (i, j)-sorted and duplicate free, so rows come out in order.
  const int64_t n = lat_.num_sites();
  rowptr.assign((size_t)(3 * n + 1), 0);
  for (int64_t p = 0; p < scaled_.size(); ++p) {
    const Mat3 &T = scaled_.values[scaled_.v[p]];
    for (int m = 0; m < 3; ++m)
      for (int c = 0; c < 3; ++c)
        if (T[m][c] != 0.0) rowptr[3 * (int64_t)scaled_.i[p] + m + 1]++;
  }
  for (int64_t r = 0; r < 3 * n; ++r) rowptr[r + 1] += rowptr[r];
  col.resize((size_t)rowptr[3 * n]);
  val.resize((size_t)rowptr[3 * n]);
  std::vector<int64_t> fill(rowptr.begin(), rowptr.end() - 1);
  for (int64_t p = 0; p < scaled_.size(); ++p) {
    const Mat3 &T = scaled_.values[scaled_.v[p]];
    for (int m = 0; m < 3; ++m)
      for (int c = 0; c < 3; ++c)
        if (T[m][c] != 0.0) {
          const int64_t r = 3 * (int64_t)scaled_.i[p] + m;
          col[fill[r]] = 3 * scaled_.j[p] + c;
          val[fill[r]] = T[m][c];
          fill[r]++;
        }
  }
}

void ExchangeHamiltonian::register_term(jh_ctx *ctx) {
  if (!ctx) return;
  std::vector<double> v9(scaled_.values.size() * 9);
  for (size_t v = 0; v < scaled_.values.size(); ++v)
    for (int q = 0; q < 9; ++q) v9[9 * v + q] = scaled_.values[v][q / 3][q % 3];
  engine_check(ctx, jh_add_exchange_list(ctx, scaled_.size(), scaled_.i.data(), scaled_.j.data(),
                                         scaled_.v.data(), (int64_t)scaled_.values.size(),
                                         v9.data(), &term_id_));
}

// ---- zeeman / applied field ------------------------------------------------

ZeemanHamiltonian::ZeemanHamiltonian(const Setting &s, const Lattice &lat, bool applied_field)
    : Hamiltonian(s, lat) {
  const int64_t n = lat.num_sites();
  dc.assign((size_t)(3 * n), 0.0);
  if (applied_field) {
    // applied_field.cc:84-148, static field only: h_i = mus_i B
    if (s.exists("type") && lowercase(s["type"].as_string()) != "static")
      throw RuntimeError("applied-field type " + s["type"].as_string() + " is not supported by this build");
    const Vec3 B = vec3(s["field"]);
    for (int64_t i = 0; i < n; ++i)
      for (int j = 0; j < 3; ++j) dc[3 * i + j] = lat.mus[i] * B[j];
    name_ += "-static";
    return;
  }
  // zeeman.cc:12-73
  const int nmat = lat.num_materials();
  if (s.exists("dc_local_field")) {
    const Setting &f = s["dc_local_field"];
    if (f.length() != nmat) throw ConfigError("dc_local_field must be specified for every material");
    for (int64_t i = 0; i < n; ++i) {
      const Vec3 B = vec3(f[lat.site_material(i)]);
      for (int j = 0; j < 3; ++j) {
        dc[3 * i + j] = B[j];
        dc[3 * i + j] *= lat.mus[i];
      }
    }
  }
  if (s.exists("ac_local_field") || s.exists("ac_local_frequency")) {
    if (!(s.exists("ac_local_field") && s.exists("ac_local_frequency")))
      throw ConfigError("ac_local_field must have a field and a frequency");
    const Setting &f = s["ac_local_field"], &w = s["ac_local_frequency"];
    if (f.length() != nmat || w.length() != nmat)
      throw ConfigError("ac_local_field and ac_local_frequency must be specified for every material");
    ac.assign((size_t)(3 * n), 0.0);
    omega.assign((size_t)n, 0.0);
    for (int64_t i = 0; i < n; ++i) {
      const Vec3 B = vec3(f[lat.site_material(i)]);
      for (int j = 0; j < 3; ++j) {
        ac[3 * i + j] = B[j];
        ac[3 * i + j] *= lat.mus[i];
      }
      omega[i] = 2.0 * kPi * w[lat.site_material(i)].as_double();
    }
  }
}

void ZeemanHamiltonian::register_term(jh_ctx *ctx) {
  if (!ctx) return;
  engine_check(ctx, jh_add_zeeman(ctx, dc.data(), ac.empty() ? nullptr : ac.data(),
                                  omega.empty() ? nullptr : omega.data(), &term_id_));
}

// ---- uniaxial ------------------------------------------------------------------

UniaxialHamiltonian::UniaxialHamiltonian(const Setting &s, const Lattice &lat) : Hamiltonian(s, lat) {
  // uniaxial_anisotropy.cc:79-113
  if (s.exists("d2z") || s.exists("d4z") || s.exists("d6z") || s.exists("K1") || s.exists("K2") ||
      s.exists("K3"))
    throw ConfigError("UniaxialHamiltonian: anisotropy should only be specified for a single K1, K2 or K3.");
  const std::string order = s["order"].as_string();
  if (order == "K1") power = 2;
  else if (order == "K2") power = 4;
  else if (order == "K3") power = 6;
  else throw RuntimeError("Unsupported anisotropy: " + order);
  const int64_t n = lat.num_sites();
  K.assign((size_t)n, 0.0);
  axis.assign((size_t)(3 * n), 0.0);
  const Setting &list = s["anisotropies"];
  for (int q = 0; q < list.length(); ++q) {
    const Setting &a = list[q];
    if (!a.is_list()) throw RuntimeError("Incorrectly formatted anisotropy setting");
    int motif = -1;
    int material = -1;
    if (a[0].is_number()) {
      motif = a[0].as_int() - 1;
      if (motif < 0 || motif >= lat.num_basis_sites()) throw RuntimeError("uniaxial anisotropy motif position is invalid");
    } else {
      material = lat.material_index(a[0].as_string());
      if (material < 0) throw RuntimeError("uniaxial anisotropy material is invalid");
    }
    Vec3 ax = vec3(a[1]);
    const double len = norm(ax);
    ax = {ax[0] / len, ax[1] / len, ax[2] / len};
    const double energy = a[2].as_double();
    for (int64_t i = 0; i < n; ++i) {
      if (lat.site_basis(i) == motif || (material >= 0 && lat.site_material(i) == material)) {
        K[i] = energy * energy_unit_conversion_;
        for (int j = 0; j < 3; ++j) axis[3 * i + j] = ax[j];
      }
    }
  }
}

void UniaxialHamiltonian::register_term(jh_ctx *ctx) {
  if (!ctx) return;
  engine_check(ctx, jh_add_uniaxial(ctx, K.data(), axis.data(), power, &term_id_));
}

// ---- dipole-fft ----------------------------------------------------------------

DipoleFFTHamiltonian::DipoleFFTHamiltonian(const Setting &s, const Lattice &lat) : Hamiltonian(s, lat) {
  r_cutoff = s["r_cutoff"].as_double();
}

void DipoleFFTHamiltonian::register_term(jh_ctx *ctx) {
  if (!ctx) return;
  int32_t size[3], per[3];
  for (int d = 0; d < 3; ++d) {
    size[d] = lat_.size(d);
    per[d] = lat_.periodic(d) ? 1 : 0;
  }
  double cell[9];
  for (int q = 0; q < 9; ++q) cell[q] = lat_.cell()[q / 3][q % 3];
  std::vector<double> frac;
  for (int m = 0; m < lat_.num_basis_sites(); ++m)
    for (int q = 0; q < 3; ++q) frac.push_back(lat_.basis_site(m).frac[q]);
  engine_check(ctx, jh_add_dipole_fft(ctx, size, per, lat_.num_basis_sites(), cell, frac.data(),
                                      lat_.parameter(), r_cutoff, &term_id_));
}

// ---- solver ----------------------------------------------------------------------

std::unique_ptr<Solver> Solver::create(const Setting &s) {
  // core/solver.cc:60-76: the GPU Heun and RK4 LLG solvers under their
  // reference names and explicit aliases
  const std::string m = lowercase(s["module"].as_string());
  if (m == "llg-heun-gpu" || m == "llg-heun-hip") return std::make_unique<HipHeunLLGSolver>();
  if (m == "llg-rk4-gpu" || m == "llg-rk4-hip") return std::make_unique<HipRK4LLGSolver>();
  throw RuntimeError("unknown solver " + s["module"].as_string() +
                     " (this build provides llg-heun-gpu / llg-heun-hip, llg-rk4-gpu / llg-rk4-hip)");
  return nullptr;
}

void HipHeunLLGSolver::initialize(const Setting &s) {
  // solvers/cuda_llg_heun.cu:21-37
  step_size_ = s["t_step"].as_double() / 1e-12;
  const double t_max = s["t_max"].as_double() / 1e-12;
  const double t_min = s.get<double>("t_min", 0.0) / 1e-12;
  max_steps_ = static_cast<int>(t_max / step_size_);
  min_steps_ = static_cast<int>(t_min / step_size_);
  gilbert_ = s.get<bool>("gilbert_prefactor", false);
  thermostat_ = s.get<std::string>("thermostat", "langevin-white-gpu");
  // core/thermostat.cc:25-27 aliases of the classical white-noise thermostat
  const std::string t = capitalize(thermostat_);
  if (t != "CLASSICAL-GPU" && t != "LANGEVIN-WHITE-GPU" && t != "CUDA_LANGEVIN_WHITE")
    throw RuntimeError("unknown thermostat " + thermostat_ +
                       " (this build provides langevin-white-gpu / classical-gpu)");
  std::cout << "  timestep (ps) " << step_size_ << "\n  t_max (ps) " << t_max << " steps "
            << max_steps_ << "\n  thermostat " << thermostat_ << "\n";
}

void Solver::run_steps(int64_t k) {
  if (k <= 0) return;
  engine_check(ctx_, jh_step(ctx_, k));
  iteration_ += k;
  time_ = (double)iteration_ * step_size_;
}

void Solver::register_monitor(std::unique_ptr<Monitor> m) { monitors_.push_back(std::move(m)); }

void Solver::resume_at(int64_t iteration) {
  if (iteration < 0) throw RuntimeError("checkpoint iteration must not be negative");
  iteration_ = iteration;
  time_ = (double)iteration_ * step_size_;
  if (ctx_) engine_check(ctx_, jh_set_iteration(ctx_, iteration));
}

void Solver::notify_monitors() {
  for (auto &m : monitors_)
    if (m->is_updating(iteration_)) m->update(*this);
}

void Solver::compute_fields(std::vector<double> &h) {
  h.resize(hamiltonians_.empty() ? 0 : hamiltonians_[0]->field().size());
  if (hamiltonians_.empty()) return;
  engine_check(ctx_, jh_compute_fields(ctx_, time_, h.data()));
}

// ---- monitors ------------------------------------------------------------------

// the closing tags of the XDMF time series, rewritten after each snapshot
static const char kXdmfTail[] = "    </Grid>\n  </Domain>\n</Xdmf>";

std::unique_ptr<Monitor> Monitor::create(const Setting &s, const std::string &prefix, Solver &solver,
                                         const Lattice &lat) {
  const std::string m = lowercase(s["module"].as_string());
  if (m == "magnetisation") return std::make_unique<MagnetisationMonitor>(s, prefix + "mag.tsv");
  if (m == "energy") return std::make_unique<EnergyMonitor>(s, prefix + "eng.tsv", solver);
  if (m == "checkpoint")
    return std::make_unique<CheckpointMonitor>(s, prefix + "checkpoint.tsv", lat.num_sites());
  if (m == "hdf5") {
    // the simulation name: the file name part of prefix, less its "_"
    std::string name = prefix.substr(0, prefix.empty() ? 0 : prefix.size() - 1);
    const size_t slash = name.find_last_of('/');
    if (slash != std::string::npos) name = name.substr(slash + 1);
    return std::make_unique<Hdf5Monitor>(s, prefix, name, lat);
  }
  throw RuntimeError("unknown monitor " + s["module"].as_string() +
                     " (this build provides magnetisation, energy, checkpoint, hdf5)");
  return nullptr;
}

MagnetisationMonitor::MagnetisationMonitor(const Setting &s, const std::string &path)
    : Monitor(s), tsv_(path) {
  normalize_ = s.get<bool>("normalize", true);
  tsv_ << std::setw(12) << "time" << "\tT\tHx\tHy\tHz\tmx\tmy\tmz\tm\n";
}

void MagnetisationMonitor::update(Solver &solver) {
  // m = sum mus s / (Kahan) sum mus, or / muB without normalisation
  double r[4];
  engine_check(solver.engine(), jh_reduce_magnetisation(solver.engine(), r));
  const double f = normalize_ ? 1.0 / r[3] : 1.0 / kBohrMagnetonIU;
  const double mag = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  char line[512];
  std::snprintf(line, sizeof line, "%.8e\t%.8e\t%.8e\t%.8e\t%.8e\t%.8e\t%.8e\t%.8e\t%.8e\n",
                solver.time(), solver.temperature(), solver.applied_field()[0],
                solver.applied_field()[1], solver.applied_field()[2], r[0] * f, r[1] * f, r[2] * f,
                mag * f);
  tsv_ << line;
  tsv_.flush();
}

void write_checkpoint(jh_ctx *ctx, int64_t n, const std::string &path, int64_t iteration,
                      double time_ps, uint64_t seed) {
  std::vector<double> s((size_t)(3 * n));
  engine_check(ctx, jh_get_spins(ctx, s.data()));
  const std::string tmp = path + ".tmp";
  {
    std::FILE *f = std::fopen(tmp.c_str(), "w");
    if (!f) throw RuntimeError("failed to open file: " + tmp);
    std::fprintf(f, "# jams checkpoint\n# num_spins %lld\n# iteration %lld\n# time_ps %.17g\n"
                    "# seed %llu\n",
                 (long long)n, (long long)iteration, time_ps, (unsigned long long)seed);
    for (int64_t i = 0; i < n; ++i)
      std::fprintf(f, "%.17g\t%.17g\t%.17g\n", s[3 * i], s[3 * i + 1], s[3 * i + 2]);
    const bool ok = std::fflush(f) == 0;
    std::fclose(f);
    if (!ok) throw RuntimeError("failed to write file: " + tmp);
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) throw RuntimeError("failed to write file: " + path);
}

CheckpointHeader read_checkpoint_header(const std::string &path) {
  std::ifstream f(path);
  if (!f.is_open()) throw RuntimeError("failed to open file: " + path);
  CheckpointHeader h;
  bool tagged = false, has_iteration = false;
  for (std::string line; std::getline(f, line);) {
    if (line.empty() || line[0] != '#') break;  // the header ends at the first data row
    std::istringstream is(line.substr(1));
    std::string key;
    is >> key;
    if (key == "jams") tagged = true;
    else if (key == "num_spins") is >> h.num_spins;
    else if (key == "iteration") has_iteration = static_cast<bool>(is >> h.iteration);
    else if (key == "seed") h.has_seed = static_cast<bool>(is >> h.seed);
  }
  if (!tagged || !has_iteration)
    throw ConfigError("sim.resume: " + path + " is not a checkpoint written by this build");
  return h;
}

void CheckpointMonitor::update(Solver &solver) {
  if (!is_updating(solver.iteration())) return;
  write_checkpoint(solver.engine(), n_, path_, solver.iteration(), solver.time(), solver.seed());
  written_ = solver.iteration();
}

void CheckpointMonitor::finish(Solver &solver) {
  if (written_ == solver.iteration()) return;
  write_checkpoint(solver.engine(), n_, path_, solver.iteration(), solver.time(), solver.seed());
  written_ = solver.iteration();
}

Hdf5Monitor::Hdf5Monitor(const Setting &s, const std::string &prefix, const std::string &sim_name,
                         const Lattice &lat)
    : Monitor(s), prefix_(prefix), sim_name_(sim_name), n_(lat.num_sites()) {
  if (output_steps_ == 0) output_steps_ = INT_MAX;  // only the final snapshot
  if (output_steps_ < 0) throw ConfigError("hdf5: output_steps must not be negative");
  compressed_ = s.get<bool>("compressed", false);
  fields_ = s.get<bool>("fields", false);
  for (const char *k : {"slice", "ds_dt", "energies"})
    if (s.exists(k) && (std::string(k) == "slice" || s[k].as_bool()))
      throw RuntimeError(std::string("hdf5: the ") + k + " setting is not supported by this build");
  // <sim>_lattice.h5 (monitors/hdf5.cc:151-193): material ids, moments and
  // positions in nm
  std::vector<int32_t> types((size_t)n_);
  std::vector<double> pos((size_t)(3 * n_));
  for (int64_t i = 0; i < n_; ++i) {
    types[(size_t)i] = lat.site_material(i);
    const Vec3 r = lat.site_position(i);
    for (int q = 0; q < 3; ++q) pos[(size_t)(3 * i + q)] = lat.parameter() * r[q] / 1e-9;
  }
  std::vector<h5::Data> sets(3);
  sets[0].path = "/types";
  sets[0].integer = true;
  sets[0].dims = {(uint64_t)n_};
  sets[0].data = types.data();
  sets[1].path = "/moments";
  sets[1].dims = {(uint64_t)n_};
  sets[1].data = lat.mus.data();
  sets[2].path = "/positions";
  sets[2].dims = {(uint64_t)n_, 3};
  sets[2].data = pos.data();
  sets[2].units = "nm";
  h5::write_file(prefix_ + "lattice.h5", sets);
  // <sim>.xdmf (monitors/hdf5.cc:197-214): an empty temporal collection,
  // each snapshot inserted before the closing tags
  const std::string xdmf = prefix_.substr(0, prefix_.size() - 1) + ".xdmf";
  xdmf_ = std::fopen(xdmf.c_str(), "w");
  if (!xdmf_) throw RuntimeError("failed to open XDMF output file: " + xdmf);
  std::fputs("<?xml version=\"1.0\"?>\n", xdmf_);
  std::fputs("<!DOCTYPE Xdmf SYSTEM \"https://gitlab.kitware.com/xdmf/xdmf/raw/master/Xdmf.dtd\"[]>\n",
             xdmf_);
  std::fputs("<Xdmf Version=\"3.0\">\n", xdmf_);
  std::fputs("  <Domain Name=\"JAMS\">\n", xdmf_);
  std::fputs("    <Information Name=\"Commit\" Value=\"jams_amd\" />\n", xdmf_);
  std::fprintf(xdmf_, "    <Information Name=\"Configuration\" Value=\"%s\" />\n", sim_name_.c_str());
  std::fputs("    <Grid Name=\"TimeSeries\" GridType=\"Collection\" CollectionType=\"Temporal\">\n",
             xdmf_);
  std::fputs(kXdmfTail, xdmf_);
  std::fflush(xdmf_);
}

Hdf5Monitor::~Hdf5Monitor() {
  if (xdmf_) std::fclose(xdmf_);
}

void Hdf5Monitor::write_spins(Solver &solver, const std::string &file) {
  std::vector<double> s((size_t)(3 * n_)), h;
  engine_check(solver.engine(), jh_get_spins(solver.engine(), s.data()));
  std::vector<h5::Data> sets(1);
  sets[0].path = "/spins";
  sets[0].dims = {(uint64_t)n_, 3};
  sets[0].data = s.data();
  if (fields_) {
    solver.compute_fields(h);
    h5::Data f;
    f.path = "/fields";
    f.dims = {(uint64_t)n_, 3};
    f.data = h.data();
    sets.push_back(f);
  }
  // chunks of up to 4095 rows, deflate level 6 (monitors/hdf5.cc:22-25)
  h5::write_file(file, sets, compressed_ ? 4095 : 0, 6);
  append_xdmf(file, solver.time());
}

void Hdf5Monitor::append_xdmf(const std::string &file, double time) {
  const size_t slash = file.find_last_of('/');
  const std::string base = slash == std::string::npos ? file : file.substr(slash + 1);
  const unsigned nn = (unsigned)n_;
  if (std::fseek(xdmf_, -(long)std::strlen(kXdmfTail), SEEK_CUR) != 0)
    throw RuntimeError("failed to seek XDMF file while appending a new timestep");
  std::fprintf(xdmf_, "      <Grid Name=\"Lattice\" GridType=\"Uniform\">\n");
  std::fprintf(xdmf_, "        <Time Value=\"%f\" />\n", time);
  std::fprintf(xdmf_, "        <Topology TopologyType=\"Polyvertex\" Dimensions=\"%u\" />\n", nn);
  std::fputs("       <Geometry GeometryType=\"XYZ\">\n", xdmf_);
  std::fprintf(xdmf_, "         <DataItem Dimensions=\"%u 3\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">\n", nn);
  std::fprintf(xdmf_, "           %s_lattice.h5:/positions\n", sim_name_.c_str());
  std::fputs("         </DataItem>\n       </Geometry>\n", xdmf_);
  std::fputs("       <Attribute Name=\"type\" AttributeType=\"Scalar\" Center=\"Node\">\n", xdmf_);
  std::fprintf(xdmf_, "         <DataItem Dimensions=\"%u\" NumberType=\"Int\" Precision=\"4\" Format=\"HDF\">\n", nn);
  std::fprintf(xdmf_, "           %s_lattice.h5:/types\n", sim_name_.c_str());
  std::fputs("         </DataItem>\n       </Attribute>\n", xdmf_);
  std::fputs("       <Attribute Name=\"moment\" AttributeType=\"Scalar\" Center=\"Node\">\n", xdmf_);
  std::fprintf(xdmf_, "         <DataItem Dimensions=\"%u\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">\n", nn);
  std::fprintf(xdmf_, "           %s_lattice.h5:/moments\n", sim_name_.c_str());
  std::fputs("         </DataItem>\n       </Attribute>\n", xdmf_);
  std::fputs("       <Attribute Name=\"spin\" AttributeType=\"Vector\" Center=\"Node\">\n", xdmf_);
  std::fprintf(xdmf_, "         <DataItem Dimensions=\"%u 3\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">\n", nn);
  std::fprintf(xdmf_, "           %s:/spins\n", base.c_str());
  std::fputs("         </DataItem>\n       </Attribute>\n", xdmf_);
  if (fields_) {
    std::fputs("       <Attribute Name=\"fields\" AttributeType=\"Vector\" Center=\"Node\">\n", xdmf_);
    std::fprintf(xdmf_, "         <DataItem Dimensions=\"%u 3\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">\n", nn);
    std::fprintf(xdmf_, "           %s:/fields\n", base.c_str());
    std::fputs("         </DataItem>\n       </Attribute>\n", xdmf_);
  }
  std::fputs("      </Grid>\n", xdmf_);
  std::fputs(kXdmfTail, xdmf_);
  std::fflush(xdmf_);
}

void Hdf5Monitor::update(Solver &solver) {
  if (!is_updating(solver.iteration())) return;
  char num[32];
  std::snprintf(num, sizeof num, "%07lld", (long long)(solver.iteration() / output_steps_));
  write_spins(solver, prefix_ + num + ".h5");
}

void Hdf5Monitor::finish(Solver &solver) {
  if (final_written_) return;
  write_spins(solver, prefix_ + "final.h5");
  final_written_ = true;
}

EnergyMonitor::EnergyMonitor(const Setting &s, const std::string &path, Solver &solver)
    : Monitor(s), tsv_(path) {
  tsv_ << "time\t";
  for (auto &h : solver.hamiltonians()) tsv_ << h->name() << "_E_meV\t";
  tsv_ << "\n";
}

void EnergyMonitor::update(Solver &solver) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%e\t", solver.time());
  tsv_ << buf;
  for (auto &h : solver.hamiltonians()) {
    std::snprintf(buf, sizeof buf, "%.15e\t", h->calculate_total_energy(solver.engine(), solver.time()));
    tsv_ << buf;
  }
  tsv_ << "\n";
  tsv_.flush();
}

}  // namespace jams
