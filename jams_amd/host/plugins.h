// plugins.h -- the reference's plugin API (Solver / Hamiltonian / Thermostat /
// Physics / Monitor, core/solver.h:15-88, core/hamiltonian.h:15-77,
// core/thermostat.h:10-39) restated on top of the gfx950 engine's C ABI.
//
// Hamiltonians are descriptors: each one registers its term with the engine
// in config order; calculate_fields and calculate_total_energy are served by
// the engine (jh_compute_term_field, jh_term_energy).  Module names are the
// reference's, so unmodified configs select these implementations.
#pragma once

#include <cstdint>
#include <fstream>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "../../include/jams_hip.h"
#include "common.h"
#include "config.h"
#include "exchange_pairs.h"
#include "interactions.h"
#include "lattice.h"

namespace jams {

class Hamiltonian {
 public:
  Hamiltonian(const Setting &s, const Lattice &lat);
  virtual ~Hamiltonian() = default;
  const std::string &name() const { return name_; }
  // add this term to the engine (ctx may be null in set-up-only mode)
  virtual void register_term(jh_ctx *ctx) = 0;
  int term_id() const { return term_id_; }
  // Hamiltonian::calculate_fields / calculate_total_energy through the engine
  void calculate_fields(jh_ctx *ctx, double t);
  double calculate_total_energy(jh_ctx *ctx, double t);
  const std::vector<double> &field() const { return field_; }

  static std::unique_ptr<Hamiltonian> create(const Setting &s, const Lattice &lat);

 protected:
  std::string name_;
  double energy_unit_conversion_ = 1.0;
  int term_id_ = -1;
  std::vector<double> field_;
  const Lattice &lat_;
};

// hamiltonian/exchange.cc:12-172 + sparse_interaction.cc
class ExchangeHamiltonian : public Hamiltonian {
 public:
  ExchangeHamiltonian(const Setting &s, const Lattice &lat);
  void register_term(jh_ctx *ctx) override;
  // neighbour list after insertion scaling and energy cutoff
  const NeighbourList &neighbours() const { return scaled_; }
  int64_t computed_interactions() const { return computed_; }
  // the reference's scalar CSR (insert_interaction_tensor zero skip, rows
  // ascending by column); int64 row pointers
  void build_csr(std::vector<int64_t> &rowptr, std::vector<int32_t> &col,
                 std::vector<double> &val) const;
  bool is_symmetric() const;

 protected:
  // a list built by another module (the distance-based producers below)
  ExchangeHamiltonian(const Setting &s, const Lattice &lat, int);
  NeighbourList scaled_;
  int64_t computed_ = 0;
};

// hamiltonian/exchange_neartree.cc: J I between sites of the listed material
// pair at a given distance (a shell of width shell_width)
class ExchangeNeartreeHamiltonian : public ExchangeHamiltonian {
 public:
  ExchangeNeartreeHamiltonian(const Setting &s, const Lattice &lat);
};

// hamiltonian/exchange_functional.cc: J(r_ij) I from a named functional of
// the pair vector, for pairs of the listed material pair within r_cutoff
class ExchangeFunctionalHamiltonian : public ExchangeHamiltonian {
 public:
  ExchangeFunctionalHamiltonian(const Setting &s, const Lattice &lat);
};

// hamiltonian/zeeman.cc and applied_field.cc: h = mus B (+ ac cos(omega t))
class ZeemanHamiltonian : public Hamiltonian {
 public:
  ZeemanHamiltonian(const Setting &s, const Lattice &lat, bool applied_field);
  void register_term(jh_ctx *ctx) override;
  std::vector<double> dc, ac, omega;
};

// hamiltonian/uniaxial_anisotropy.cc
class UniaxialHamiltonian : public Hamiltonian {
 public:
  UniaxialHamiltonian(const Setting &s, const Lattice &lat);
  void register_term(jh_ctx *ctx) override;
  std::vector<double> K, axis;
  int power = 2;
};

// hamiltonian/dipole_fft.cc
class DipoleFFTHamiltonian : public Hamiltonian {
 public:
  DipoleFFTHamiltonian(const Setting &s, const Lattice &lat);
  void register_term(jh_ctx *ctx) override;
  double r_cutoff = 0.0;
};

class Monitor;

class Solver {
 public:
  virtual ~Solver() = default;
  virtual void initialize(const Setting &s) = 0;
  virtual void run() = 0;                 // exactly one step
  virtual void run_steps(int64_t k);      // k steps with no monitor in between
  virtual std::string name() const = 0;
  virtual bool is_cuda_solver() const { return true; }
  // the engine integrator this solver drives (JH_INTEGRATOR_*)
  virtual int integrator() const { return JH_INTEGRATOR_HEUN; }
  bool is_running() const { return iteration_ < max_steps_; }
  int64_t iteration() const { return iteration_; }
  double time() const { return time_; }
  double time_step() const { return step_size_; }
  int64_t max_steps() const { return max_steps_; }
  double temperature() const { return temperature_; }
  const Vec3 &applied_field() const { return applied_field_; }
  bool gilbert_prefactor() const { return gilbert_; }
  uint64_t seed() const { return seed_; }
  void set_seed(uint64_t seed) { seed_ = seed; }
  const std::string &thermostat_name() const { return thermostat_; }

  void attach(jh_ctx *ctx) { ctx_ = ctx; }
  jh_ctx *engine() const { return ctx_; }
  void register_hamiltonian(std::unique_ptr<Hamiltonian> h) { hamiltonians_.push_back(std::move(h)); }
  void register_monitor(std::unique_ptr<Monitor> m);
  std::vector<std::unique_ptr<Hamiltonian>> &hamiltonians() { return hamiltonians_; }
  std::vector<std::unique_ptr<Monitor>> &monitors() { return monitors_; }
  void set_physics(double temperature, const Vec3 &field) {
    temperature_ = temperature;
    applied_field_ = field;
  }
  void notify_monitors();
  // continue a run from a checkpoint: iteration (and time) here and in the
  // engine, whose thermostat stream is keyed by the step counter
  void resume_at(int64_t iteration);
  // Solver::compute_fields (core/solver.cc:43-57) into h (N x 3)
  void compute_fields(std::vector<double> &h);

  static std::unique_ptr<Solver> create(const Setting &s);

 protected:
  jh_ctx *ctx_ = nullptr;
  int64_t iteration_ = 0, max_steps_ = 0, min_steps_ = 0;
  double time_ = 0.0, step_size_ = 1.0;
  double temperature_ = 0.0;
  Vec3 applied_field_ = {0, 0, 0};
  bool gilbert_ = false;
  uint64_t seed_ = 0;
  std::string thermostat_ = "langevin-white-gpu";
  std::vector<std::unique_ptr<Hamiltonian>> hamiltonians_;
  std::vector<std::unique_ptr<Monitor>> monitors_;
};

// CUDAHeunLLGSolver (solvers/cuda_llg_heun.cu) on the gfx950 engine
class HipHeunLLGSolver : public Solver {
 public:
  void initialize(const Setting &s) override;
  void run() override { run_steps(1); }
  std::string name() const override { return "llg-heun-gpu"; }
};

// CUDALLGRK4Solver (solvers/cuda_llg_rk4.cu over solvers/cuda_rk4_base.cu) on
// the same engine: the settings it reads (cuda_rk4_base.cu:10-48) are the Heun
// solver's, the engine runs the RK4 stages
class HipRK4LLGSolver : public HipHeunLLGSolver {
 public:
  std::string name() const override { return "llg-rk4-gpu"; }
  int integrator() const override { return JH_INTEGRATOR_RK4; }
};

class Monitor {
 public:
  Monitor(const Setting &s) : output_steps_(s.get<int>("output_steps", 100)) {}
  virtual ~Monitor() = default;
  bool is_updating(int64_t iteration) const { return iteration % output_steps_ == 0; }
  int64_t output_steps() const { return output_steps_; }
  virtual void update(Solver &solver) = 0;
  // after the run loop (the reference's Monitor::post_process)
  virtual void finish(Solver &solver) {}
  static std::unique_ptr<Monitor> create(const Setting &s, const std::string &prefix,
                                         Solver &solver, const Lattice &lat);

 protected:
  int64_t output_steps_;
};

// Checkpoint of the spin state (an addition: the reference restarts from
// lattice.spins = <file> with the time and the noise stream reset, and its
// HDF5 snapshots, written here by the hdf5 monitor, hold no step counter).
// The file is the
// reference's TSV spin format (helpers/load.h:22-51: N rows of sx sy sz,
// '#' lines skipped), written lossless (%.17g), headed by comment lines with
// the iteration, time and seed; sim.resume = <file> continues from it.
class CheckpointMonitor : public Monitor {
 public:
  CheckpointMonitor(const Setting &s, const std::string &path, int64_t n)
      : Monitor(s), path_(path), n_(n) {}
  void update(Solver &solver) override;
  void finish(Solver &solver) override;

 private:
  std::string path_;
  int64_t n_;
  int64_t written_ = -1;
};

// write / read the checkpoint file (n spins read back from the engine)
void write_checkpoint(jh_ctx *ctx, int64_t n, const std::string &path, int64_t iteration,
                      double time_ps, uint64_t seed);
struct CheckpointHeader {
  int64_t iteration = 0;
  int64_t num_spins = -1;
  bool has_seed = false;
  uint64_t seed = 0;
};
CheckpointHeader read_checkpoint_header(const std::string &path);

// monitors/hdf5.cc:27-52, :64-148, :151-289 (the "hdf5" module): at
// construction <sim>_lattice.h5 (types, moments, positions in nm) and the
// XDMF time series <sim>.xdmf; every output_steps (0: only the final one)
// <sim>_<count, 7 digits>.h5 with /spins (and /fields with fields = true),
// each appended to the XDMF; after the run <sim>_final.h5.  compressed =
// true stores the datasets chunked, shuffled and deflated.  The ds_dt,
// energies and slice settings are refused.  The HDF5 C library is loaded at
// run time (h5.h).
class Hdf5Monitor : public Monitor {
 public:
  Hdf5Monitor(const Setting &s, const std::string &prefix, const std::string &sim_name,
              const Lattice &lat);
  ~Hdf5Monitor() override;
  void update(Solver &solver) override;
  void finish(Solver &solver) override;

 private:
  void write_spins(Solver &solver, const std::string &file);
  void append_xdmf(const std::string &file, double time);
  std::string prefix_, sim_name_;
  std::FILE *xdmf_ = nullptr;
  int64_t n_ = 0;
  bool compressed_ = false, fields_ = false, final_written_ = false;
};

// monitors/magnetisation.cc:77-104 (one group: all spins)
class MagnetisationMonitor : public Monitor {
 public:
  MagnetisationMonitor(const Setting &s, const std::string &path);
  void update(Solver &solver) override;

 private:
  std::ofstream tsv_;
  bool normalize_ = true;
};

// monitors/energy.cc:22-33
class EnergyMonitor : public Monitor {
 public:
  EnergyMonitor(const Setting &s, const std::string &path, Solver &solver);
  void update(Solver &solver) override;

 private:
  std::ofstream tsv_;
};

void engine_check(jh_ctx *ctx, int rc);

}  // namespace jams
