// simulation.cc -- the JAMS-compatible set-up and run loop
// (core/jams++.cc:206-377) behind include/jams_host.h.
#include <chrono>
#include <cstring>
#include <iostream>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/jams_host.h"
#include "config.h"
#include "lattice.h"
#include "plugins.h"

using namespace jams;

namespace {
thread_local std::string g_error;
}

struct jams_sim {
  std::string err;
  Config cfg;
  Lattice lattice;
  std::unique_ptr<Solver> solver;
  jh_ctx *ctx = nullptr;
  uint64_t seed = 0;
  std::string prefix;

  ~jams_sim() {
    if (ctx) jh_destroy(ctx);
  }
};

namespace {

ExchangeHamiltonian *exchange_term(const jams_sim *s, int index) {
  int k = 0;
  for (auto &h : s->solver->hamiltonians())
    if (auto *x = dynamic_cast<ExchangeHamiltonian *>(h.get())) {
      if (k == index) return x;
      ++k;
    }
  throw RuntimeError("no exchange Hamiltonian with index " + std::to_string(index));
}

template <class F>
int guarded(jams_sim *s, F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    if (s) s->err = e.what();
    else g_error = e.what();
    return 1;
  }
}

void initialize(jams_sim *s, const std::vector<std::string> &sources, int device) {
  s->cfg.merge_sources(sources);
  const Config &cfg = s->cfg;
  // seeds: sim.seed drives the thermostat stream and "random" spins; without
  // it the seed is drawn once and printed so the run can be repeated
  // sim.resume = <checkpoint file>: spins, iteration and (unless sim.seed
  // names it) the thermostat seed of a run to continue
  std::string resume;
  CheckpointHeader ck;
  if (cfg.exists("sim.resume")) {
    resume = cfg.lookup("sim.resume").as_string();
    ck = read_checkpoint_header(resume);
  }
  if (cfg.exists("sim.seed")) {
    s->seed = (uint64_t)cfg.lookup("sim.seed").as_int64();
    if (!resume.empty() && ck.has_seed && ck.seed != s->seed)
      throw ConfigError("sim.seed differs from the seed of the checkpoint in sim.resume; the "
                        "thermostat stream would not continue");
  } else if (!resume.empty() && ck.has_seed) {
    s->seed = ck.seed;
  } else {
    std::random_device rd;
    s->seed = ((uint64_t)rd() << 32) ^ rd();
  }
  std::cout << "seed    " << s->seed << "\n";
  const Setting &solver_s = cfg.lookup("solver");
  const bool gilbert = solver_s.get<bool>("gilbert_prefactor", false);

  std::cout << "init lattice\n";
  s->lattice.init_from_config(cfg, gilbert, s->seed);
  const int64_t n = s->lattice.num_sites();
  std::cout << "  computed lattice positions " << n << "\n";
  if (!resume.empty()) {
    if (ck.num_spins >= 0 && ck.num_spins != n)
      throw ConfigError("sim.resume: the checkpoint holds " + std::to_string(ck.num_spins) +
                        " spins, the lattice " + std::to_string(n));
    s->lattice.spins = load_spins_tsv(resume, n);
    std::cout << "  resume from " << resume << " at iteration " << ck.iteration << "\n";
  }

  std::cout << "init solver\n";
  s->solver = Solver::create(solver_s);
  s->solver->initialize(solver_s);
  s->solver->set_seed(s->seed);
  // physics (core/physics.cc): the empty module with a constant temperature
  const Setting &phys = cfg.lookup("physics");
  const std::string pm = phys.get<std::string>("module", "empty");
  if (lowercase(pm) != "empty")
    throw RuntimeError("physics module " + pm + " is not supported by this build (empty only)");
  Vec3 applied = {0, 0, 0};
  if (phys.exists("applied_field"))
    for (int q = 0; q < 3; ++q) applied[q] = phys["applied_field"][q].as_double();
  s->solver->set_physics(phys["temperature"].as_double(), applied);

  if (device >= 0) {
    engine_check(nullptr, jh_create(device, &s->ctx));
    engine_check(s->ctx, jh_set_sites(s->ctx, n, s->lattice.mus.data(), s->lattice.gyro.data(),
                                      s->lattice.alpha.data()));
    const int32_t lsize[3] = {s->lattice.size(0), s->lattice.size(1), s->lattice.size(2)};
    const int32_t lper[3] = {s->lattice.periodic(0) ? 1 : 0, s->lattice.periodic(1) ? 1 : 0,
                             s->lattice.periodic(2) ? 1 : 0};
    engine_check(s->ctx, jh_set_lattice(s->ctx, lsize, lper, s->lattice.num_basis_sites()));
    engine_check(s->ctx, jh_set_time_step(s->ctx, s->solver->time_step()));
    engine_check(s->ctx, jh_set_spins(s->ctx, s->lattice.spins.data()));
    s->solver->attach(s->ctx);
  }

  std::cout << "init hamiltonians\n";
  if (!cfg.exists("hamiltonians")) throw ConfigError("No hamiltonians group in config");
  const Setting &hs = cfg.lookup("hamiltonians");
  for (int q = 0; q < hs.length(); ++q) {
    auto h = Hamiltonian::create(hs[q], s->lattice);
    std::cout << "  " << h->name() << " hamiltonian\n";
    h->register_term(s->ctx);
    s->solver->register_hamiltonian(std::move(h));
  }
  if (s->ctx) {
    engine_check(s->ctx, jh_set_thermostat(s->ctx, s->solver->temperature(), s->seed,
                                           s->solver->gilbert_prefactor() ? 1 : 0));
    // solver.precision selects the engine's precision form (an addition:
    // the reference has FP64 only)
    const std::string p = lowercase(solver_s.get<std::string>("precision", "fp64"));
    if (p != "fp64" && p != "fp32-field")
      throw ConfigError("solver.precision must be \"fp64\" or \"fp32-field\"");
    engine_check(s->ctx, jh_set_precision(s->ctx, p == "fp32-field" ? JH_PREC_FP32_FIELD : JH_PREC_FP64));
    engine_check(s->ctx, jh_set_integrator(s->ctx, s->solver->integrator()));
  }
  if (!resume.empty()) s->solver->resume_at(ck.iteration);
  if (cfg.exists("monitors") && s->ctx) {
    const Setting &ms = cfg.lookup("monitors");
    for (int q = 0; q < ms.length(); ++q)
      s->solver->register_monitor(Monitor::create(ms[q], s->prefix, *s->solver, s->lattice));
  }
}

void run(jams_sim *s, int64_t limit) {
  if (!s->ctx) throw RuntimeError("set-up-only simulation cannot run");
  Solver &solver = *s->solver;
  const int64_t end = limit >= 0 ? std::min(solver.max_steps(), solver.iteration() + limit)
                                 : solver.max_steps();
  const auto t0 = std::chrono::steady_clock::now();
  // run_simulation (core/jams++.cc:333-347): monitors are notified before
  // each step; steps with no monitor update in between are handed to the
  // engine as one batch, which changes nothing in the trajectory.
  while (solver.iteration() < end) {
    solver.notify_monitors();
    int64_t next = end;
    for (auto &m : solver.monitors()) {
      const int64_t k = m->output_steps();
      const int64_t up = (solver.iteration() / k + 1) * k;
      next = std::min(next, up);
    }
    solver.run_steps(next - solver.iteration());
  }
  if (solver.iteration() == solver.max_steps()) solver.notify_monitors();
  for (auto &m : solver.monitors()) m->finish(solver);
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cout << "runtime " << sec << " seconds\n";
}

}  // namespace

extern "C" {

int jams_sim_create(const char *const *sources, int nsources, const char *output_dir,
                    const char *name, int device, jams_sim **out) {
  if (!out || !sources || nsources <= 0) return 1;
  *out = nullptr;
  auto s = std::make_unique<jams_sim>();
  std::string dir = output_dir ? output_dir : ".";
  if (!dir.empty() && dir.back() != '/') dir += '/';
  s->prefix = dir + (name ? name : "jams") + "_";
  std::vector<std::string> src(sources, sources + nsources);
  try {
    initialize(s.get(), src, device);
  } catch (const std::exception &e) {
    g_error = e.what();
    return 1;
  }
  *out = s.release();
  return 0;
}

int jams_sim_destroy(jams_sim *sim) {
  delete sim;
  return 0;
}

const char *jams_sim_last_error(const jams_sim *sim) { return sim ? sim->err.c_str() : g_error.c_str(); }

int64_t jams_sim_num_spins(const jams_sim *sim) { return sim ? sim->lattice.num_sites() : 0; }

int64_t jams_sim_max_steps(const jams_sim *sim) { return sim ? sim->solver->max_steps() : 0; }

jh_ctx *jams_sim_engine(jams_sim *sim) { return sim ? sim->ctx : nullptr; }

int jams_sim_run(jams_sim *sim, int64_t limit) {
  return guarded(sim, [&] { run(sim, limit); });
}

int jams_sim_sites(const jams_sim *sim, double *mus, double *gyro, double *alpha, double *spins) {
  return guarded(const_cast<jams_sim *>(sim), [&] {
    const Lattice &l = sim->lattice;
    const size_t n = (size_t)l.num_sites();
    if (mus) std::memcpy(mus, l.mus.data(), n * 8);
    if (gyro) std::memcpy(gyro, l.gyro.data(), n * 8);
    if (alpha) std::memcpy(alpha, l.alpha.data(), n * 8);
    if (spins) std::memcpy(spins, l.spins.data(), 3 * n * 8);
  });
}

int jams_sim_exchange_counts(const jams_sim *sim, int index, int64_t *computed, int64_t *npairs,
                             int64_t *nnz) {
  return guarded(const_cast<jams_sim *>(sim), [&] {
    ExchangeHamiltonian *x = exchange_term(sim, index);
    if (computed) *computed = x->computed_interactions();
    if (npairs) *npairs = x->neighbours().size();
    if (nnz) {
      std::vector<int64_t> rp;
      std::vector<int32_t> c;
      std::vector<double> v;
      x->build_csr(rp, c, v);
      *nnz = rp.back();
    }
  });
}

int jams_sim_exchange_pairs(const jams_sim *sim, int index, int32_t *pi, int32_t *pj, double *J9) {
  return guarded(const_cast<jams_sim *>(sim), [&] {
    const NeighbourList &nl = exchange_term(sim, index)->neighbours();
    for (int64_t p = 0; p < nl.size(); ++p) {
      if (pi) pi[p] = nl.i[p];
      if (pj) pj[p] = nl.j[p];
      if (J9)
        for (int q = 0; q < 9; ++q) J9[9 * p + q] = nl.values[nl.v[p]][q / 3][q % 3];
    }
  });
}

int jams_sim_exchange_values(const jams_sim *sim, int index, int32_t *vidx, double *values9,
                             int64_t *nvalues) {
  return guarded(const_cast<jams_sim *>(sim), [&] {
    const NeighbourList &nl = exchange_term(sim, index)->neighbours();
    if (nvalues) *nvalues = (int64_t)nl.values.size();
    if (vidx) std::memcpy(vidx, nl.v.data(), sizeof(int32_t) * nl.size());
    if (values9)
      for (size_t v = 0; v < nl.values.size(); ++v)
        for (int q = 0; q < 9; ++q) values9[9 * v + q] = nl.values[v][q / 3][q % 3];
  });
}

int jams_sim_exchange_csr(const jams_sim *sim, int index, int64_t *rowptr, int32_t *col, double *val) {
  return guarded(const_cast<jams_sim *>(sim), [&] {
    std::vector<int64_t> rp;
    std::vector<int32_t> c;
    std::vector<double> v;
    exchange_term(sim, index)->build_csr(rp, c, v);
    if (rowptr) std::memcpy(rowptr, rp.data(), rp.size() * 8);
    if (col) std::memcpy(col, c.data(), c.size() * 4);
    if (val) std::memcpy(val, v.data(), v.size() * 8);
  });
}

int jams_sim_iteration(const jams_sim *sim, int64_t *iteration, double *time_ps) {
  if (!sim) return 1;
  if (iteration) *iteration = sim->solver->iteration();
  if (time_ps) *time_ps = sim->solver->time();
  return 0;
}

int jams_sim_write_checkpoint(jams_sim *sim, const char *path) {
  return guarded(sim, [&] {
    if (!sim->ctx) throw RuntimeError("set-up-only simulation has no state to write");
    if (!path) throw RuntimeError("checkpoint path is null");
    const Solver &so = *sim->solver;
    write_checkpoint(sim->ctx, sim->lattice.num_sites(), path, so.iteration(), so.time(), sim->seed);
  });
}

int jams_sim_params(const jams_sim *sim, double *dt, double *T, uint64_t *seed) {
  if (!sim) return 1;
  if (dt) *dt = sim->solver->time_step();
  if (T) *T = sim->solver->temperature();
  if (seed) *seed = sim->seed;
  return 0;
}

}  // extern "C"
