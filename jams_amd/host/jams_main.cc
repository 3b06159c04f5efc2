// jams_main.cc -- command-line driver (main.cc, core/args.cc:13-60 of the
// reference): jams [--output=DIR] [--name=NAME] [--setup-only] [--version]
//                  config.cfg ['patch string' ...]
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/jams_host.h"

int main(int argc, char **argv) {
  std::vector<std::string> sources;
  std::string output = ".", name;
  bool setup_only = false;
  for (int a = 1; a < argc; ++a) {
    std::string s = argv[a];
    if (s == "--version") {
      std::printf("jams (gfx950 Heun-LLG engine) %d\n", jh_version());
      return 0;
    } else if (s == "--setup-only") {
      setup_only = true;
    } else if (s.rfind("--output=", 0) == 0) {
      output = s.substr(9);
    } else if (s.rfind("--name=", 0) == 0) {
      name = s.substr(7);
    } else if (s.rfind("--", 0) == 0) {
      std::fprintf(stderr, "RUNTIME ERROR: Unknown flag '%s'\n", s.c_str());
      return 1;
    } else {
      sources.push_back(s);
    }
  }
  if (sources.empty()) {
    std::fprintf(stderr, "RUNTIME ERROR: No config file specified\n");
    return 1;
  }
  if (name.empty()) {
    // the simulation is named after the first config file (core/jams++.cc:188-204)
    std::string f = sources[0];
    size_t sl = f.find_last_of('/');
    if (sl != std::string::npos) f = f.substr(sl + 1);
    size_t dot = f.find_last_of('.');
    name = dot == std::string::npos ? f : f.substr(0, dot);
  }
  std::vector<const char *> src;
  for (auto &s : sources) src.push_back(s.c_str());
  jams_sim *sim = nullptr;
  if (jams_sim_create(src.data(), (int)src.size(), output.c_str(), name.c_str(),
                      setup_only ? -1 : 0, &sim) != 0) {
    std::fprintf(stderr, "ERROR: %s\n", jams_sim_last_error(nullptr));
    return 1;
  }
  int rc = 0;
  if (!setup_only && jams_sim_run(sim, -1) != 0) {
    std::fprintf(stderr, "ERROR: %s\n", jams_sim_last_error(sim));
    rc = 1;
  }
  jams_sim_destroy(sim);
  return rc;
}
