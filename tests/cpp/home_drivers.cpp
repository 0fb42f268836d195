// A host program written the way the reference's api::Build*FromParameter functions build and
// run their drivers (lmc/api/src/Home.cpp:159-189, 240-292): elements as std::set<Element>, the
// configuration read with ReadConfig + ReassignLatticeVector, the driver brace-initialised from
// the parameter values, then Simulate(). The only additions are SetSeed (a reproducible Philox
// stream in place of the reference's clock seed) and the final-state line printed for the test.
//
//   home_drivers kmc|chain <config> <json> <log> <cfgdump> <max> <restart_steps> <restart_energy>
//                <restart_time> <T> <time_temperature_file|-> <rate_corrector> <early_stop>
//                <solute_disp> <seed>
//   home_drivers cmc|cmc_omp <config> <json> <log> <cfgdump> <max> <thermo> <restart_steps>
//                <restart_energy> <T> <seed>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "lmc.hpp"

namespace {
struct Parameter {  // the fields of api::Parameter the builders read
  std::string config_filename_, json_coefficients_filename_, time_temperature_filename_;
  std::vector<std::string> element_set_{"Al", "Mg", "Zn"};
  unsigned long long log_dump_steps_ = 0, config_dump_steps_ = 0, maximum_steps_ = 0,
                     thermodynamic_averaging_steps_ = 0, restart_steps_ = 0;
  double restart_energy_ = 0, restart_time_ = 0, temperature_ = 0;
  bool rate_corrector_ = false, early_stop_ = false, solute_disp_ = false;
};

template <typename Driver>
Driver BuildKineticMcFromParameter(const Parameter &parameter) {
  std::set<Element> element_set;
  for (const auto &element_string : parameter.element_set_) {
    element_set.insert(Element(element_string));
  }
  cfg::Config config;
  config = cfg::Config::ReadConfig(parameter.config_filename_);
  config.ReassignLatticeVector();
  return Driver{config,
                parameter.log_dump_steps_,
                parameter.config_dump_steps_,
                parameter.maximum_steps_,
                parameter.thermodynamic_averaging_steps_,
                parameter.restart_steps_,
                parameter.restart_energy_,
                parameter.restart_time_,
                parameter.temperature_,
                element_set,
                parameter.json_coefficients_filename_,
                parameter.time_temperature_filename_,
                parameter.rate_corrector_,
                parameter.early_stop_,
                parameter.solute_disp_};
}

template <typename Driver>
Driver BuildCanonicalMcFromParameter(const Parameter &parameter) {
  std::set<Element> element_set;
  for (const auto &element_string : parameter.element_set_) {
    element_set.insert(Element(element_string));
  }
  cfg::Config config;
  config = cfg::Config::ReadConfig(parameter.config_filename_);
  config.ReassignLatticeVector();
  return Driver{config,
                parameter.log_dump_steps_,
                parameter.config_dump_steps_,
                parameter.maximum_steps_,
                parameter.thermodynamic_averaging_steps_,
                parameter.restart_steps_,
                parameter.restart_energy_,
                parameter.temperature_,
                element_set,
                parameter.json_coefficients_filename_};
}

unsigned long long U(const char *s) { return std::strtoull(s, nullptr, 10); }
double D(const char *s) { return std::strtod(s, nullptr); }
}  // namespace

int main(int argc, char **argv) {
  try {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    Parameter p;
    if ((mode == "kmc" || mode == "chain") && argc == 17) {
      p.config_filename_ = argv[2];
      p.json_coefficients_filename_ = argv[3];
      p.log_dump_steps_ = U(argv[4]);
      p.config_dump_steps_ = U(argv[5]);
      p.maximum_steps_ = U(argv[6]);
      p.restart_steps_ = U(argv[7]);
      p.restart_energy_ = D(argv[8]);
      p.restart_time_ = D(argv[9]);
      p.temperature_ = D(argv[10]);
      p.time_temperature_filename_ = std::string(argv[11]) == "-" ? "" : argv[11];
      p.rate_corrector_ = U(argv[12]) != 0;
      p.early_stop_ = U(argv[13]) != 0;
      p.solute_disp_ = U(argv[14]) != 0;
      const uint64_t seed = U(argv[15]);
      (void)argv[16];
      auto run = [&](auto &&driver) {
        driver.SetSeed(seed);
        driver.Simulate();
        const auto &st = driver.GetState();
        std::printf("final steps=%llu vacancy=%lld time=%.17g energy=%.17g abs_energy=%.17g\n",
                    static_cast<unsigned long long>(st.steps[0]), static_cast<long long>(st.vacancy[0]),
                    st.time[0], st.energy[0], st.absolute_energy[0]);
      };
      if (mode == "kmc")
        run(BuildKineticMcFromParameter<mc::KineticMcFirstOmp>(p));
      else
        run(BuildKineticMcFromParameter<mc::KineticMcChainOmpi>(p));
      return 0;
    }
    if ((mode == "cmc" || mode == "cmc_omp") && argc == 12) {
      p.config_filename_ = argv[2];
      p.json_coefficients_filename_ = argv[3];
      p.log_dump_steps_ = U(argv[4]);
      p.config_dump_steps_ = U(argv[5]);
      p.maximum_steps_ = U(argv[6]);
      p.thermodynamic_averaging_steps_ = U(argv[7]);
      p.restart_steps_ = U(argv[8]);
      p.restart_energy_ = D(argv[9]);
      p.temperature_ = D(argv[10]);
      auto run = [&](auto &&driver) {
        driver.SetSeed(U(argv[11]));
        driver.Simulate();
        std::printf("final energy=%.17g\n", driver.GetEnergy());
      };
      if (mode == "cmc")
        run(BuildCanonicalMcFromParameter<mc::CanonicalMcSerial>(p));
      else
        run(BuildCanonicalMcFromParameter<mc::CanonicalMcOmp>(p));
      return 0;
    }
    std::cerr << "bad arguments\n";
    return 2;
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
}
