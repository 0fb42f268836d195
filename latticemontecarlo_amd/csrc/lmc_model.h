// Coefficient model: the reference's JSON schema (lmc/pred/src/EnergyPredictor.cpp:33-37,
// VacancyMigrationPredictorE0.cpp:28-40, VacancyMigrationPredictorQuartic.cpp:44-63) loaded into
// the flat tables the GPU kernels read.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "lmc_tables.h"

namespace lmc {

enum ModelKind { kEnergy = 0, kE0 = 1, kQuartic = 2, kPairSite = 3 };

struct BarrierParams {
  bool present = false;
  std::vector<double> mu_mmm, sigma_mmm, mu_mm2, sigma_mm2;
  std::vector<std::vector<double>> U_mmm, U_mm2;
  std::vector<double> theta_a, theta_b;  // E0: theta_e0 | Quartic: theta_D, theta_Ks
  double mu_a = 0, sigma_a = 0, mu_b = 0, sigma_b = 0;
};

struct DeviceTables;  // defined in lmc_device.h

struct Model {
  int kind = kEnergy;
  std::vector<std::string> elements;  // without X, string order
  EciSpace eci;
  int S = 0, ne = 0, x_code = -1;
  std::vector<int> onehot;            // [S] rank among elements, -1 for X
  std::vector<double> theta;          // Base.theta
  std::vector<BarrierParams> bar;     // [S] by migrating species code

  // Derived tables.
  std::vector<double> w;              // theta/N per (label, codes): w1[S] | w2[3][S][S] | w3[4][S^3]
  std::vector<double> norm_total;     // per ECI index: EnergyPredictor normaliser
  std::vector<int> base_mmm, len_mmm, base_mm2, len_mm2;
  int F_mmm = 0, F_mm2 = 0;
  std::vector<double> b_mmm, b_mm2;   // [S][F]: weight of one cluster of feature f
  std::vector<double> par;            // [S][6]: c0_a, mu_a, sigma_a, c0_b, mu_b, sigma_b
  std::vector<int8_t> type_tab;       // [3][S][S]: one-hot type index (kind 0 uses row cA)

  DeviceTables *dev = nullptr;        // lazily uploaded
  int32_t *dev_err = nullptr;         // device flag: unknown cluster / missing parameters
  int formulation = 0;                // jump kernels: 0 dense, 1 solvent baseline
  ~Model();
};

// Throws std::runtime_error with the reference's messages where they exist.
Model *LoadModel(const std::string &json_path, const std::vector<std::string> &elements, int kind);

}  // namespace lmc
