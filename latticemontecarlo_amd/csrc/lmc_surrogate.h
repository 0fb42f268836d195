// Host-side launchers of the surrogate kernels (lmc_surrogate.hip, lmc_surrogate_kmc.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lmc_device.h"
#include "lmc_mlp.h"

namespace lmc {

// Mutable view of an ensemble (training).
struct MlpMut {
  int M, C, Kp;
  MlpLayout lay;
  float *master, *adam_m, *adam_v;
  uint16_t *w1t, *w2t, *w2k, *w3, *w1f, *w2f;
};

struct AdamStep {
  float lr, beta1, beta2, eps, corr1, corr2;  // corr = 1 / (1 - beta^t)
};

// Fused forward + backward of every member over rows row0 .. row0+n-1 of `order` (or of env
// when order is null); per-64-row-tile gradient partials [tiles][M][P] and loss partials.
hipError_t LaunchMlpTrain(const MlpDev &w, const uint8_t *env, const int32_t *order, int row0,
                          int n, const float *tgt, float *grad_part, float *loss_part,
                          hipStream_t s);
// Normalised outputs [n][M][2] of every member.
hipError_t LaunchMlpForward(const MlpDev &w, const uint8_t *env, int n, float *out, hipStream_t s);
// Sum of `nparts` partials (tile order) -> grad_out (optional); Adam when update; bf16 copies.
hipError_t LaunchMlpReduceAdam(const MlpMut &w, const float *grad_part, int nparts,
                               float *grad_out, int update, const AdamStep &a, hipStream_t s);
int MlpTileRows();

// Surrogate in the KMC step (lmc_surrogate_kmc.hip).
struct SurDev {
  MlpDev w;
  double mu[2], sigma[2];
  int log_ekra;      // output 0 is ln E_kra
  int base_code;     // species of the tied-to-zero input channel (the solvent)
  int max_solutes;   // environments with more non-base, non-vacancy slots fall back
  double threshold;  // member disagreement (eV) above which a relevant event falls back; < 0: all
  double window;     // relevance window in kT (<= 0: every event is relevant)
  uint64_t *n_sur, *n_fb;  // per-replica counters (optional)
};
struct SurOut {  // stateless evaluation, event (rank) order
  int64_t *nb;
  double *ea, *de;
  int8_t *flag;
};
hipError_t LaunchSurrogateKmc(const KTables &t, const KmcDev &k, const SurDev &sd, uint64_t steps,
                              uint64_t trace_base, hipStream_t s);
hipError_t PrepareSurrogateKmc(const KTables &t, int C);
hipError_t LaunchSurrogateEval(const KTables &t, const KmcDev &k, const SurDev &sd,
                               const SurOut &out, hipStream_t s);

}  // namespace lmc
