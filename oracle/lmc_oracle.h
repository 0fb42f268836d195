// CPU oracle for the LatticeMC event-evaluation path.
//
// TEST INFRASTRUCTURE ONLY. This library restates, in plain C++ and FP64, the algorithms of
// zhucongx/LatticeMonteCarlo that the GPU engine replaces. It exists to check the engine: only
// tests, the smoke check and the CPU-baseline leg of bench.py load it. The engine never links,
// calls or falls back to it.
//
// Occupancy convention shared with the engine's C ABI: one uint8 per lattice site, value =
// index of the species in sorted(element_set + {"X"}) by element string (e.g. Al=0 Mg=1 X=2 Zn=3).
// Lattice ids follow the reference convention (lmc/cfg/src/Config.cpp:466-470: x-major sort of
// fractional positions).
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Lattice: GenerateFCC (Config.cpp:1060-1095) + UpdateNeighbors (:955-1045) +
// ReassignLatticeVector (:466-552), cubic n x n x n conventional cells.
void *orc_lattice_new(int n);
void orc_lattice_free(void *lat);
int orc_lattice_num_sites(const void *lat);
// shell 1/2/3 -> neighbour ids in ascending id order (Config.cpp:1037-1044); returns count.
int orc_lattice_neighbors(const void *lat, int site, int shell, int32_t *out);
void orc_lattice_position(const void *lat, int site, double *rel3);

// Model kinds.
enum { ORC_ENERGY = 0, ORC_E0 = 1, ORC_QUARTIC = 2, ORC_PAIRSITE = 3 };
// Load JSON coefficients (reference schema) for the given element set (comma separated, no X).
// Returns NULL on error and writes a message into err.
void *orc_model_new(const void *lat, const char *json_path, const char *elements_csv, int kind,
                    char *err, int errlen);
void orc_model_free(void *model);
int orc_model_num_eci(const void *model);
int orc_model_num_features(const void *model, int which /*0=MMM,1=MM2*/);
int orc_model_num_species(const void *model);
// Species code of an element string (code convention above); -1 if absent.
int orc_model_code(const void *model, const char *element);

// VacancyMigrationPredictor{E0,Quartic}::GetBarrierAndDiffFromLatticeIdPair.
// Returns 0 on success, -1 on error (e.g. unknown cluster type).
int orc_barrier_de(const void *model, const void *lat, const uint8_t *occ, int v, int nb,
                   double *ea, double *de);
// Feature vectors of one jump (MMM for E0/Quartic D, MM2 forward/backward for Quartic Ks).
int orc_encode(const void *model, const void *lat, const uint8_t *occ, int v, int nb, int which,
               double *out);
// EnergyChangePredictorPairSite::GetDeFromLatticeIdPair / GetDeFromLatticeIdSite.
int orc_swap_de(const void *model, const void *lat, const uint8_t *occ, int p, int q, double *de);
int orc_site_de(const void *model, const void *lat, const uint8_t *occ, int site, int new_code,
                double *de);
// EnergyPredictor::GetEnergy and GetEncode.
int orc_total_energy(const void *model, const void *lat, const uint8_t *occ, double *e);
int orc_total_encode(const void *model, const void *lat, const uint8_t *occ, double *out);

// Cluster tables of one jump pair, as lattice ids, for comparison with the engine's tables.
// State clusters: records (label, s0, s1, s2) with -1 padding; returns count.
int orc_state_clusters(const void *model, const void *lat, int v, int nb, int32_t *out, int cap);
// MMM/MM2 orbit clusters: records (orbit, sym, s0, s1) with s0,s1 in the reference's sorted
// order (-1 padding for singlets); returns count.
int orc_orbit_clusters(const void *model, const void *lat, int v, int nb, int which, int32_t *out,
                       int cap);

// Philox4x32-10 (Salmon et al., SC'11) used for every random draw.
void orc_philox(uint32_t k0, uint32_t k1, const uint32_t ctr[4], uint32_t out[4]);

// KineticMcFirst driver (KineticMcAbstract.cpp:140-190, KineticMcFirstOmp.cpp:52-82).
typedef struct {
  double temperature;
  uint64_t seed;
  uint32_t replica;
  uint64_t restart_steps;
  double restart_energy;
  double restart_time;
  int rate_corrector;      // RateCorrector.hpp
  int early_stop;          // KineticMcAbstract.cpp:192-223
  int solute_disp;         // KineticMcAbstract.cpp:163-177
  int n_tt;                // time-temperature table points (0 = constant T)
  const double *tt_time;   // TimeTemperatureInterpolator.cpp:44-65
  const double *tt_temp;
  int lru_capacity;        // 0 = no memo (VacancyMigrationPredictor*Lru)
  int omp_threads;         // threads for the 12-event loop (1 = serial)
  int chain;               // 1: second-order residence-time steps (KineticMcChainOmpi.cpp:56-151)
} orc_kmc_params;

void *orc_kmc_new(const void *barrier_model, const void *energy_model, const void *lat,
                  const uint8_t *occ, const orc_kmc_params *p, char *err, int errlen);
void orc_kmc_free(void *kmc);
// Runs until `steps` steps were executed or the run ends (steps_ > maximum_steps).
// Optional per-step trace arrays (length >= steps, may be NULL): selected neighbour id, Ea, dE,
// time after the step, total rate.
int orc_kmc_run(void *kmc, uint64_t steps, uint64_t maximum_steps, int32_t *tr_sel, double *tr_ea,
                double *tr_de, double *tr_time, double *tr_rate);
// State: steps, time, energy, absolute_energy, temperature, vacancy site, vacancy
// unwrapped displacement (3 x half-lattice units), solute COM trajectory (3), escaped flag.
void orc_kmc_state(const void *kmc, uint64_t *steps, double *time, double *energy,
                   double *abs_energy, double *temperature, int32_t *vac, int64_t *disp3,
                   double *com3, int *escaped);
void orc_kmc_occupancy(const void *kmc, uint8_t *occ);
// Events of the current step (12 x {neighbour id, Ea, dE, rate, cumulative probability}).
// (chain mode: barriers, energy changes and forward rates of the events k -> i, second-order
// cumulative probabilities)
int orc_kmc_events(void *kmc, int32_t *nb, double *ea, double *de, double *rate, double *cum);
// Residence time of the current step: the second-order time t_2 in chain mode, otherwise the
// mean 1 / (nu total rate). -1 on error.
double orc_kmc_step_time(void *kmc);
uint64_t orc_kmc_lru_hits(const void *kmc);

// CanonicalMcSerial driver (CanonicalMcAbstract.cpp:43-101, CanonicalMcSerial.cpp:40-51) with an
// optional linear temperature ramp T0 -> T1 over maximum_steps.
typedef struct {
  double t_start;
  double t_end;
  uint64_t seed;
  uint32_t replica;
  uint64_t restart_steps;
  double restart_energy;
  int anneal;  // 1: SimulatedAnnealing's adaptive schedule from T = t_start (t_end unused)
} orc_cmc_params;
void *orc_cmc_new(const void *pairsite_model, const void *energy_model, const void *lat,
                  const uint8_t *occ, const orc_cmc_params *p, char *err, int errlen);
// Annealing state: current temperature, lowest energy seen at a step start, reheats done.
void orc_cmc_anneal_state(const void *cmc, double *temperature, double *lowest, uint32_t *reheats);
void orc_cmc_free(void *cmc);
// Optional per-trial trace: site pair, dE, accepted flag.
int orc_cmc_run(void *cmc, uint64_t trials, uint64_t maximum_steps, int32_t *tr_p, int32_t *tr_q,
                double *tr_de, int8_t *tr_acc);
/* CanonicalMcOmp trials: rounds of up to K swap pairs under the 3NN exclusion set, all dE on the
 * round's starting occupancy, then the Metropolis steps in order; rounds carry over calls. */
int orc_cmc_omp_run(void *cmc, uint64_t trials, uint64_t maximum_steps, int K, int32_t *tr_p,
                    int32_t *tr_q, double *tr_de, int8_t *tr_acc);
int orc_cmc_round_left(const void *cmc);
void orc_cmc_state(const void *cmc, uint64_t *steps, double *energy, double *abs_energy,
                   uint64_t *accepted);
void orc_cmc_occupancy(const void *cmc, uint8_t *occ);

// Wall-clock timing helper for the CanonicalMc / SimulatedAnnealing CPU baseline: runs
// `replicas` independent replicas of occ concurrently, one thread each, `trials` trials each;
// returns seconds of the trial loops (site tables and setup excluded).
double orc_cmc_baseline(const void *pairsite_model, const void *lat, const uint8_t *occ,
                        double t_start, uint64_t seed, int replicas, uint64_t trials,
                        uint64_t maximum_steps, int anneal);

// Wall-clock timing helper for the CPU baseline: runs `replicas` independent KMC trajectories
// (each with `threads_per_replica` OpenMP threads on the 12-event loop) concurrently for
// `steps` steps each, returns seconds of the step loops (setup excluded).
double orc_kmc_baseline(const void *barrier_model, const void *energy_model, const void *lat,
                        const uint8_t *occ, double temperature, uint64_t seed, int replicas,
                        int threads_per_replica, int lru_capacity, uint64_t steps,
                        uint64_t *lru_hits, int chain);

#ifdef __cplusplus
}
#endif
