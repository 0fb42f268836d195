// Device-side tables and kernel launchers (implemented in lmc_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lmc_model.h"

namespace lmc {

// Packed table entries (all slots < 64, 63 = none):
//   offset   : int8 x | int8 y << 8 | int8 z << 16
//   state    : other1[0:8] | other2[8:16] | byte offset of the reduced weight block in wd
//              [16:29] | n_endpoint[31]
//   orbit    : s0[0:8] | s1[8:16] | feature base[16:28] | kind[28:30]
//   mm2      : the forward orbit entries, then the backward ones (slots mirrored)
//   site     : label[0:3] | other1[4:10] | other2[10:16]
constexpr int kMaxStateClusters = 512;
constexpr int kMaxOrbitClusters = 768;
// Cluster lists are padded to multiples of this (4 entries per lane of an event's 21 lanes).
constexpr int kClusterPad = 84;
// Copies for the wide kernel shape: 4 entries per lane of an event's 64 lanes.
constexpr int kClusterPadWide = 256;
constexpr int kMaxSiteClusters = 256;
constexpr int kMaxW = 1024;  // w1 | w2 | w3 for S <= 6

// Kernel view of a model (passed by value).
struct KTables {
  const int32_t *off;       // [12][60]
  const uint8_t *mirror;    // [60]
  const uint32_t *state;    // one-endpoint state clusters
  const uint32_t *mmm;
  const uint32_t *mm2;
  const uint32_t *state_w, *mmm_w, *mm2_w;  // the same lists padded to kClusterPadWide
  const uint32_t *site_cl;
  const int32_t *site_off;  // [43]
  const double *w;          // theta/N by (label, codes): w1[S] | w2[3][S][S] | w3[4][S][S][S]
  // jump dE over element ranks: [2 ne][wrow] w(.|M) - w(.|X) by the migrating species' rank,
  // rows ne .. 2 ne - 1 negated; wrow = 2 + 3 ne + 4 ne^2 (last: 0), nwd entries
  const double *wd;
  const double *b1;         // [ne][F1] expanded barrier weights (last: 0), row = rank of the migrating species
  const double *b2;         // [ne][F2]
  const double *par;        // [S][6]
  const double *theta;      // [n_eci]
  const double *norm_tot;   // [n_eci]
  const int32_t *eci1;      // [S]
  const int32_t *eci2;      // [3][S][S]
  const int32_t *eci3;      // [4][S][S][S]
  // Solvent-baseline formulation: for every (solvent s, migrating species M), the sums (dE, MMM
  // logit, MM2 logit) of an environment that is all solvent (base), and per environment slot j
  // the clusters containing j (CSR inc_off[3][61] -> inc, lists: state, MMM, MM2 forward and
  // backward); a real environment adds value - value(all solvent) for each cluster touching a
  // non-solvent slot, once (at its lowest such slot).
  //   state entry: reduced weight-block byte offset [0:13] | j is the second other site [13]
  //                | other site [14:20] (63 = none) | neighbour-endpoint side [31]
  //   orbit entry: expanded-weight byte offset [0:16] | j takes the rank*8 role [16]
  //                | other site [17:23] (63 = none)
  const uint16_t *inc_off;  // [3][61]
  const uint32_t *inc;
  const double *base;       // [S][S][3]
  int n_inc, solvent_default;
  uint64_t vac_shell_mask;  // env slots within 3NN of the vacancy
  int n_state, n_mmm, n_mm2, n_site, n_eci;  // n_mm2 counts forward + backward entries
  int n_state_v;            // state entries of the vacancy endpoint (the rest: neighbour's)
  int n_state_w, n_mmm_w, n_mm2_w;
  int v_slot, n_slot;
  int kind, S, ne, x_code, F1, F2, nw;
  int nwd, wrow;
  uint32_t rank_lut;        // species code -> element rank, 4 bits per code (vacancy: 0)
};

struct DeviceTables {
  int device = -1;
  void *blob = nullptr;  // one allocation holding every table
  KTables k{};
};

// Uploads (once) and returns the kernel view of a model on the current device. Thread-safe, as
// are the two queries below.
const KTables &Upload(Model &m);
bool IsUploaded(const Model &m);
// The kernel view when the model is already on `device`, else null (never uploads).
const KTables *UploadedOn(const Model &m, int device);

// Per-replica KMC state, structure of arrays in device memory.
// One line of kmc_log.txt (KineticMcFirstAbstract::Dump, KineticMcAbstract.cpp:59-104), taken
// after the step's event is chosen and before it is applied.
struct KmcLogRecord {
  uint64_t steps;
  int64_t atom;  // the migrating atom (-1 without atom tracking)
  double time, temperature, energy, ea, de;
  double vac[3];  // unwrapped Cartesian position of the vacancy (A)
  double com[3];  // solute centre-of-mass trajectory
};

struct KmcDev {
  int32_t R = 0;
  uint8_t *occ = nullptr;       // concatenated occupancy
  int64_t *off = nullptr;       // [R] occupancy offsets
  int32_t *ncell = nullptr;     // [R]
  int64_t *vac = nullptr;       // [R]
  uint64_t *steps = nullptr;    // [R]
  uint64_t *max_steps = nullptr;
  uint64_t *seed = nullptr;
  uint32_t *rid = nullptr;      // Philox replica id
  double *time = nullptr, *energy = nullptr, *abs_energy = nullptr, *temp = nullptr;
  double *cv = nullptr, *cs = nullptr, *solute_mass = nullptr;
  int32_t *flags = nullptr;     // bit0 rate corrector, bit1 early stop, bit2 solute disp
  int32_t *solvent = nullptr;   // species code of the solvent
  int32_t *escaped = nullptr;
  // [R] the escape check's moment (KineticMcAbstract.cpp:210-218): the time and step it found
  // the vacancy's 1-3NN shells all solvent, and the vacancy's site before that step's jump
  double *esc_time = nullptr;
  uint64_t *esc_step = nullptr;
  int64_t *esc_site = nullptr;
  int64_t *disp = nullptr;      // [R][3] vacancy displacement, a/2 units
  double *com = nullptr;        // [R][3] solute centre-of-mass trajectory
  int32_t *error = nullptr;     // [1] first error code
  int64_t *prev_j = nullptr;    // [R] chain steps: the site the vacancy came from
  // optional atom tracking: atom id on every site (ids local to the replica) and map_shift
  // [site offset + atom id][3]
  int32_t *atom = nullptr, *shift = nullptr;
  // optional log records [R][log_cap] at the reference's dump cadence of log_every
  uint64_t log_every = 0, log_cap = 0;
  uint64_t *log_n = nullptr;      // [R] records taken since the last drain
  KmcLogRecord *log = nullptr;
  double *log_edge = nullptr;     // [R] cubic box edge (A)
  // optional temperature schedule (shared)
  int32_t n_tt = 0;
  double *tt_time = nullptr, *tt_temp = nullptr;
  // optional trace [R][cap]
  uint64_t trace_cap = 0;
  int64_t *tr_sel = nullptr;
  double *tr_ea = nullptr, *tr_de = nullptr, *tr_time = nullptr;
  double mass[kMaxSpecies] = {};  // element masses by species code
};

// SimulatedAnnealing's adaptive schedule state of one replica
// (lmc/mc/include/SimulatedAnnealing.h:31-71).
struct CmcAnneal {
  double T, lowest, recent_best;
  uint64_t last_improvement, last_reheat;
  uint32_t window_trials, window_accepts, reheats, pad;
};

// One line of cmc_log.txt (CanonicalMcAbstract::Dump, CanonicalMcAbstract.cpp:53-84) or of the
// simulated-annealing log (SimulatedAnnealing::Dump, SimulatedAnnealing.cpp:75-97): the step, the
// temperature, the energy before the step's trial, `second` (canonical MC: the thermodynamic
// average over the averaging window, filled in after the launch; annealing: the lowest energy)
// and the absolute energy. `count`: energies in the averaging ring up to this step.
struct CmcLogRecord {
  uint64_t steps;
  double temperature, energy, second, abs_energy;
  uint64_t count;
};

struct CmcDev {
  int32_t R = 0;
  int64_t max_sites = 0;   // largest replica, in sites
  int lds_occupancy = 1;   // hold each replica's occupancy in LDS while it runs, where it fits
  int32_t *atom = nullptr, *shift = nullptr;  // optional atom tracking, as in KmcDev
  CmcAnneal *anneal = nullptr;  // [R]; null: the linear ramp t0 -> t1
  uint8_t *occ = nullptr;
  int64_t *off = nullptr;
  int32_t *ncell = nullptr;
  uint64_t *steps = nullptr, *max_steps = nullptr, *seed = nullptr, *accepted = nullptr;
  uint32_t *rid = nullptr;
  double *t0 = nullptr, *t1 = nullptr, *energy = nullptr, *abs_energy = nullptr;
  int32_t *error = nullptr;
  uint64_t trace_cap = 0;
  int64_t *tr_p = nullptr, *tr_q = nullptr;
  double *tr_de = nullptr;
  int8_t *tr_acc = nullptr;
  // optional log: records on the Dump cadence (log_every > 0), log_n counts past log_cap;
  // log_lo: the replica's record count when the current launch began
  uint64_t log_every = 0, log_cap = 0;
  uint64_t *log_n = nullptr, *log_lo = nullptr;
  CmcLogRecord *log = nullptr;
  // ThermodynamicAveraging over the last avg_m energies: ring [R][ring_cap] of the energy at
  // every step (ring_cap >= avg_m + trials per launch), ring_n energies written per replica
  uint64_t avg_m = 0, ring_cap = 0;
  double *ring = nullptr;
  uint64_t *ring_n = nullptr;
  // CanonicalMcOmp rounds (omp_k > 0 trials per round; 0: serial trials): per replica the
  // round's pairs and dE [R][omp_k], its length and the next event to try
  int32_t omp_k = 0;
  int64_t *omp_p = nullptr, *omp_q = nullptr;
  double *omp_de = nullptr;
  int32_t *omp_len = nullptr, *omp_pos = nullptr;
};

// env_out (optional): [12 R][64] environment rows of the events in rank order; ekra_out
// (optional): [12 R] kinetically resolved barrier (E0 model: E0; quartic: Ea - dE/2).
hipError_t LaunchEvalJumps(const KTables &t, const uint8_t *occ, const int64_t *off,
                           const int32_t *ncell, const int64_t *vac, int32_t R, int64_t *nb,
                           double *ea, double *de, int32_t *err, hipStream_t s, bool solvent,
                           uint8_t *env_out = nullptr, double *ekra_out = nullptr);
// The 12 jumps of a vacancy placed at each of sites[K] in the single occupancy occ (n^3 cells),
// with site sub_site holding species sub_code instead (the ExitTime analysis' configuration);
// outputs [K][12] in neighbour-id order.
hipError_t LaunchVacancyScan(const KTables &t, const uint8_t *occ, int32_t n, int64_t sub_site,
                             int sub_code, const int64_t *sites, int32_t K, int64_t *nb,
                             double *ea, double *de, int32_t *err, hipStream_t s);
// form: 0 dense (256 threads per replica), 1 solvent baseline, 2 dense wide shape (768 threads
// per replica, one workgroup per CU), 3 second-order chain steps (KineticMcChainOmpi).
hipError_t LaunchKmc(const KTables &t, const KmcDev &k, uint64_t steps, uint64_t trace_base,
                     hipStream_t s, int form);
// Everything LaunchKmc sets up outside the stream (LDS plan, kernel attributes), so that its
// launches can then be captured in a graph.
hipError_t PrepareKmc(const KTables &t, int32_t R, int form);
hipError_t LaunchSwapDe(const KTables &t, const uint8_t *occ, const int64_t *off,
                        const int32_t *ncell, const int64_t *pairs, int32_t R, double *de,
                        hipStream_t s);
// from: optional species the sites change from (null: the occupancy's).
hipError_t LaunchSiteDe(const KTables &t, const uint8_t *occ, const int64_t *off,
                        const int32_t *ncell, const int64_t *sites, const uint8_t *from,
                        const uint8_t *codes, int32_t R, double *de, hipStream_t s);
// The thermodynamic averages of the canonical-MC log records of the last launch
// (ThermodynamicAveraging::GetThermodynamicAverage, ThermodynamicAveraging.cpp:25-40).
hipError_t LaunchCmcAverage(const CmcDev &c, hipStream_t s);
hipError_t LaunchCmc(const KTables &t, const CmcDev &c, uint64_t trials, uint64_t trace_base,
                     hipStream_t s);
// GetEncode of every replica: encode[R][n_eci] and energy[R] (either may be null).
hipError_t LaunchTotalEncode(const KTables &t, const uint8_t *occ, const int64_t *off,
                             const int32_t *ncell, int32_t R, unsigned long long *counts,
                             double *encode, double *energy, hipStream_t s);
// GetEncodeOfCluster over the sorted unique sites set[m] of one n^3-cell occupancy.
hipError_t LaunchClusterEncode(const KTables &t, const uint8_t *occ, int32_t n, const int64_t *set,
                               int32_t m, unsigned long long *counts, double *encode,
                               double *energy, hipStream_t s);
hipError_t LaunchTotalEnergy(const KTables &t, const uint8_t *occ, const int64_t *off,
                             const int32_t *ncell, int32_t R, unsigned long long *counts,
                             double *energy, hipStream_t s);
// The per-device address of the fixed cluster-count scratch array of lmc_total_energy_noalloc
// (looked up once per device) and its capacity in counts.
hipError_t TotalEnergyScratch(unsigned long long **ptr);
size_t TotalEnergyScratchCapacity();

// Limits of the current device that the launch plans read: compute units, LDS per CU and LDS
// one workgroup may use (bytes). Queried once per device under a lock, so that launches from
// several host threads or on several devices each see their own device's values.
struct DeviceLimits {
  int cus;
  size_t lds_cu, lds_block;
};
const DeviceLimits &CurrentDeviceLimits();

}  // namespace lmc
