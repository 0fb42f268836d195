/* C ABI of the MI355X LatticeMC event-evaluation engine.
 *
 * Plain C types only (no torch, no STL). Every function returns LMC_OK (0) or a negative
 * status; lmc_last_error() then returns a thread-local message. Pointers named d_* are device
 * pointers (hipMalloc'ed or torch CUDA tensors); `stream` is a hipStream_t (NULL = default).
 * Calls over a batch reject a negative count, and a null buffer when the count is positive,
 * before touching the device; a call over zero items does nothing. A model may be shared by
 * host threads (its first use uploads its tables once, under a lock); a driver may not.
 *
 * Occupancy format: one uint8 per lattice site, indexed by the reference's lattice id
 * (x-major rank of the fractional position, lmc/cfg/src/Config.cpp:466-470); the value is the
 * species code = index of the element in sorted(element_set + {"X"}) by element string
 * (lmc/cfg/include/Element.hpp:44-49), e.g. {Al, Mg, Zn} -> Al=0 Mg=1 X=2 Zn=3. Lattices are
 * cubic FCC boxes of n x n x n conventional cells (N = 4 n^3 sites, n >= 4), a = 4.046 A.
 * A batch of R replicas stores its occupancies back to back; d_offset[r] is the first byte of
 * replica r and d_ncell[r] its n.
 */
#ifndef LMC_CAPI_H_
#define LMC_CAPI_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LMC_OK 0
#define LMC_ERR_ARG (-1)      /* invalid argument */
#define LMC_ERR_IO (-2)       /* file missing / unreadable / malformed */
#define LMC_ERR_HIP (-3)      /* HIP runtime error (no GPU, launch failure, ...) */
#define LMC_ERR_MODEL (-4)    /* unknown cluster type or missing element parameters */

enum {
  LMC_MODEL_ENERGY = 0,   /* pred::EnergyPredictor */
  LMC_MODEL_E0 = 1,       /* pred::VacancyMigrationPredictorE0 */
  LMC_MODEL_QUARTIC = 2,  /* pred::VacancyMigrationPredictorQuartic */
  LMC_MODEL_PAIRSITE = 3  /* pred::EnergyChangePredictorPairSite */
};

typedef struct lmc_model lmc_model;
typedef struct lmc_kmc lmc_kmc;
typedef struct lmc_cmc lmc_cmc;

const char *lmc_last_error(void);
int lmc_version(void);

/* ---- Device buffers -------------------------------------------------------------------------
 * For host code that does not link the HIP runtime itself (the C++ mirror in lmc.hpp): plain
 * device allocations and synchronous copies. No reference counterpart: the reference keeps its
 * configuration in host memory (cfg::Config, lmc/cfg/include/Config.h:99-107). */
int lmc_device_alloc(uint64_t bytes, void **d_ptr);
int lmc_device_free(void *d_ptr);
int lmc_copy_to_device(void *d_dst, const void *src, uint64_t bytes);
int lmc_copy_to_host(void *dst, const void *d_src, uint64_t bytes);
int lmc_synchronize(void *stream);

/* ---- Models --------------------------------------------------------------------------------
 * Replaces the predictor constructors: EnergyPredictor(json, element_set)
 * (lmc/pred/include/EnergyPredictor.h:16-17), VacancyMigrationPredictorE0(json, reference_config,
 * element_set) (VacancyMigrationPredictorE0.h:15-17), VacancyMigrationPredictorQuartic(...)
 * (VacancyMigrationPredictorQuartic.h:18-20), EnergyChangePredictorPairSite(...)
 * (EnergyChangePredictorPairSite.h:16-18). The reference config only supplied geometry; the
 * engine's tables are independent of lattice size, so it is not needed here.
 * elements: the element set without "X". Unknown element strings become "X" (Element.hpp:25-26).
 */
int lmc_model_load(const char *json_path, const char *const *elements, int32_t n_elements,
                   int32_t kind, lmc_model **out);
void lmc_model_free(lmc_model *m);
int32_t lmc_model_num_species(const lmc_model *m);
/* Species code of an element string, -1 if not in the model. */
int32_t lmc_model_species_code(const lmc_model *m, const char *element);
/* Formulation of the jump kernels used by this model's stateless calls and as the default of
 * new KMC drivers: 0 = dense (every cluster of every jump; default), 1 = solvent baseline
 * (all-solvent sums precomputed per (solvent, migrating species); only clusters touching a
 * non-solvent site are evaluated). Both are exact; they differ in FP64 summation order. */
int lmc_model_set_formulation(lmc_model *m, int32_t formulation);
/* Element string of a species code (NULL if out of range); valid while m lives. */
const char *lmc_model_species_name(const lmc_model *m, int32_t code);
/* Sizes: number of ECIs (Base.theta), MMM and MM2 feature lengths. */
int lmc_model_layout(const lmc_model *m, int32_t *n_eci, int32_t *n_mmm, int32_t *n_mm2);
/* ECI key `index`: label (0..7) and its sorted species codes (-1 padded to 3). */
int lmc_model_eci_key(const lmc_model *m, int32_t index, int32_t *label, int32_t *codes3);
/* Orbit layout of the MMM (which=0) or MM2 (which=1) features; arrays of capacity `cap`:
 * orbit size (clusters), kind (0 singlet, 1 pair, 2 symmetric pair), shell, feature base and
 * length. Returns the number of orbits in *n_orbits. */
int lmc_model_orbits(const lmc_model *m, int32_t which, int32_t cap, int32_t *n_orbits,
                     int32_t *size, int32_t *kind, int32_t *shell, int32_t *base, int32_t *len);

/* ---- Random numbers ------------------------------------------------------------------------
 * The engine's Philox4x32-10 block for (seed, counter, replica, tag): key = seed (low, high
 * words), counter words = (counter low, counter high, replica, tag). The reference draws from a
 * clock-seeded mt19937_64 (lmc/mc/src/McAbstract.cpp:32); this replaces it so that every
 * replica's stream is reproducible and resumable. */
void lmc_philox4x32(uint64_t seed, uint64_t counter, uint32_t replica, uint32_t tag,
                    uint32_t *out4);

/* ---- Geometry (host) --------------------------------------------------------------------- */
int64_t lmc_fcc_num_sites(int32_t n);
/* Neighbour ids of `site` in shell 1/2/3, ascending (Config::UpdateNeighbors,
 * Config.cpp:955-1045). Returns the count (12/6/24) or a negative status. */
int32_t lmc_fcc_neighbors(int32_t n, int64_t site, int32_t shell, int64_t *out);
/* Cluster tables of the jump v -> nb as lattice ids (for validation):
 * state clusters as (label, s0, s1, s2) and MMM/MM2 orbit clusters as (orbit, sym, s0, s1). */
int32_t lmc_jump_state_clusters(int32_t n, int64_t v, int64_t nb, int64_t *out, int32_t cap);
/* Lattice ids of the 60 environment slots of the jump v -> nb in the engine's slot order (the
 * order of surrogate input rows); GetSortedLatticeVectorStateOfPair (EnergyUtility.cpp:261-287)
 * gives the same site set. Returns 60 or a negative status. */
int32_t lmc_jump_environment(int32_t n, int64_t v, int64_t nb, int64_t *out60);
/* The four symmetries of a jump environment that fix both endpoints (identity and the three
 * mirror/rotation operations of the bond's C2v group), as slot permutations perm[4][60]: the
 * environment with species e'[j] = e[perm[g][j]] is the image of e under operation g and has
 * the same E_kra and dE. No reference counterpart (used to augment surrogate training data). */
int lmc_jump_symmetries(int32_t *perm);
int32_t lmc_jump_orbit_clusters(int32_t n, int64_t v, int64_t nb, int32_t which, int64_t *out,
                                int32_t cap);

/* ---- Configuration files (host) -----------------------------------------------------------
 * A configuration of the cubic n x n x n-cell FCC box: the element of every atom, the atom on
 * every lattice site (lattice ids x-major, the order ReassignLatticeVector gives,
 * Config.cpp:466-552) and every atom's periodic image map_shift. Atom ids are the file's line
 * order, as in the reference. Other boxes are rejected with LMC_ERR_ARG. */
typedef struct lmc_config lmc_config;
/* cfg::GenerateFCC({n, n, n}, element) (Config.cpp:1060-1095): basis 4.046 n, atom ids in
 * GenerateFCC's creation order (cells z-major, basis sites (0,0,0) (1,1,0) (1,0,1) (0,1,1) in a/2
 * units), map_shift zero. */
int lmc_config_generate(int32_t n, const char *element, lmc_config **out);
/* Config::ReadConfig(filename) followed by ReassignLatticeVector (Config.cpp:554-661, gzip by
 * the .gz extension). "Cannot open <filename>" as the reference; unknown element strings are X
 * (Element.hpp:25-26). */
int lmc_config_read(const char *filename, lmc_config **out);
/* Config::ReadMap(lattice_filename, element_filename, map_filename) (Config.cpp:814-885). The
 * lattice file's positions must be in x-major order (as WriteLattice writes them after
 * ReassignLatticeVector); its neighbour lists are implied by the positions. Map errors carry the
 * reference's messages ("Lattice id out of bounds in map file: ", "Duplicate lattice id in map
 * file: "). */
int lmc_config_read_map(const char *lattice_filename, const char *element_filename,
                        const char *map_filename, lmc_config **out);
void lmc_config_free(lmc_config *c);
/* Cells per edge, number of sites and the row-major 3 x 3 basis (A); any pointer may be null. */
int lmc_config_info(const lmc_config *c, int32_t *ncell, int64_t *num_sites, double *basis);
/* Species code of every lattice id for the species list species[n_species] (a model's sorted
 * species with X, see lmc_model_species_name): codes[site] indexes that list. LMC_ERR_ARG if an element
 * of the configuration is not in the list. */
int lmc_config_codes(const lmc_config *c, int32_t n_species, const char *const *species,
                     uint8_t *codes);
/* Atom on every lattice site, lattice_to_atom[num_sites], and map_shift[num_sites][3] by atom
 * id (Config::GetLatticeToAtomVector, GetMapShiftList); either pointer may be null. */
int lmc_config_atoms(const lmc_config *c, int64_t *lattice_to_atom, int32_t *map_shift);
/* The state after a run: species codes by lattice id and, optionally (null keeps the current
 * ones), the atom on every site (a permutation of the atom ids) and map_shift by atom id. Each
 * atom takes the element of its site. */
int lmc_config_update(lmc_config *c, int32_t n_species, const char *const *species,
                      const uint8_t *codes, const int64_t *lattice_to_atom,
                      const int32_t *map_shift);
/* The row-major 3 x 3 basis (A) of a cubic box: diagonal, equal edges. */
int lmc_config_set_basis(lmc_config *c, const double *basis);
/* Config::WriteConfig(filename) (Config.cpp:663-710, gzip by the .gz extension): the same text,
 * byte for byte, as the reference writes for the same configuration. */
int lmc_config_write(const lmc_config *c, const char *filename);
/* Config::WriteExtendedXyz(filename, {}, {}) (Config.cpp:712-809): extended XYZ with the
 * lattice, periodic flags and per atom (atom-id order) species, Cartesian position and map_shift;
 * gzip by the .gz extension. Auxiliary and global lists are not written. */
int lmc_config_write_xyz(const lmc_config *c, const char *filename);
/* Config::WriteLattice / WriteElement / WriteMap (Config.cpp:887-922); a null name skips that
 * file. */
int lmc_config_write_map(const lmc_config *c, const char *lattice_filename,
                         const char *element_filename, const char *map_filename);

/* ---- Stateless evaluation on the GPU ------------------------------------------------------
 * The 12 vacancy jumps of each replica, in the reference's event order (neighbour ids
 * ascending): replaces the 12 calls of GetBarrierAndDiffFromLatticeIdPair made by
 * KineticMcFirstOmp::BuildEventList (lmc/mc/src/KineticMcFirstOmp.cpp:52-68,
 * VacancyMigrationPredictorE0.cpp:153-159, VacancyMigrationPredictorQuartic.cpp:247-276).
 * d_vac[r]: vacancy site; outputs d_nb/d_ea/d_de of length 12 R (eV).
 * Asynchronous on `stream`; errors of the model (unknown cluster type, missing element
 * parameters) are reported by lmc_check_errors(). */
int lmc_eval_jumps(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                   const int32_t *d_ncell, const int64_t *d_vac, int32_t n_replicas,
                   int64_t *d_nb, double *d_ea, double *d_de, void *stream);
/* Vacancy scan of the ExitTime analysis: in the configuration d_occ (one replica of ncell^3
 * cells) with its vacancy `vacancy` replaced by species `solvent` (-1: no vacancy), a vacancy
 * placed at each of d_sites[n_sites] in turn, and its 12 jumps in neighbour-id order:
 * d_nb/d_ea/d_de of length 12 n_sites. Replaces the loops over sites of
 * ExitTime::GetBarrierListAndExitTime and ExitTime::GetJumpEnergetics
 * (lmc/ansys/src/ExitTime.cpp:216-311), each jump a GetBarrierAndDiffFromLatticeIdPair call on
 * the modified configuration. Asynchronous; an invalid site is reported by lmc_check_errors(). */
int lmc_vacancy_scan(const lmc_model *m, const uint8_t *d_occ, int32_t ncell, int64_t vacancy,
                     int32_t solvent, const int64_t *d_sites, int32_t n_sites, int64_t *d_nb,
                     double *d_ea, double *d_de, void *stream);
/* One jump v -> nb of one replica, synchronous, host outputs: replaces
 * GetBarrierAndDiffFromLatticeIdPair(config, {v, nb}). */
int lmc_eval_pair(const lmc_model *m, const uint8_t *d_occ, int32_t ncell, int64_t v, int64_t nb,
                  double *ea, double *de);
/* Swap energy change of d_pairs[2r], d_pairs[2r+1] per replica:
 * EnergyChangePredictorPairSite::GetDeFromLatticeIdPair (EnergyChangePredictorPairSite.cpp:69-81). */
int lmc_swap_de(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                const int32_t *d_ncell, const int64_t *d_pairs, int32_t n_replicas, double *d_de,
                void *stream);
/* Energy change of setting d_sites[r] to species d_codes[r]:
 * EnergyChangePredictorPairSite::GetDeFromLatticeIdSite (EnergyChangePredictorPairSite.cpp:154-192). */
int lmc_site_de(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                const int32_t *d_ncell, const int64_t *d_sites, const uint8_t *d_codes,
                int32_t n_replicas, double *d_de, void *stream);
/* Energy change of site d_sites[r] going from species d_from[r] to d_codes[r], the rest of the
 * occupancy as it is: GetDeFromLatticeIdSite on the configuration with that site set to
 * d_from[r] first (the ExitTime analyses, ExitTime.cpp:401-512, evaluate every site this way).
 * Replicas may share one occupancy (equal offsets). */
int lmc_site_change_de(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                       const int32_t *d_ncell, const int64_t *d_sites, const uint8_t *d_from,
                       const uint8_t *d_codes, int32_t n_replicas, double *d_de, void *stream);
/* Total cluster-expansion energy per replica: EnergyPredictor::GetEnergy
 * (lmc/pred/src/EnergyPredictor.cpp:173-177). */
int lmc_total_energy(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                     const int32_t *d_ncell, int32_t n_replicas, double *d_energy, void *stream);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): lmc_total_energy
 * with the same kernels and results, its cluster-count scratch a fixed per-device array of
 * 2^19 counts (4 MiB, held until the process ends) instead of a hipMalloc'd buffer. It allocates
 * nothing, does not synchronise, and rejects with LMC_ERR_ARG
 * - n_replicas * n_eci above 2^19 (n_eci = 95 for Al/Mg/Zn: at most 5518 replicas; 167 for
 *   Al/Cu/Mg/Zn: 3139), and
 * - a model whose tables are not yet on the current device (any other compute entry point with
 *   that model uploads them).
 * Every call on one device shares that device's array: calls that overlap on one device race. */
int lmc_total_energy_noalloc(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                             const int32_t *d_ncell, int32_t n_replicas, double *d_energy,
                             void *stream);
/* EnergyPredictor::GetEncode (lmc/pred/src/EnergyPredictor.cpp:40-96): per replica the ordered
 * cluster counts divided by the total-energy normalisers, d_encode[R][n_eci] (n_eci from
 * lmc_model_layout, in ECI key order); theta . encode is lmc_total_energy's value. */
int lmc_total_encode(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                     const int32_t *d_ncell, int32_t n_replicas, double *d_encode, void *stream);
/* EnergyPredictor::GetEncodeOfCluster / GetEnergyOfCluster (EnergyPredictor.cpp:97-187): the
 * encode (host encode[n_eci], may be null) and energy (host, may be null) of the clusters whose
 * sites all lie in the union of sites[k] and their three neighbour shells, in the device
 * occupancy d_occ of one n^3-cell box. Sites are lattice ids (the reference takes atom ids).
 * Synchronous. */
int lmc_cluster_energy(const lmc_model *m, const uint8_t *d_occ, int32_t n, const int64_t *sites,
                       int32_t k, double *encode, double *energy);
/* EnergyPredictor::GetChemicalPotential(solvent) (EnergyPredictor.cpp:194-212): mu[code] =
 * E(15^3-cell solvent box with one site of that species) - E(pure solvent box) for every other
 * species of the model, X included, and 0 for the solvent; the S boxes are one GPU batch.
 * Synchronous. */
int lmc_chemical_potential(const lmc_model *m, const char *solvent, double *mu);
/* Returns LMC_ERR_MODEL (and clears the flag) if a kernel of model m met an unknown cluster
 * type or an element without barrier parameters since the last check. Synchronous. */
int lmc_check_errors(const lmc_model *m);

/* ---- Batched KineticMcFirst driver ------------------------------------------------------
 * Replaces KineticMcFirstOmp(config, ..., maximum_steps, ..., restart_steps, restart_energy,
 * restart_time, temperature, element_set, json, time_temperature_filename, rate_corrector,
 * early_stop, solute_disp) (lmc/mc/include/KineticMcFirstOmp.h:11-25) for R independent
 * replicas, and Simulate() (KineticMcAbstract.cpp:184-190) by lmc_kmc_run. The initial
 * absolute energy is EnergyPredictor::GetEnergy of the start config (McAbstract.cpp:26).
 * Random numbers: Philox4x32-10 keyed by seed, counter (step, replica_id, 0). */
typedef struct {
  double temperature;      /* K */
  uint64_t seed;
  uint32_t replica_id;     /* Philox stream of this replica */
  uint64_t restart_steps;
  double restart_energy;
  double restart_time;
  uint64_t maximum_steps;  /* the run ends after step maximum_steps (steps_ <= maximum_steps) */
  int32_t rate_corrector;  /* RateCorrector (lmc/pred/include/RateCorrector.hpp:17-24) */
  int32_t early_stop;      /* IsEscaped (KineticMcAbstract.cpp:192-223) */
  int32_t solute_disp;     /* solute centre-of-mass trajectory (KineticMcAbstract.cpp:163-177) */
} lmc_kmc_replica;

/* occ: host array of all replicas back to back (ncell[r] gives the sizes). */
int lmc_kmc_create(const lmc_model *barrier, const lmc_model *energy, int32_t n_replicas,
                   const int32_t *ncell, const uint8_t *occ, const lmc_kmc_replica *params,
                   lmc_kmc **out);
void lmc_kmc_free(lmc_kmc *k);
/* Piecewise-linear T(time) shared by all replicas (TimeTemperatureInterpolator). */
int lmc_kmc_set_schedule(lmc_kmc *k, int32_t n_points, const double *time, const double *temp);
/* Per-replica trace of the next `capacity` steps (selected neighbour, Ea, dE, time). 0 = off. */
int lmc_kmc_set_trace(lmc_kmc *k, uint64_t capacity);
/* Kernel of this driver: 0 dense (256 threads per replica, up to four replicas per CU),
 * 1 solvent baseline (see lmc_model_set_formulation), 2 dense wide shape (768 threads per
 * replica, one per CU: for batches of at most one replica per CU); these are first-order steps
 * and have no reference counterpart. 3: second-order residence-time steps, replacing
 * KineticMcChainOmpi (lmc/mc/include/KineticMcChainOmpi.h, KineticMcChainOmpi.cpp:56-151):
 * the 12 x 12 jumps out of the vacancy's first neighbours after each virtual jump, the
 * second-order probabilities and the deterministic time t_2; the previous site starts as the
 * vacancy's first neighbour in id order (KineticMcAbstract.cpp:252). Takes no surrogate. */
int lmc_kmc_set_formulation(lmc_kmc *k, int32_t formulation);
/* Autotuner (no reference counterpart): runs `steps` KMC steps with each kernel shape on a
 * scratch copy of the batch state (HIP events), keeps the fastest and leaves the replicas
 * untouched. ms_per_step[3]: per-step times of formulations 0, 1, 2 (NaN when not tried: the
 * wide shape is tried only when every replica has a CU of its own). */
int lmc_kmc_autotune_shapes(lmc_kmc *k, int32_t steps, int32_t *chosen, double *ms_per_step);
/* lmc_kmc_autotune_shapes reporting formulations 0 and 1 only. */
int lmc_kmc_autotune(lmc_kmc *k, int32_t steps, int32_t *chosen, double *ms_dense,
                     double *ms_solvent);
/* Advance every replica by `steps` steps (fewer if it reaches maximum_steps or escapes).
 * Asynchronous on `stream`. */
int lmc_kmc_run(lmc_kmc *k, uint64_t steps, void *stream);
/* `launches` launches of `steps_per_launch` steps each, captured once in a HIP graph (rebuilt when
 * the arguments, the formulation or the surrogate change) and replayed on `stream`; the same
 * work as that many lmc_kmc_run calls. Needs tracing off. Asynchronous. */
int lmc_kmc_run_graph(lmc_kmc *k, uint64_t steps_per_launch, int32_t launches, void *stream);
/* Captures that graph without running it (lmc_kmc_run_graph with the same arguments replays it). */
int lmc_kmc_build_graph(lmc_kmc *k, uint64_t steps_per_launch, int32_t launches);
/* Host copies of the state arrays (each of length R, or 3R for disp/com); any may be NULL.
 * disp: vacancy displacement in units of a/2; com: solute COM trajectory (A). Synchronous. */
int lmc_kmc_state(lmc_kmc *k, uint64_t *steps, double *time, double *energy, double *abs_energy,
                  double *temperature, int64_t *vacancy, int64_t *disp, double *com,
                  int32_t *escaped);
/* The escape of an early-stop replica (KineticMcFirstAbstract::IsEscaped, KineticMcAbstract.cpp:
 * 192-223), arrays of length R, any may be NULL: t_exit, the time the check found every 1-3NN
 * site of the vacancy solvent (the reference prints it as "t_exit: "), the step it ran in
 * ("steps: "), and the vacancy's site before that step's jump (-1: no escape), so that the
 * configuration the reference writes to escaped.cfg.gz is the final one with that last jump
 * undone (LatticeJump of the final vacancy site and site_before). Synchronous. */
int lmc_kmc_escape_info(lmc_kmc *k, double *t_exit, uint64_t *step, int64_t *site_before);
int lmc_kmc_occupancy(lmc_kmc *k, uint8_t *occ);
/* Trace arrays [R][capacity]. */
int lmc_kmc_trace(lmc_kmc *k, int64_t *sel, double *ea, double *de, double *time);
/* Atom tracking, replacing the atom ids and map_shift that Config::LatticeJump keeps
 * (Config.cpp:431-456; lattice_to_atom_vector_, map_shift_list_): lattice_to_atom[total] gives
 * the atom on every site of the batch, ids local to each replica (a permutation of 0..N_r-1, as
 * lmc_config_atoms returns them), and map_shift[total][3] the periodic image of every atom (row
 * offset_r + atom id; null: zero). Every jump then swaps the two atoms and moves the image change
 * of the vacancy -> target displacement from the vacancy's atom to the migrating one. Null
 * lattice_to_atom turns tracking off. Synchronous. */
int lmc_kmc_set_atoms(lmc_kmc *k, const int64_t *lattice_to_atom, const int32_t *map_shift);
/* The tracked atom ids and map_shift, in the layout of lmc_kmc_set_atoms (either may be null). */
int lmc_kmc_atoms(lmc_kmc *k, int64_t *lattice_to_atom, int32_t *map_shift);
/* kmc_log.txt, replacing KineticMcFirstAbstract::Dump's log (KineticMcAbstract.cpp:59-104):
 * with log_dump_steps > 0 the kernels keep a record at the reference's cadence (every step below
 * step 99, then every 10^(floor(log10(steps + 1)) - 1) steps up to log_dump_steps, and every
 * log_dump_steps steps past 10 log_dump_steps), up to `capacity` records per replica between
 * writes. A record is taken after the step's event is chosen and before it is applied: steps,
 * time, temperature, energy, Ea, dE, the migrating atom (-1 without lmc_kmc_set_atoms), the
 * vacancy's unwrapped Cartesian position (box_edge[r] A, null: 4.046 n, times the relative
 * position plus the vacancy atom's map_shift) and, with solute_disp, the solute centre of mass.
 * log_dump_steps = 0 turns it off; at most 1e17. Synchronous. */
int lmc_kmc_set_log(lmc_kmc *k, uint64_t log_dump_steps, uint64_t capacity,
                    const double *box_edge);
/* Appends the records taken since the last call to filenames[r] (null: skip and drop that
 * replica's records) in the reference's text format: tab-separated columns, precision 16, the
 * default float format up to the first vector and std::fixed after it, the header line and a
 * truncated file for a run from step 0, and for a restarted run an appended file without its
 * first dump. LMC_ERR_ARG after writing if records were dropped for lack of capacity. */
int lmc_kmc_write_log(lmc_kmc *k, const char *const *filenames);
/* Device views of the batch occupancy (for callers that evaluate it with lmc_eval_jumps). */
int lmc_kmc_device_views(lmc_kmc *k, uint8_t **d_occ, const int64_t **d_offset,
                         const int32_t **d_ncell, const int64_t **d_vac);

/* ---- Batched CanonicalMc driver -----------------------------------------------------------
 * Replaces CanonicalMcSerial(config, ..., maximum_steps, ..., restart_steps, restart_energy,
 * temperature, element_set, json) (lmc/mc/include/CanonicalMcSerial.h:11-20) and Simulate()
 * (CanonicalMcSerial.cpp:40-51). Temperature: t_start + (t_end - t_start) * step/maximum_steps
 * (t_end = t_start gives the reference's constant temperature). Random numbers: Philox
 * counter (step, replica_id, 2*attempt+1) for the site pair, (step, replica_id, 0xFFFFFFFE) for
 * the Metropolis test. */
typedef struct {
  double t_start, t_end;
  uint64_t seed;
  uint32_t replica_id;
  uint64_t restart_steps;
  double restart_energy;
  uint64_t maximum_steps;
} lmc_cmc_replica;
int lmc_cmc_create(const lmc_model *pairsite, const lmc_model *energy, int32_t n_replicas,
                   const int32_t *ncell, const uint8_t *occ, const lmc_cmc_replica *params,
                   lmc_cmc **out);
void lmc_cmc_free(lmc_cmc *c);
int lmc_cmc_set_trace(lmc_cmc *c, uint64_t capacity);
/* Simulated annealing, replacing SimulatedAnnealing (lmc/mc/include/SimulatedAnnealing.h,
 * SimulatedAnnealing.cpp:99-185): the same swap trials with the reference's adaptive schedule
 * instead of the ramp, from T = t_start: cooling by 0.99 after each window of N/1000 trials
 * whose acceptance exceeds 0.5, up to 5 reheats by 1.10 after N/20 trials without improvement
 * when the acceptance falls below 0.05, and the baseline cooling exp(-3/N) per trial
 * (N = maximum_steps). Energies are relative to the starting configuration (the reference
 * subtracts chemical potentials). enable = 0 returns to the ramp. Synchronous. */
int lmc_cmc_set_annealing(lmc_cmc *c, int32_t enable);
/* Per replica: current temperature, lowest energy at a trial start, reheats done. */
int lmc_cmc_annealing_state(lmc_cmc *c, double *temperature, double *lowest, uint32_t *reheats);
/* No reference counterpart (a placement choice): enable != 0, the default, holds each
 * replica's occupancy in LDS while a run works on it, when the largest replica fits beside the
 * kernel's tables (about 150 000 sites in a 160 KB workgroup); 0 keeps it in HBM and leaves the
 * CU's LDS to kernels that run beside it. Results are the same either way. */
int lmc_cmc_set_lds_occupancy(lmc_cmc *c, int32_t enable);
/* CanonicalMcOmp (lmc/mc/include/CanonicalMcOmp.h:9-29, CanonicalMcOmp.cpp:40-92) in place of
 * the serial trials: rounds of up to trials_per_round (1..64; the reference's OpenMP thread
 * count) swap pairs, each event's pair the first draw with neither site on or within 3NN of an
 * earlier event's sites, the 50th draw of an event ending the round; every dE on the round's
 * starting occupancy, then the Metropolis steps in trial order. A new round starts only while
 * steps <= maximum_steps; a started round runs to its end, across runs. The attempts of the
 * round starting at step s0 are numbered through the round: attempt a is Philox counter (s0,
 * replica_id, 2a+1); the Metropolis uniform of step s stays (s, replica_id, 0xFFFFFFFE), so one
 * trial per round gives the serial trajectory. 0 restores the serial trials. No annealing with
 * rounds. Synchronous. */
int lmc_cmc_set_parallel_trials(lmc_cmc *c, int32_t trials_per_round);
/* Per replica, the trials of the current CanonicalMcOmp round not yet run (0 for serial trials);
 * left[R]. Synchronous. */
int lmc_cmc_round_left(lmc_cmc *c, int32_t *left);
int lmc_cmc_run(lmc_cmc *c, uint64_t trials, void *stream);
int lmc_cmc_state(lmc_cmc *c, uint64_t *steps, double *energy, double *abs_energy,
                  uint64_t *accepted);
int lmc_cmc_occupancy(lmc_cmc *c, uint8_t *occ);
int lmc_cmc_trace(lmc_cmc *c, int64_t *p, int64_t *q, double *de, int8_t *accepted);
/* Atom tracking as lmc_kmc_set_atoms / lmc_kmc_atoms: an accepted swap of p and q is
 * Config::LatticeJump({p, q}) (CanonicalMcAbstract.cpp:86-100). */
int lmc_cmc_set_atoms(lmc_cmc *c, const int64_t *lattice_to_atom, const int32_t *map_shift);
int lmc_cmc_atoms(lmc_cmc *c, int64_t *lattice_to_atom, int32_t *map_shift);
/* cmc_log.txt, replacing CanonicalMcAbstract::Dump's log (CanonicalMcAbstract.cpp:53-84), and
 * with annealing on the log of SimulatedAnnealing::Dump (SimulatedAnnealing.cpp:75-97). With
 * log_dump_steps > 0 the kernel keeps a record of the step about to be tried, up to `capacity`
 * per replica between writes: steps, temperature, energy, the fourth column and the absolute
 * energy.
 * - canonical MC: on KineticMc's cadence (every step below step 99, then every
 *   10^(floor(log10(steps + 1)) - 1) steps up to log_dump_steps, every log_dump_steps steps past
 *   10 log_dump_steps); the fourth column is ThermodynamicAveraging's average over the last
 *   thermodynamic_averaging_steps energies (ThermodynamicAveraging.cpp:6-40; 0 when that is 0),
 *   counted from this call. The averages are formed after each launch from a ring of
 *   (thermodynamic_averaging_steps + 65536) doubles per replica, so runs are split into launches
 *   of 65536 trials while it is on.
 * - annealing: every log_dump_steps steps and at every step whose energy is below the lowest so
 *   far by more than 1e-4 eV; the fourth column is the lowest energy.
 * log_dump_steps = 0 turns it off; at most 1e17 (thermodynamic_averaging_steps at most 2^31).
 * Synchronous. */
int lmc_cmc_set_log(lmc_cmc *c, uint64_t log_dump_steps, uint64_t capacity,
                    uint64_t thermodynamic_averaging_steps);
/* Appends the records taken since the last call to filenames[r] (null: skip and drop that
 * replica's records) in the reference's text format: tab-separated columns, precision 16 in the
 * default float format, the header line and a truncated file for a run from step 0, an appended
 * file for a restarted run (canonical MC also without its first dump). LMC_ERR_ARG after
 * writing if records were dropped for lack of capacity. */
int lmc_cmc_write_log(lmc_cmc *c, const char *const *filenames);

/* ---- Learned surrogate of the jump evaluation ------------------------------------------
 * No reference counterpart: an ensemble of BF16 MLPs (one-hot jump environment -> 128 -> 128
 * -> (E_kra, dE), SiLU, FP32 accumulation on the matrix cores) that stands in for
 * GetBarrierAndDiffFromLatticeIdPair (VacancyMigrationPredictorE0.cpp:153-159,
 * VacancyMigrationPredictorQuartic.cpp:247-276), with the cluster-expansion predictor as its
 * fallback and the reference of its errors. Input rows are the 60 environment slots of a jump in the engine's fixed
 * per-direction slot order (one species code per byte, rows of 64 bytes). Ea = max(0, E_kra +
 * dE/2); for the quartic model E_kra = Ea - dE/2. Supports 3 or 4 elements (plus vacancy). */
typedef struct lmc_surrogate lmc_surrogate;
int lmc_surrogate_create(const lmc_model *barrier, int32_t n_members, uint64_t seed,
                         lmc_surrogate **out);
void lmc_surrogate_free(lmc_surrogate *s);
/* Members, padded input width, parameters per member, channels per slot and their species. */
int lmc_surrogate_layout(const lmc_surrogate *s, int32_t *n_members, int32_t *n_inputs,
                         int32_t *n_params, int32_t *n_channels, int32_t *channel_codes4,
                         int32_t *log_ekra);
/* FP32 parameters [n_members][n_params] (layout W1[K][128] b1 W2[128][128] b2 W3[128][2] b3). */
int lmc_surrogate_get_params(const lmc_surrogate *s, float *params);
int lmc_surrogate_set_params(lmc_surrogate *s, const float *params);
/* Output normalisation: dE = y1 * sigma[1] + mu[1]; E_kra = y0 * sigma[0] + mu[0] for a
 * quartic model and exp(y0 * sigma[0] + mu[0]) for an E0 model (whose E_kra is the exponential
 * of its cluster sums, E0 = exp(logit * sigma_e0 + mu_e0), VacancyMigrationPredictorE0.cpp:147). */
int lmc_surrogate_set_normalization(lmc_surrogate *s, const double *mu2, const double *sigma2);
int lmc_surrogate_get_normalization(const lmc_surrogate *s, double *mu2, double *sigma2);
/* Exact training samples: for each replica, the 12 jumps in event order: environment rows
 * d_env[12 R][64], E_kra and dE (the plain predictor, FP64). */
int lmc_jump_samples(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                     const int32_t *d_ncell, const int64_t *d_vac, int32_t n_replicas,
                     uint8_t *d_env, double *d_ekra, double *d_de, void *stream);
/* Normalised outputs of every member for n_rows environment rows: d_out[n_rows][M][2]. */
int lmc_surrogate_forward(const lmc_surrogate *s, const uint8_t *d_env, int32_t n_rows,
                          float *d_out, void *stream);
/* Gradient of each member's loss sum_rows,outputs (y - t)^2 / (2 n_rows) with respect to its
 * parameters (d_grad[M][n_params]) for normalised targets d_target[n_rows][2]; host loss[M]
 * (optional) receives the sum of squared errors. Synchronous. */
int lmc_surrogate_gradients(lmc_surrogate *s, const uint8_t *d_env, const float *d_target,
                            int32_t n_rows, float *d_grad, double *loss);
/* Adam training of every member on (environment, E_kra, dE) samples: sets the normalisation
 * from the samples if none is set, shuffles each epoch (Philox, shuffle_seed), cosine-decays
 * the learning rate from lr to lr/20. Host mse[M] (optional): mean squared error per output of
 * the last epoch, in normalised units. Synchronous. */
int lmc_surrogate_train(lmc_surrogate *s, const uint8_t *d_env, const double *d_ekra,
                        const double *d_de, int32_t n_rows, int32_t epochs, int32_t batch,
                        double lr, uint64_t shuffle_seed, double *mse);
/* The surrogate's evaluation of the 12 jumps of each replica, in event order, exactly as the
 * KMC step makes it: Ea = max(0, E_kra + dE/2) from the member means; an event falls back to the
 * model m (exact, d_fallback = 1) when its environment is out of distribution (a vacancy off the
 * endpoints, more solute slots than any training row) or when the members disagree on Ea or dE
 * by more than `threshold` eV (< 0: every event falls back) while the event's lowest member
 * barrier lies within `window` kT (T = d_temp[r]) of the replica's lowest barrier (window <= 0:
 * every event counts). Needs 4 members. Asynchronous. */
int lmc_surrogate_eval_jumps(const lmc_surrogate *s, const lmc_model *m, const uint8_t *d_occ,
                             const int64_t *d_offset, const int32_t *d_ncell,
                             const int64_t *d_vac, const double *d_temp, int32_t n_replicas,
                             double threshold, double window, int64_t *d_nb, double *d_ea,
                             double *d_de, int8_t *d_fallback, void *stream);
/* Use the surrogate (NULL: the cluster expansion again) in every later lmc_kmc_run of k, with
 * the fallback rule of lmc_surrogate_eval_jumps. The selected jump's dE is always recomputed
 * exactly, so energies stay those of the cluster expansion. */
int lmc_kmc_set_surrogate(lmc_kmc *k, const lmc_surrogate *s, double threshold, double window);
/* Events evaluated since the surrogate was attached, and how many of them fell back. */
int lmc_kmc_surrogate_stats(lmc_kmc *k, uint64_t *events, uint64_t *fallbacks);

#ifdef __cplusplus
}
#endif
#endif /* LMC_CAPI_H_ */
