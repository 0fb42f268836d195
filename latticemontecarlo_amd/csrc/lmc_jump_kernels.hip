// gfx950 kernels of the vacancy-jump evaluation and the KineticMcFirst step.
//
// Work mapping: one 256-thread workgroup per replica. Per KMC step
//   1. gather   : the 12 jump environments (12 x 60 sites) -> LDS (bytes, from HBM/L2);
//   2. clusters : per-cluster weights of dE (one-endpoint state clusters,
//                 VacancyMigrationPredictorQuartic.cpp:112-166), of the barrier logit (MMM
//                 orbit clusters, VacancyMigrationPredictorE0.cpp:128-151) and of the Ks logit
//                 (MM2 forward + backward, VacancyMigrationPredictorQuartic.cpp:167-215), in one
//                 of two exact formulations:
//                   dense   - 21 lanes per jump sum every cluster (4 accumulators per lane);
//                   solvent - per (solvent, migrating species) sums of an all-solvent
//                             environment are precomputed; 16 lanes per jump add, for each
//                             environment site holding another species, the changes of the
//                             clusters it touches;
//   3. events   : 12 lanes form dE, Ea and the rate exp(-Ea beta) in a fixed order and store
//                 them at the jump's rank (neighbour ids ascending);
//   4. select   : the 12 event lanes form the ordered cumulative probabilities (FP64); the lane
//                 holding the chosen event advances the residence-time clock (Philox u1/u2
//                 drawn by an idle lane during the gather) and applies the jump
//                 (KineticMcFirstOmp.cpp:52-82, KineticMcAbstract.cpp:106-182).
// Clusters containing both jump endpoints have the same species multiset before and after the
// jump and are skipped. Tables live in dynamic LDS; the barrier weight tables go there when they
// fit the workgroup's share of the CU's LDS, otherwise they are read from global memory.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <type_traits>

#include "lmc_jump_device.h"

namespace lmc {
namespace {

// One workgroup per replica r (vacancy vac[r] in occupancy occ + off[r] of ncell[r]^3 cells).
// kScan: one workgroup per scanned site vac[r] of the single occupancy occ (scan_n^3 cells), with
// the gather substitution `sub`.
template <bool kSolvent, bool kBL, bool kScan>
__global__ __launch_bounds__(kThreads) void EvalJumpsKernel(KTables t, DynLayout lay,
                                                            const uint8_t *occ,
                                                            const int64_t *off,
                                                            const int32_t *ncell,
                                                            const int64_t *vac, int64_t *nb,
                                                            double *ea, double *de, int32_t *err,
                                                            uint8_t *env_out, double *ekra_out,
                                                            int scan_n, GatherSub sub) {
  __shared__ JumpLds L;
  extern __shared__ char dyn[];
  const int r = blockIdx.x;
  const int n = kScan ? scan_n : ncell[r];
  if (!ValidSite(n, vac[r])) {
    if (threadIdx.x == 0) atomicOr(err, 8);
    return;
  }
  const Tabs T =
      LoadJumpTables<kSolvent, kBL>(t, L, dyn, lay, threadIdx.x, kThreads, threadIdx.x, kThreads);
  __syncthreads();
  EvaluateJumps<kSolvent, kScan>(t, L, T, kScan ? occ : occ + off[r], n, ToXYZ(vac[r], n),
                                 t.solvent_default, 0.0, err, threadIdx.x, nullptr, sub);
  const int k = threadIdx.x;
  if (k < 12) {
    const int64_t o = static_cast<int64_t>(r) * 12 + k;
    nb[o] = L.snb[k];
    ea[o] = L.ea[k];
    de[o] = L.de[k];
    if (ekra_out) ekra_out[o] = L.ekra[k];
  }
  if (env_out) {  // environment rows of the events in rank order (60 slot codes, zero padded)
    for (int i = threadIdx.x; i < 12 * 64; i += kThreads) {
      const int q = i >> 6, j = i & 63;
      env_out[(static_cast<int64_t>(r) * 12 + q) * 64 + j] = j < kEnvSlots ? L.env[L.dir[q]][j] : 0;
    }
  }
}

// Replica r's state into LDS (one thread).
__device__ void LoadReplica(const KmcDev &k, int r, int n, ReplicaLds &Q) {
  Q.steps = k.steps[r];
  Q.max_steps = k.max_steps[r];
  Q.seed = k.seed[r];
  Q.rid = k.rid[r];
  Q.time = k.time[r];
  Q.energy = k.energy[r];
  Q.abse = k.abs_energy[r];
  Q.T = k.temp[r];
  Q.cv = k.cv[r];
  Q.cs = k.cs[r];
  Q.smass = k.solute_mass[r];
  Q.vac = k.vac[r];
  Q.flags = k.flags[r];
  Q.esc = k.escaped[r];
  Q.solvent = k.solvent[r];
  for (int a = 0; a < 3; ++a) {
    Q.disp[a] = k.disp[3 * r + a];
    Q.com[a] = k.com[3 * r + a];
  }
  Q.done = (Q.steps > Q.max_steps || Q.esc || !ValidSite(n, Q.vac)) ? 1 : 0;
  Q.beta = 1.0 / kBoltzmann / Q.T;
  SetVacancyXYZ(Q, n);
}

// Replica r's state back to HBM (one thread).
__device__ void StoreReplica(const KmcDev &k, int r, const ReplicaLds &Q) {
  k.steps[r] = Q.steps;
  k.time[r] = Q.time;
  k.energy[r] = Q.energy;
  k.abs_energy[r] = Q.abse;
  k.temp[r] = Q.T;
  k.vac[r] = Q.vac;
  k.escaped[r] = Q.esc;
  for (int a = 0; a < 3; ++a) {
    k.disp[3 * r + a] = Q.disp[a];
    k.com[3 * r + a] = Q.com[a];
  }
}

// UpdateTemperature (KineticMcAbstract.cpp:45-50), before any lane reads beta.
__device__ __forceinline__ void UpdateTemperature(const KmcDev &k, ReplicaLds &Q, int tid) {
  if (k.n_tt > 0) {
    if (tid == 0) {
      Q.T = ScheduleTemperature(k.tt_time, k.tt_temp, k.n_tt, Q.time);
      Q.beta = 1.0 / kBoltzmann / Q.T;
    }
    __syncthreads();
  }
}

// Persistent KMC: one 256-thread workgroup per replica, up to four per CU.
template <bool kSolvent, bool kBL, bool kRecord>
__global__ __launch_bounds__(kThreads, 4) void KmcKernel(KTables t, DynLayout lay, KmcDev k,
                                                      uint64_t nsteps, uint64_t trace_base) {
  __shared__ JumpLds L;
  __shared__ ReplicaLds Q;
  extern __shared__ char dyn[];
  const int r = blockIdx.x;
  const int tid = threadIdx.x;
  const Tabs T = LoadJumpTables<kSolvent, kBL>(t, L, dyn, lay, tid, kThreads, tid, kThreads);
  uint8_t *occ = k.occ + k.off[r];
  const int n = k.ncell[r];
  if (tid == 0) LoadReplica(k, r, n, Q);
  __syncthreads();
  const int sv = Q.solvent;
  for (uint64_t it = 0; it < nsteps; ++it) {
    if (Q.done) break;
    RotatePriority(it);
    UpdateTemperature(k, Q, tid);
    // the selection reads the events in wave 0, which finished them (dense shape)
    EvaluateJumps<kSolvent, false, false>(t, L, T, occ, n, VacancyXYZ(Q), sv, Q.beta, k.error,
                                          tid, &Q);
    SelectAndApply<kRecord>(t, k, L, Q, occ, r, trace_base + it, tid);
    __syncthreads();
  }
  if (tid == 0) StoreReplica(k, r, Q);
}

// Persistent KMC, wide shape: one 768-thread workgroup per replica (wave k evaluates event k),
// for batches of at most one replica per CU, where a step's latency sets the rate.
template <bool kBL, bool kRecord>
__global__ __launch_bounds__(kWideThreads, 1) void KmcWideKernel(KTables t, DynLayout lay,
                                                                 KmcDev k, uint64_t nsteps,
                                                                 uint64_t trace_base) {
  __shared__ JumpLds L;
  __shared__ ReplicaLds Q;
  extern __shared__ char dyn[];
  const int r = blockIdx.x;
  const int tid = threadIdx.x;
  const Tabs T =
      LoadJumpTables<false, kBL, true>(t, L, dyn, lay, tid, kWideThreads, tid, kWideThreads);
  uint8_t *occ = k.occ + k.off[r];
  const int n = k.ncell[r];
  if (tid == 0) LoadReplica(k, r, n, Q);
  __syncthreads();
  for (uint64_t it = 0; it < nsteps; ++it) {
    if (Q.done) break;
    UpdateTemperature(k, Q, tid);
    {
      Gathered<kWideThreads> G;
      GatherIssue<kWideThreads>(L, occ, n, VacancyXYZ(Q), tid, G);
      DrawSteps<kWideThreads>(L, Q, tid);
      GatherStore<kWideThreads>(t, L, k.error, tid, G);
    }
    __syncthreads();
    WideSums(t, L, T, tid);
    __syncthreads();
    if (tid < 64) {  // wave 0 finishes the 12 events, then selects
      FinishFromSums(t, L, Q.beta, k.error, tid);
      WaveFence();
    }
    SelectAndApply<kRecord>(t, k, L, Q, occ, r, trace_base + it, tid);
    __syncthreads();
  }
  if (tid == 0) StoreReplica(k, r, Q);
}

// Persistent second-order (chain) KMC, KineticMcChainOmpi: one 256-thread workgroup per replica.
// Each step evaluates the 12 jumps out of every first neighbour i of the vacancy k in the
// configuration after the jump k -> i (12 passes of the dense evaluation: vacancy at i, site k
// holding i's species), then lane 0 forms the second-order probabilities and time and jumps.
template <bool kBL, bool kRecord>
__global__ __launch_bounds__(kThreads, 4) void KmcChainKernel(KTables t, DynLayout lay, KmcDev k,
                                                           uint64_t nsteps, uint64_t trace_base) {
  __shared__ JumpLds L;
  __shared__ ReplicaLds Q;
  __shared__ ChainLds C;
  extern __shared__ char dyn[];
  const int r = blockIdx.x;
  const int tid = threadIdx.x;
  const Tabs T = LoadJumpTables<false, kBL>(t, L, dyn, lay, tid, kThreads, tid, kThreads);
  uint8_t *occ = k.occ + k.off[r];
  const int n = k.ncell[r];
  if (tid == 0) {
    LoadReplica(k, r, n, Q);
    C.prev = k.prev_j[r];
    L.bad = 0;
  }
  __syncthreads();
  const int sv = Q.solvent;
  for (uint64_t it = 0; it < nsteps; ++it) {
    if (Q.done) break;
    RotatePriority(it);
    UpdateTemperature(k, Q, tid);
    ChainNeighbours(t, L, C, Q, occ, n, tid);
    if (tid == kThreads - 1) DrawStep(L, Q);
    __syncthreads();
    for (int i = 0; i < 12; ++i) {
      int dx, dy, dz;
      Unpack16(C.soff[i], dx, dy, dz);
      const XYZ vi{Wrap(Q.vx + dx, Q.m2), Wrap(Q.vy + dy, Q.m2), Wrap(Q.vz + dz, Q.m2)};
      EvaluateJumps<false, true>(t, L, T, occ, n, vi, sv, Q.beta, k.error, tid, nullptr,
                                 GatherSub{Q.vac, C.sm[i]});
      if (tid == 0) ChainPass(L, C, Q, i);
      __syncthreads();
    }
    if (tid == 0) ChainSelectAndApply<kRecord>(t, k, L, C, Q, occ, n, r, trace_base + it);
    __syncthreads();
  }
  if (tid == 0) {
    StoreReplica(k, r, Q);
    k.prev_j[r] = C.prev;
  }
}

// Dynamic LDS plan: weight tables go to LDS while the workgroup stays within its share of the
// CU's LDS (up to 4 resident workgroups per CU when the batch fills the chip).
// form: 0 dense, 1 solvent baseline, 2 dense wide shape (one workgroup per CU), 3 chain steps.
DynLayout PlanLds(const KTables &t, int32_t R, int form) {
  const bool solvent = form == 1;
  const DeviceLimits &dl = CurrentDeviceLimits();
  const int cus = dl.cus;
  const size_t lds_cu = dl.lds_cu, lds_block = dl.lds_block;
  const int per_cu = form == 2 ? 1 : std::min(4, std::max(1, (R + cus - 1) / cus));
  const size_t stat = sizeof(JumpLds) + sizeof(ReplicaLds) + (form == 3 ? sizeof(ChainLds) : 0) + 256;
  const size_t share = std::min(lds_block, lds_cu / per_cu);
  auto al = [](size_t x) { return (x + 15) & ~size_t(15); };
  DynLayout lay{};
  size_t at = 0;
  lay.w = static_cast<int>(at);
  at = al(at + sizeof(double) * t.nwd);
  if (solvent) {
    lay.base = static_cast<int>(at);
    at = al(at + sizeof(double) * 3 * t.S * t.S);
    lay.lists = static_cast<int>(at);
    at = al(at + sizeof(uint32_t) * t.n_inc);
    lay.offs = static_cast<int>(at);
    at = al(at + sizeof(uint16_t) * 3 * (kEnvSlots + 1));
  } else {
    lay.lists = static_cast<int>(at);
    if (form == 2)
      at = al(at + sizeof(uint32_t) * (t.n_state_w + t.n_mmm_w + (t.kind == kQuartic ? t.n_mm2_w : 0)));
    else
      at = al(at + sizeof(uint32_t) * (t.n_state + t.n_mmm + (t.kind == kQuartic ? t.n_mm2 : 0)));
    // (n_mm2 counts the forward and the backward entries)
  }
  const size_t b1 = sizeof(double) * static_cast<size_t>(t.ne) * t.F1;
  const size_t b2 = sizeof(double) * static_cast<size_t>(t.ne) * t.F2;
  if ((t.kind == kE0 || t.kind == kQuartic) && stat + at + b1 <= share) {
    lay.b1_lds = 1;
    lay.b1 = static_cast<int>(at);
    at = al(at + b1);
  }
  if (t.kind == kQuartic && lay.b1_lds) {
    if (stat + at + b2 <= share) {
      lay.b2_lds = 1;
      lay.b2 = static_cast<int>(at);
      at = al(at + b2);
    } else {  // both barrier tables in LDS or neither
      at = lay.b1;
      lay.b1_lds = 0;
    }
  }
  lay.bytes = static_cast<int>(at);
  return lay;
}

template <typename K>
hipError_t AllowLds(K kernel, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
}

}  // namespace

hipError_t LaunchEvalJumps(const KTables &t, const uint8_t *occ, const int64_t *off,
                           const int32_t *ncell, const int64_t *vac, int32_t R, int64_t *nb,
                           double *ea, double *de, int32_t *err, hipStream_t s, bool solvent,
                           uint8_t *env_out, double *ekra_out) {
  if (R <= 0) return hipSuccess;
  const DynLayout lay = PlanLds(t, R, solvent ? 1 : 0);
  auto go = [&](auto kernel) {
    hipError_t e = AllowLds(kernel, lay.bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kernel, dim3(R), dim3(kThreads), lay.bytes, s, t, lay, occ, off, ncell,
                       vac, nb, ea, de, err, env_out, ekra_out, 0, GatherSub{-1, 0});
    return hipGetLastError();
  };
  if (lay.b1_lds)
    return solvent ? go(EvalJumpsKernel<true, true, false>) : go(EvalJumpsKernel<false, true, false>);
  return solvent ? go(EvalJumpsKernel<true, false, false>) : go(EvalJumpsKernel<false, false, false>);
}

hipError_t LaunchVacancyScan(const KTables &t, const uint8_t *occ, int32_t n, int64_t sub_site,
                             int sub_code, const int64_t *sites, int32_t K, int64_t *nb,
                             double *ea, double *de, int32_t *err, hipStream_t s) {
  if (K <= 0) return hipSuccess;
  const DynLayout lay = PlanLds(t, K, 0);
  auto go = [&](auto kernel) {
    hipError_t e = AllowLds(kernel, lay.bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kernel, dim3(K), dim3(kThreads), lay.bytes, s, t, lay, occ, nullptr,
                       nullptr, sites, nb, ea, de, err, nullptr, nullptr, n,
                       GatherSub{sub_site, sub_code});
    return hipGetLastError();
  };
  return lay.b1_lds ? go(EvalJumpsKernel<false, true, true>) : go(EvalJumpsKernel<false, false, true>);
}

// The instantiation of form (0 dense, 1 solvent baseline, 2 dense wide, 3 second-order chain
// steps) and LDS placement, with its block size; f(kernel, threads).
template <typename F>
hipError_t WithKmcKernel(int form, bool b1_lds, bool record, F &&f) {
  auto pick = [&](auto rec) {
    constexpr bool kR = decltype(rec)::value;
    if (form == 3)
      return b1_lds ? f(KmcChainKernel<true, kR>, kThreads) : f(KmcChainKernel<false, kR>, kThreads);
    if (form == 2)
      return b1_lds ? f(KmcWideKernel<true, kR>, kWideThreads) : f(KmcWideKernel<false, kR>, kWideThreads);
    if (form == 1)
      return b1_lds ? f(KmcKernel<true, true, kR>, kThreads) : f(KmcKernel<true, false, kR>, kThreads);
    return b1_lds ? f(KmcKernel<false, true, kR>, kThreads) : f(KmcKernel<false, false, kR>, kThreads);
  };
  return record ? pick(std::true_type{}) : pick(std::false_type{});
}

hipError_t PrepareKmc(const KTables &t, int32_t R, int form) {
  if (R <= 0) return hipSuccess;
  const DynLayout lay = PlanLds(t, R, form);
  auto allow = [&](auto kernel, int) { return AllowLds(kernel, lay.bytes); };
  const hipError_t e = WithKmcKernel(form, lay.b1_lds != 0, false, allow);
  return e != hipSuccess ? e : WithKmcKernel(form, lay.b1_lds != 0, true, allow);
}

hipError_t LaunchKmc(const KTables &t, const KmcDev &k, uint64_t steps, uint64_t trace_base,
                     hipStream_t s, int form) {
  if (k.R <= 0 || steps == 0) return hipSuccess;
  const DynLayout lay = PlanLds(t, k.R, form);
  const bool record = k.atom != nullptr || k.log_every != 0;
  return WithKmcKernel(form, lay.b1_lds != 0, record, [&](auto kernel, int threads) {
    hipError_t e = AllowLds(kernel, lay.bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kernel, dim3(k.R), dim3(threads), lay.bytes, s, t, lay, k, steps,
                       trace_base);
    return hipGetLastError();
  });
}

const DeviceLimits &CurrentDeviceLimits() {
  static std::mutex mu;
  static std::map<int, DeviceLimits> known;  // entries are never erased: references stay valid
  static const DeviceLimits fallback{256, 65536, 65536};
  int dev = 0;
  hipDeviceProp_t p;
  if (hipGetDevice(&dev) != hipSuccess) return fallback;
  std::lock_guard<std::mutex> lock(mu);
  const auto it = known.find(dev);
  if (it != known.end()) return it->second;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) return fallback;
  DeviceLimits d;
  d.cus = p.multiProcessorCount;
  d.lds_cu = p.maxSharedMemoryPerMultiProcessor ? p.maxSharedMemoryPerMultiProcessor : 65536;
  d.lds_block = p.sharedMemPerBlock;
  return known.emplace(dev, d).first->second;
}

}  // namespace lmc
