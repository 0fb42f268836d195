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
//                             environment are precomputed; the block lists the environment
//                             sites that hold another species and 8 lanes per such site add the
//                             changes of the clusters it touches;
//   3. events   : 12 lanes form dE, Ea and the rate exp(-Ea beta) in a fixed order and store
//                 them at the jump's rank (neighbour ids ascending);
//   4. select   : one lane sums the 12 rates in that order (FP64), draws Philox u1/u2,
//                 advances the residence-time clock and applies the jump
//                 (KineticMcFirstOmp.cpp:52-82, KineticMcAbstract.cpp:106-182).
// Clusters containing both jump endpoints have the same species multiset before and after the
// jump and are skipped. Tables live in dynamic LDS; the barrier weight tables go there when they
// fit the workgroup's share of the CU's LDS, otherwise they are read from global memory.
#include <hip/hip_runtime.h>

#include "lmc_dev_common.h"
#include "lmc_device.h"
#include "lmc_rng.h"

namespace lmc {
namespace {

constexpr int kLanes = 21;      // dense: lanes per jump event (12 * 21 = 252 <= 256)
constexpr int kGroup = 8;       // solvent: lanes per non-solvent site
constexpr int kPass = kThreads / kGroup;  // sites handled per pass
constexpr int kEnvPad = 768;    // 12 x 60 environment entries padded to whole waves

struct JumpLds {
  int32_t off[12 * kEnvSlots];
  double par[kMaxSpecies * 6];
  uint8_t mirror[64];
  int8_t tt[3 * kMaxSpecies * kMaxSpecies];
  uint8_t env[12][64];
  int64_t nb[12];   // by direction
  // by rank (neighbour-id order)
  double ea[12], de[12], rate[12];
  int64_t snb[12];
  int dir[12];
  // per-event sums (dE, logit MMM, logit MM2)
  double acc[12][3];
  // dense partials / solvent-baseline work list
  union {
    double part[12][kLanes][3];
    struct {
      uint16_t inc[kEnvPad];
      unsigned long long mask[12];
      int prefix[13];
      int start[13];
      double part[kPass][3];
    } sp;
  } u;
};

// Replica state of the persistent KMC kernel (lane 0 reads and writes it in phase 4).
struct ReplicaLds {
  uint64_t steps, max_steps, seed;
  double time, energy, abse, T, beta, cv, cs, smass;
  double com[3];
  int64_t vac, disp[3];
  uint32_t rid;
  int flags, esc, solvent, done;
};

// Byte offsets of the tables in dynamic LDS (computed on the host).
struct DynLayout {
  int w, b1, b2, base, lists, offs, bytes;
  int b1_lds, b2_lds;
};

struct Tabs {  // resolved table pointers (LDS or global)
  const double *w, *b1, *b2, *base;
  const uint32_t *state, *mmm, *mm2;  // dense lists
  const uint32_t *inc;                // solvent-baseline entries
  const uint16_t *inc_off;
};

template <bool kSolvent>
__device__ Tabs LoadJumpTables(const KTables &t, JumpLds &L, char *dyn, const DynLayout &lay) {
  const int tid = threadIdx.x;
  for (int i = tid; i < 12 * kEnvSlots; i += kThreads) L.off[i] = t.off[i];
  for (int i = tid; i < kEnvSlots; i += kThreads) L.mirror[i] = t.mirror[i];
  for (int i = tid; i < 3 * t.S * t.S; i += kThreads) L.tt[i] = t.tt[i];
  for (int i = tid; i < 6 * t.S; i += kThreads) L.par[i] = t.par[i];
  Tabs T{};
  double *w = reinterpret_cast<double *>(dyn + lay.w);
  for (int i = tid; i < t.nw; i += kThreads) w[i] = t.w[i];
  T.w = w;
  if (lay.b1_lds) {
    double *p = reinterpret_cast<double *>(dyn + lay.b1);
    for (int i = tid; i < t.ne * t.F1; i += kThreads) p[i] = t.b1[i];
    T.b1 = p;
  } else {
    T.b1 = t.b1;
  }
  if (lay.b2_lds) {
    double *p = reinterpret_cast<double *>(dyn + lay.b2);
    for (int i = tid; i < t.ne * t.F2; i += kThreads) p[i] = t.b2[i];
    T.b2 = p;
  } else {
    T.b2 = t.b2;
  }
  if (kSolvent) {
    double *bs = reinterpret_cast<double *>(dyn + lay.base);
    for (int i = tid; i < 3 * t.S * t.S; i += kThreads) bs[i] = t.base[i];
    T.base = bs;
    uint32_t *in = reinterpret_cast<uint32_t *>(dyn + lay.lists);
    for (int i = tid; i < t.n_inc; i += kThreads) in[i] = t.inc[i];
    T.inc = in;
    uint16_t *io = reinterpret_cast<uint16_t *>(dyn + lay.offs);
    for (int i = tid; i <= kEnvSlots; i += kThreads) io[i] = t.inc_off[i];
    T.inc_off = io;
  } else {
    uint32_t *p = reinterpret_cast<uint32_t *>(dyn + lay.lists);
    for (int i = tid; i < t.n_state; i += kThreads) p[i] = t.state[i];
    T.state = p;
    p += t.n_state;
    for (int i = tid; i < t.n_mmm; i += kThreads) p[i] = t.mmm[i];
    T.mmm = p;
    p += t.n_mmm;
    if (t.kind == kQuartic)
      for (int i = tid; i < t.n_mm2; i += kThreads) p[i] = t.mm2[i];
    T.mm2 = p;
  }
  return T;
}

// Phase 1: species of the 12 environments; neighbour ids of the 12 jumps. 720 sites over 256
// threads: three independent loads per thread, issued together.
__device__ __forceinline__ void GatherEnv(const KTables &t, JumpLds &L, const uint8_t *occ, int n,
                                          int64_t vac, int32_t *err) {
  const XYZ c = ToXYZ(vac, n);
  constexpr int kTotal = 12 * kEnvSlots;
  int64_t s[3];
  uint8_t v[3];
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const int i = threadIdx.x + u * kThreads;
    if (i < kTotal) {
      int dx, dy, dz;
      Unpack(L.off[i], dx, dy, dz);
      s[u] = SiteAt(c, n, dx, dy, dz);
      v[u] = occ[s[u]];
    }
  }
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const int i = threadIdx.x + u * kThreads;
    if (i < kTotal) {
      const int k = i / kEnvSlots, j = i - k * kEnvSlots;
      uint8_t x = v[u];
      if (x >= t.S) {  // corrupt occupancy: flag it and keep every table index in range
        atomicOr(err, 4);
        x = 0;
      }
      L.env[k][j] = x;
      if (j == t.n_slot) L.nb[k] = s[u];
    }
  }
}

__device__ __forceinline__ double WIndex(const double *w, int label, int a, int c1, int c2, int S) {
  if (label == 0) return w[a];
  if (label < 4) return w[S + (label - 1) * S * S + a * S + c1];
  return w[S + 3 * S * S + (label - 4) * S * S * S + (a * S + c1) * S + c2];
}

__device__ __forceinline__ double StateTerm(const double *w, const uint8_t *e, uint32_t d, int M,
                                            int X, int S) {
  const int label = d & 7;
  const int o1 = (d >> 4) & 63, o2 = (d >> 10) & 63;
  const int c1 = label ? e[o1] : 0;
  const int c2 = label >= 4 ? e[o2] : 0;
  const double v = WIndex(w, label, M, c1, c2, S) - WIndex(w, label, X, c1, c2, S);
  return (d & 8) ? -v : v;
}

__device__ __forceinline__ double OrbitTerm(const JumpLds &L, const double *bm, const uint8_t *e,
                                            uint32_t d, int S, bool backward) {
  const int base = d & 4095, kind = (d >> 12) & 3;
  int s0 = (d >> 14) & 63, s1 = (d >> 20) & 63;
  if (backward) {
    s0 = L.mirror[s0];
    s1 = kind ? L.mirror[s1] : 0;
  }
  const int c0 = e[s0];
  const int c1 = kind ? e[s1] : 0;
  const int ty = L.tt[(kind * S + c0) * S + c1];
  // A vacancy off the jump endpoints has no one-hot type (the reference throws).
  return ty < 0 ? __builtin_nan("") : bm[base + ty];
}

// Phase 2, dense: per-lane partial sums of every cluster's weight, then per-event sums.
__device__ void DenseClusters(const KTables &t, JumpLds &L, const Tabs &T) {
  const int tid = threadIdx.x;
  if (tid < 12 * kLanes) {
    const int k = tid / kLanes, lane = tid - k * kLanes;
    const uint8_t *e = L.env[k];
    const int S = t.S, X = t.x_code, M = e[t.n_slot];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int c = lane;
    for (; c + 3 * kLanes < t.n_state; c += 4 * kLanes) {
      a0 += StateTerm(T.w, e, T.state[c], M, X, S);
      a1 += StateTerm(T.w, e, T.state[c + kLanes], M, X, S);
      a2 += StateTerm(T.w, e, T.state[c + 2 * kLanes], M, X, S);
      a3 += StateTerm(T.w, e, T.state[c + 3 * kLanes], M, X, S);
    }
    for (; c < t.n_state; c += kLanes) a0 += StateTerm(T.w, e, T.state[c], M, X, S);
    const double de = (a0 + a1) + (a2 + a3);
    double la = 0.0, lb = 0.0;
    const int row = L.tt[M * S];  // one-hot rank of the migrating species (-1 for X)
    if (t.kind == kE0 || t.kind == kQuartic) {
      if (row < 0) {
        la = __builtin_nan("");
      } else {
        const double *bm = T.b1 + static_cast<size_t>(row) * t.F1;
        a0 = a1 = a2 = a3 = 0.0;
        c = lane;
        for (; c + 3 * kLanes < t.n_mmm; c += 4 * kLanes) {
          a0 += OrbitTerm(L, bm, e, T.mmm[c], S, false);
          a1 += OrbitTerm(L, bm, e, T.mmm[c + kLanes], S, false);
          a2 += OrbitTerm(L, bm, e, T.mmm[c + 2 * kLanes], S, false);
          a3 += OrbitTerm(L, bm, e, T.mmm[c + 3 * kLanes], S, false);
        }
        for (; c < t.n_mmm; c += kLanes) a0 += OrbitTerm(L, bm, e, T.mmm[c], S, false);
        la = (a0 + a1) + (a2 + a3);
      }
    }
    if (t.kind == kQuartic && row >= 0) {
      const double *bm = T.b2 + static_cast<size_t>(row) * t.F2;
      a0 = a1 = a2 = a3 = 0.0;
      c = lane;
      for (; c + kLanes < t.n_mm2; c += 2 * kLanes) {
        const uint32_t x = T.mm2[c], y = T.mm2[c + kLanes];
        a0 += OrbitTerm(L, bm, e, x, S, false);
        a1 += OrbitTerm(L, bm, e, x, S, true);
        a2 += OrbitTerm(L, bm, e, y, S, false);
        a3 += OrbitTerm(L, bm, e, y, S, true);
      }
      for (; c < t.n_mm2; c += kLanes) {
        a0 += OrbitTerm(L, bm, e, T.mm2[c], S, false);
        a1 += OrbitTerm(L, bm, e, T.mm2[c], S, true);
      }
      lb = (a0 + a1) + (a2 + a3);
    }
    L.u.part[k][lane][0] = de;
    L.u.part[k][lane][1] = la;
    L.u.part[k][lane][2] = lb;
  }
  __syncthreads();
  if (tid < 12) {
    double de = 0.0, la = 0.0, lb = 0.0;
#pragma unroll 7
    for (int l = 0; l < kLanes; ++l) {
      de += L.u.part[tid][l][0];
      la += L.u.part[tid][l][1];
      lb += L.u.part[tid][l][2];
    }
    L.acc[tid][0] = de;
    L.acc[tid][1] = la;
    L.acc[tid][2] = lb;
  }
}

// Change of one touched cluster relative to the all-solvent environment.
__device__ __forceinline__ void SolventEntry(const JumpLds &L, const Tabs &T, const KTables &t,
                                             const uint8_t *e, uint32_t d, int j, int cj, int sv,
                                             int M, int row, double &de, double &la, double &lb) {
  const int S = t.S;
  const int type = d & 3, other = (d >> 6) & 63;
  const int co = other == 63 ? sv : e[other];
  if (other != 63 && other < j && co != sv) return;  // evaluated at the lower non-solvent slot
  if (type == 0) {
    const int label = (d >> 2) & 7;
    const int X = t.x_code;
    const double v = (WIndex(T.w, label, M, cj, co, S) - WIndex(T.w, label, X, cj, co, S)) -
                     (WIndex(T.w, label, M, sv, sv, S) - WIndex(T.w, label, X, sv, sv, S));
    de += (d & 32) ? -v : v;
    return;
  }
  const int kind = (d >> 2) & 3;
  const int base = d >> 12;
  const bool second = (d & 32) != 0;
  const int c0 = second ? co : cj, c1 = kind ? (second ? cj : co) : 0;
  const int ty = L.tt[(kind * S + c0) * S + c1];
  const int tp = L.tt[(kind * S + sv) * S + (kind ? sv : 0)];
  if (type == 1) {
    const double *bm = T.b1 + static_cast<size_t>(row) * t.F1;
    la += ty < 0 ? __builtin_nan("") : bm[base + ty] - bm[base + tp];
  } else {
    const double *bm = T.b2 + static_cast<size_t>(row) * t.F2;
    lb += ty < 0 ? __builtin_nan("") : bm[base + ty] - bm[base + tp];
  }
}

// Phase 2, solvent baseline: list the environment entries (event k, slot j) that hold a
// non-solvent species, then 8 lanes per entry add the changes of the clusters touching it.
__device__ void SolventClusters(const KTables &t, JumpLds &L, const Tabs &T, int sv) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  bool flag[3];
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const int idx = tid + u * kThreads;
    bool f = false;
    if (idx < 12 * kEnvSlots) {
      const int k = idx / kEnvSlots, j = idx - k * kEnvSlots;
      f = j != t.v_slot && j != t.n_slot && L.env[k][j] != sv;
    }
    flag[u] = f;
    const unsigned long long m = __ballot(f);
    if (lane == 0) L.u.sp.mask[wave + 4 * u] = m;
  }
  __syncthreads();
  if (tid == 0) {
    int s = 0;
    for (int c = 0; c < 12; ++c) {
      L.u.sp.prefix[c] = s;
      s += __popcll(L.u.sp.mask[c]);
    }
    L.u.sp.prefix[12] = s;
  }
  __syncthreads();
  if (tid <= 12) {  // first list position of event k's entries
    const int e0 = tid * kEnvSlots, c = e0 >> 6, b = e0 & 63;
    L.u.sp.start[tid] = c >= 12 ? L.u.sp.prefix[12]
                                : L.u.sp.prefix[c] + __popcll(L.u.sp.mask[c] & ((1ull << b) - 1ull));
  }
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int u = 0; u < 3; ++u)
    if (flag[u]) {
      const int c = wave + 4 * u;
      L.u.sp.inc[L.u.sp.prefix[c] + __popcll(L.u.sp.mask[c] & below)] =
          static_cast<uint16_t>(tid + u * kThreads);
    }
  if (tid < 12) L.acc[tid][0] = L.acc[tid][1] = L.acc[tid][2] = 0.0;
  __syncthreads();
  const int total = L.u.sp.prefix[12];
  const int grp = tid / kGroup, g = tid - grp * kGroup;
  for (int p0 = 0; p0 < total; p0 += kPass) {
    double de = 0.0, la = 0.0, lb = 0.0;
    const int i = p0 + grp;
    if (i < total) {
      const int idx = L.u.sp.inc[i];
      const int k = idx / kEnvSlots, j = idx - k * kEnvSlots;
      const uint8_t *e = L.env[k];
      const int M = e[t.n_slot];
      const int row = L.tt[M * t.S];
      const int cj = e[j];
      if (row < 0) {
        de = __builtin_nan("");
      } else {
        const int q1 = T.inc_off[j + 1];
        for (int q = T.inc_off[j] + g; q < q1; q += kGroup)
          SolventEntry(L, T, t, e, T.inc[q], j, cj, sv, M, row, de, la, lb);
      }
    }
    // sum over the 8 lanes of the group (fixed butterfly order)
#pragma unroll
    for (int m = kGroup / 2; m > 0; m >>= 1) {
      de += __shfl_xor(de, m);
      la += __shfl_xor(la, m);
      lb += __shfl_xor(lb, m);
    }
    if (g == 0) {
      L.u.sp.part[grp][0] = de;
      L.u.sp.part[grp][1] = la;
      L.u.sp.part[grp][2] = lb;
    }
    __syncthreads();
    if (tid < 12) {
      const int lo = max(p0, L.u.sp.start[tid]), hi = min(p0 + kPass, L.u.sp.start[tid + 1]);
      for (int q = lo; q < hi; ++q) {
        L.acc[tid][0] += L.u.sp.part[q - p0][0];
        L.acc[tid][1] += L.u.sp.part[q - p0][1];
        L.acc[tid][2] += L.u.sp.part[q - p0][2];
      }
    }
    __syncthreads();
  }
  if (tid < 12) {  // all-solvent sums of this (solvent, migrating species)
    const int M = L.env[tid][t.n_slot];
    const double *b = T.base + (static_cast<size_t>(sv) * t.S + M) * 3;
    L.acc[tid][0] += b[0];
    L.acc[tid][1] += b[1];
    L.acc[tid][2] += b[2];
  }
}

// Phase 3: Ea, dE and rate per jump, stored at the jump's rank among the 12 neighbour ids.
__device__ __forceinline__ void FinishEvents(const KTables &t, JumpLds &L, double beta,
                                             int32_t *err) {
  const int k = threadIdx.x;
  if (k >= 12) return;
  const double de = L.acc[k][0], la = L.acc[k][1], lb = L.acc[k][2];
  const int M = L.env[k][t.n_slot];
  const double *p = L.par + 6 * M;
  double ea = 0.0;
  if (t.kind == kE0) {
    // E0 = exp(logit * sigma_e0 + mu_e0); Ea = max(0, E0 + dE/2) (...E0.cpp:147-158)
    const double e0 = exp((p[0] + la) * p[2] + p[1]);
    ea = fmax(0.0, e0 + de / 2);
  } else if (t.kind == kQuartic) {
    // Quartic barrier (...Quartic.cpp:266-275)
    const double D = exp((p[0] + la) * p[2] + p[1]);
    const double Ks = exp((p[3] + lb) * p[5] + p[4]);
    const double b = 4 * de / (D * D * D);
    const double a = Ks / (4 * D * D);
    const double c = (9 * b * b - 16 * a * a * D * D) / (32 * a);
    const double delta = sqrt(fabs(9 * b * b - 32 * a * c));
    ea = (3 * b + delta) * (3 * b + delta) * (3 * b * b - 16 * a * c + b * delta) / (a * a * a) /
         2048;
  }
  if (!(ea == ea) || !(de == de) || !(la == la)) atomicOr(err, 1);  // unknown cluster type
  int r = 0;
  const int64_t mine = L.nb[k];
#pragma unroll
  for (int j = 0; j < 12; ++j) r += L.nb[j] < mine;
  L.ea[r] = ea;
  L.de[r] = de;
  L.rate[r] = exp(-ea * beta);  // JumpEvent forward rate (JumpEvent.cpp:6-13)
  L.snb[r] = mine;
  L.dir[r] = k;
}

template <bool kSolvent>
__device__ __forceinline__ void EvaluateJumps(const KTables &t, JumpLds &L, const Tabs &T,
                                              const uint8_t *occ, int n, int64_t vac, int sv,
                                              double beta, int32_t *err) {
  GatherEnv(t, L, occ, n, vac, err);
  __syncthreads();
  if (kSolvent) SolventClusters(t, L, T, sv);
  else DenseClusters(t, L, T);
  __syncthreads();
  FinishEvents(t, L, beta, err);
  __syncthreads();
}

template <bool kSolvent>
__global__ __launch_bounds__(kThreads) void EvalJumpsKernel(KTables t, DynLayout lay,
                                                            const uint8_t *occ,
                                                            const int64_t *off,
                                                            const int32_t *ncell,
                                                            const int64_t *vac, int64_t *nb,
                                                            double *ea, double *de, int32_t *err) {
  __shared__ JumpLds L;
  extern __shared__ char dyn[];
  const int r = blockIdx.x;
  if (!ValidSite(ncell[r], vac[r])) {
    if (threadIdx.x == 0) atomicOr(err, 8);
    return;
  }
  const Tabs T = LoadJumpTables<kSolvent>(t, L, dyn, lay);
  __syncthreads();
  EvaluateJumps<kSolvent>(t, L, T, occ + off[r], ncell[r], vac[r], t.solvent_default, 0.0, err);
  const int k = threadIdx.x;
  if (k < 12) {
    const int64_t o = static_cast<int64_t>(r) * 12 + k;
    nb[o] = L.snb[k];
    ea[o] = L.ea[k];
    de[o] = L.de[k];
  }
}

// TimeTemperatureInterpolator::GetTemperature (TimeTemperatureInterpolator.cpp:44-65)
__device__ double ScheduleTemperature(const double *tt, const double *tv, int n, double time) {
  int i = 0;
  while (i < n && tt[i] < time) ++i;
  if (i == n) return tv[n - 1];
  if (i == 0 && time <= tt[0]) return tv[0];
  return tv[i - 1] + ((time - tt[i - 1]) / (tt[i] - tt[i - 1])) * (tv[i] - tv[i - 1]);
}

// Phase 4 of one KMC step (lane 0): KineticMcFirstAbstract::OneStepSimulation after
// BuildEventList (KineticMcAbstract.cpp:147-181).
__device__ void SelectAndApply(const KTables &t, const KmcDev &k, JumpLds &L, ReplicaLds &Q,
                               uint8_t *occ, int r, uint64_t trace_index) {
  double total = 0.0;
  for (int j = 0; j < 12; ++j) total += L.rate[j];
  const uint64_t steps = Q.steps;
  const U4 u = Philox4x32_10(Q.seed, steps, Q.rid, 0u);
  const double u1 = UnitOpenLeft(Join64(u.x, u.y));
  const double u2 = UnitOpenRight(Join64(u.z, u.w));
  double dt = -log(u1) / total / kPrefactor;  // KineticMcFirstOmp::CalculateTime (:79-82)
  if (Q.flags & 1)  // RateCorrector::GetTimeCorrectionFactor (RateCorrector.hpp:17-24)
    dt *= Q.cv / (1.64 * exp(-(0.66 / kBoltzmann / Q.T - 0.7))) / (1 - 13 * Q.cs);
  int sel = 11;  // SelectEvent: first cumulative probability >= u2, else the last event
  double cum = 0.0;
  for (int j = 0; j < 12; ++j) {
    cum += L.rate[j] / total;
    if (!(cum < u2)) {
      sel = j;
      break;
    }
  }
  if (Q.flags & 2) {  // IsEscaped (KineticMcAbstract.cpp:192-223)
    bool all = true;
    for (int j = 0; j < kEnvSlots; ++j)
      if (((t.vac_shell_mask >> j) & 1ull) && L.env[0][j] != Q.solvent) all = false;
    if (all) {
      Q.esc = 1;
      Q.steps = Q.max_steps + 1;
    }
  }
  const int dir = L.dir[sel];
  const int64_t jump = L.snb[sel];
  const double de = L.de[sel];
  if (trace_index < k.trace_cap) {
    const uint64_t o = static_cast<uint64_t>(r) * k.trace_cap + trace_index;
    k.tr_sel[o] = jump;
    k.tr_ea[o] = L.ea[sel];
    k.tr_de[o] = de;
    k.tr_time[o] = Q.time + dt;
  }
  Q.time += dt;
  Q.energy += de;
  Q.abse += de;
  int dx, dy, dz;
  Unpack(L.off[dir * kEnvSlots + t.n_slot], dx, dy, dz);
  const int mig = L.env[dir][t.n_slot];
  if ((Q.flags & 4) && mig != Q.solvent) {  // solute COM (KineticMcAbstract.cpp:163-177)
    const double wgt = k.mass[mig] / Q.smass;
    Q.com[0] -= dx * kHalfLattice * wgt;
    Q.com[1] -= dy * kHalfLattice * wgt;
    Q.com[2] -= dz * kHalfLattice * wgt;
  }
  Q.disp[0] += dx;
  Q.disp[1] += dy;
  Q.disp[2] += dz;
  occ[Q.vac] = static_cast<uint8_t>(mig);  // Config::LatticeJump (Config.cpp:431-456)
  occ[jump] = static_cast<uint8_t>(t.x_code);
  Q.steps += 1;
  Q.vac = jump;
  if (Q.steps > Q.max_steps || Q.esc) Q.done = 1;
}

template <bool kSolvent>
__global__ __launch_bounds__(kThreads) void KmcKernel(KTables t, DynLayout lay, KmcDev k,
                                                      uint64_t nsteps, uint64_t trace_base) {
  __shared__ JumpLds L;
  __shared__ ReplicaLds Q;
  extern __shared__ char dyn[];
  const int r = blockIdx.x;
  const int tid = threadIdx.x;
  const Tabs T = LoadJumpTables<kSolvent>(t, L, dyn, lay);
  uint8_t *occ = k.occ + k.off[r];
  const int n = k.ncell[r];
  if (tid == 0) {
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
  }
  __syncthreads();
  const int sv = Q.solvent;
  for (uint64_t it = 0; it < nsteps; ++it) {
    if (Q.done) break;
    if (tid == 0 && k.n_tt > 0) {
      Q.T = ScheduleTemperature(k.tt_time, k.tt_temp, k.n_tt, Q.time);
      Q.beta = 1.0 / kBoltzmann / Q.T;
    }
    EvaluateJumps<kSolvent>(t, L, T, occ, n, Q.vac, sv, Q.beta, k.error);
    if (tid == 0) SelectAndApply(t, k, L, Q, occ, r, trace_base + it);
    __syncthreads();
  }
  if (tid == 0) {
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
}

// Dynamic LDS plan: weight tables go to LDS while the workgroup stays within its share of the
// CU's LDS (up to 4 resident workgroups per CU when the batch fills the chip).
DynLayout PlanLds(const KTables &t, int32_t R, bool solvent) {
  static int cus = 0;
  static size_t lds_cu = 0, lds_block = 0;
  if (!cus) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) {
      cus = p.multiProcessorCount;
      lds_cu = p.maxSharedMemoryPerMultiProcessor ? p.maxSharedMemoryPerMultiProcessor : 65536;
      lds_block = p.sharedMemPerBlock;
    } else {
      cus = 256;
      lds_cu = lds_block = 65536;
    }
  }
  const int per_cu = std::min(4, std::max(1, (R + cus - 1) / cus));
  const size_t stat = sizeof(JumpLds) + sizeof(ReplicaLds) + 256;
  const size_t share = std::min(lds_block, lds_cu / per_cu);
  auto al = [](size_t x) { return (x + 15) & ~size_t(15); };
  DynLayout lay{};
  size_t at = 0;
  lay.w = static_cast<int>(at);
  at = al(at + sizeof(double) * t.nw);
  if (solvent) {
    lay.base = static_cast<int>(at);
    at = al(at + sizeof(double) * 3 * t.S * t.S);
    lay.lists = static_cast<int>(at);
    at = al(at + sizeof(uint32_t) * t.n_inc);
    lay.offs = static_cast<int>(at);
    at = al(at + sizeof(uint16_t) * (kEnvSlots + 1));
  } else {
    lay.lists = static_cast<int>(at);
    at = al(at + sizeof(uint32_t) * (t.n_state + t.n_mmm + (t.kind == kQuartic ? t.n_mm2 : 0)));
  }
  const size_t b1 = sizeof(double) * static_cast<size_t>(t.ne) * t.F1;
  const size_t b2 = sizeof(double) * static_cast<size_t>(t.ne) * t.F2;
  if ((t.kind == kE0 || t.kind == kQuartic) && stat + at + b1 <= share) {
    lay.b1_lds = 1;
    lay.b1 = static_cast<int>(at);
    at = al(at + b1);
  }
  if (t.kind == kQuartic && lay.b1_lds && stat + at + b2 <= share) {
    lay.b2_lds = 1;
    lay.b2 = static_cast<int>(at);
    at = al(at + b2);
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
                           double *ea, double *de, int32_t *err, hipStream_t s, bool solvent) {
  if (R <= 0) return hipSuccess;
  const DynLayout lay = PlanLds(t, R, solvent);
  auto go = [&](auto kernel) {
    hipError_t e = AllowLds(kernel, lay.bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kernel, dim3(R), dim3(kThreads), lay.bytes, s, t, lay, occ, off, ncell,
                       vac, nb, ea, de, err);
    return hipGetLastError();
  };
  return solvent ? go(EvalJumpsKernel<true>) : go(EvalJumpsKernel<false>);
}

hipError_t LaunchKmc(const KTables &t, const KmcDev &k, uint64_t steps, uint64_t trace_base,
                     hipStream_t s, bool solvent) {
  if (k.R <= 0 || steps == 0) return hipSuccess;
  const DynLayout lay = PlanLds(t, k.R, solvent);
  auto go = [&](auto kernel) {
    hipError_t e = AllowLds(kernel, lay.bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kernel, dim3(k.R), dim3(kThreads), lay.bytes, s, t, lay, k, steps,
                       trace_base);
    return hipGetLastError();
  };
  return solvent ? go(KmcKernel<true>) : go(KmcKernel<false>);
}

}  // namespace lmc
