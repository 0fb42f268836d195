// gfx950 kernels of the swap / site energy change (EnergyChangePredictorPairSite), the
// CanonicalMc driver and the full-lattice cluster-expansion energy (EnergyPredictor).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "lmc_dev_common.h"
#include "lmc_device.h"
#include "lmc_rng.h"

namespace lmc {
namespace {
// ---------------------------------------------------------------------------------------------
// Swap / site energy change (EnergyChangePredictorPairSite.cpp:69-192).
struct SiteLds {
  uint32_t cl[kMaxSiteClusters];
  int32_t off[kSiteSlots];
  double w[kMaxW];
  uint8_t env[4][48];           // environments of up to two swaps (p, q of each)
  int slot[4];                  // partner slots (q around p, p around q) of each swap
  double part[kThreads / 64];   // per-wave partial sums
  double part2[kThreads / 64];  // the same for the second swap
};

__device__ void LoadSiteTables(const KTables &t, SiteLds &L) {
  for (int i = threadIdx.x; i < t.n_site; i += blockDim.x) L.cl[i] = t.site_cl[i];
  for (int i = threadIdx.x; i < kSiteSlots; i += blockDim.x) L.off[i] = t.site_off[i];
  for (int i = threadIdx.x; i < t.nw; i += blockDim.x) L.w[i] = t.w[i];
}

__device__ __forceinline__ double SiteWeight(const KTables &t, const SiteLds &L, const uint8_t *e,
                                             uint32_t d, int centre) {
  const int S = t.S, label = d & 7, o1 = (d >> 4) & 63, o2 = (d >> 10) & 63;
  if (label == 0) return L.w[centre];
  if (label < 4) return L.w[S + (label - 1) * S * S + centre * S + e[o1]];
  return L.w[S + 3 * S * S + (label - 4) * S * S * S + (centre * S + e[o1]) * S + e[o2]];
}

// Sum over clusters containing site x of w(x -> en) - w(x -> eo), skipping clusters that
// contain env slot `skip` (the partner of a coupled swap): this thread's share.
__device__ double SitePartial(const KTables &t, const SiteLds &L, int which, int eo, int en,
                              int skip) {
  const uint8_t *e = L.env[which];
  double acc = 0.0;
  for (int c = threadIdx.x; c < t.n_site; c += blockDim.x) {
    const uint32_t d = L.cl[c];
    const int o1 = (d >> 4) & 63, o2 = (d >> 10) & 63;
    if (skip >= 0 && (o1 == skip || o2 == skip)) continue;
    acc += SiteWeight(t, L, e, d, en) - SiteWeight(t, L, e, d, eo);
  }
  return acc;
}

// Fixed-order block sum, valid in thread 0: DPP sums within each wave, then the wave sums in
// wave order. One barrier; L.part is rewritten only after the caller's next barrier.
__device__ double BlockSum(SiteLds &L, double v) {
  v = WaveSum(v);
  if ((threadIdx.x & 63) == 0) L.part[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kThreads / 64; ++w) r += L.part[w];
  return r;
}

// Species in the 43-site environment of x (slot i read by lane i of the calling wave).
__device__ void GatherSite(const SiteLds &L, uint8_t *dst, const uint8_t *occ, int n,
                           const XYZ &c, int S, int lane) {
  if (lane < kSiteSlots) {
    int dx, dy, dz;
    Unpack(L.off[lane], dx, dy, dz);
    const uint8_t v = occ[SiteAt(c, n, dx, dy, dz)];
    dst[lane] = v < S ? v : 0;  // codes are validated by the callers; keep LDS indices in range
  }
}

// Periodic minimum image of a coordinate difference on the 2n axis, in [-n, n).
__device__ __forceinline__ int MinImage(int v, int n) {  // |v| < 2n
  const int m = 2 * n;
  v += v < 0 ? m : 0;
  return v >= n ? v - m : v;
}

// Slot of b in the environment of a: the first slot whose offset is the minimum-image
// difference b - a, or -1 beyond the third shell. Lane i of the calling wave tests slot i.
__device__ int PartnerSlot(const SiteLds &L, int n, const XYZ &a, const XYZ &b, int lane) {
  const int dx = MinImage(b.X - a.X, n), dy = MinImage(b.Y - a.Y, n), dz = MinImage(b.Z - a.Z, n);
  bool hit = false;
  if (lane >= 1 && lane < kSiteSlots) {
    int x, y, z;
    Unpack(L.off[lane], x, y, z);
    hit = x == dx && y == dy && z == dz;
  }
  const uint64_t m = __ballot(hit);
  return m ? __builtin_ctzll(m) : -1;
}

// Environments and partner slots of a swap p <-> q: wave 0 gathers around p, wave 1 around q.
// Results in L.env and L.slot after the caller's barrier.
__device__ void GatherPair(SiteLds &L, const uint8_t *occ, int n, int64_t p, int64_t q, int S) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (wave > 1) return;
  const XYZ a = ToXYZ(wave ? q : p, n), b = ToXYZ(wave ? p : q, n);
  GatherSite(L, L.env[wave], occ, n, a, S, lane);
  const int s = PartnerSlot(L, n, a, b, lane);
  if (lane == 0) L.slot[wave] = s;
}

// dE of swapping species ep at p and eq at q, from the environments of GatherPair. Far pairs:
// dE(p) + dE(q) on the unswapped config (...WithoutCoupling :136-146). Near pairs: clusters
// containing both keep their multiset and are skipped (...WithCoupling :96-134). All threads
// call; the result is valid in thread 0.
__device__ double SwapDe(const KTables &t, SiteLds &L, int ep, int eq) {
  const double v =
      SitePartial(t, L, 0, ep, eq, L.slot[0]) + SitePartial(t, L, 1, eq, ep, L.slot[1]);
  return BlockSum(L, v);
}

__global__ __launch_bounds__(kThreads) void SwapDeKernel(KTables t, const uint8_t *occ,
                                                         const int64_t *off, const int32_t *ncell,
                                                         const int64_t *pairs, double *de) {
  __shared__ SiteLds L;
  const int r = blockIdx.x;
  const int64_t p = pairs[2 * r], q = pairs[2 * r + 1];
  const uint8_t *o = occ + off[r];
  if (!ValidSite(ncell[r], p) || !ValidSite(ncell[r], q) || o[p] >= t.S || o[q] >= t.S) {
    if (threadIdx.x == 0) de[r] = __builtin_nan("");
    return;
  }
  const int ep = o[p], eq = o[q];
  if (ep == eq) {
    if (threadIdx.x == 0) de[r] = 0.0;
    return;
  }
  LoadSiteTables(t, L);
  __syncthreads();
  GatherPair(L, o, ncell[r], p, q, t.S);
  __syncthreads();
  const double v = SwapDe(t, L, ep, eq);
  if (threadIdx.x == 0) de[r] = v;
}

// GetDeFromLatticeIdSite (EnergyChangePredictorPairSite.cpp:154-192): site sites[r] changes from
// species from[r] (null: the species it holds) to codes[r], the rest of the occupancy as is.
__global__ __launch_bounds__(kThreads) void SiteDeKernel(KTables t, const uint8_t *occ,
                                                         const int64_t *off, const int32_t *ncell,
                                                         const int64_t *sites, const uint8_t *from,
                                                         const uint8_t *codes, double *de) {
  __shared__ SiteLds L;
  const int r = blockIdx.x;
  const uint8_t *o = occ + off[r];
  if (!ValidSite(ncell[r], sites[r]) || o[sites[r]] >= t.S || codes[r] >= t.S ||
      (from && from[r] >= t.S)) {
    if (threadIdx.x == 0) de[r] = __builtin_nan("");
    return;
  }
  LoadSiteTables(t, L);
  __syncthreads();
  const int eo = from ? from[r] : o[sites[r]], en = codes[r];
  if (eo == en) {
    if (threadIdx.x == 0) de[r] = 0.0;
    return;
  }
  if (threadIdx.x < 64)
    GatherSite(L, L.env[0], o, ncell[r], ToXYZ(sites[r], ncell[r]), t.S, threadIdx.x);
  __syncthreads();
  const double v = BlockSum(L, SitePartial(t, L, 0, eo, en, -1));
  if (threadIdx.x == 0) de[r] = v;
}

// SimulatedAnnealing::UpdateTemperature (SimulatedAnnealing.cpp:99-133) with maximum_steps N:
// the window, trigger and cooldown lengths and the per-step decay factor exp(-3/N).
struct AnnealSchedule {
  uint64_t window, trigger, cooldown;
  double decay;
};
__device__ AnnealSchedule MakeAnnealSchedule(uint64_t N) {
  AnnealSchedule s;
  s.window = max(1ull, static_cast<unsigned long long>(N * 0.001));
  s.trigger = max(1ull, static_cast<unsigned long long>(N * 0.05));
  s.cooldown = max(1ull, static_cast<unsigned long long>(N * 0.10));
  s.decay = exp(-3.0 / static_cast<double>(max(1ull, static_cast<unsigned long long>(N))));
  return s;
}

// The temperature update after the trial of step `steps` with the energy after the trial.
__device__ void AnnealUpdate(CmcAnneal &A, const AnnealSchedule &s, bool accepted,
                             uint64_t steps, double energy) {
  ++A.window_trials;
  if (accepted) {
    ++A.window_accepts;
    if (energy < A.recent_best - 1e-4) {
      A.recent_best = energy;
      A.last_improvement = steps;
    }
  }
  if (A.window_trials >= s.window) {
    const double acc = static_cast<double>(A.window_accepts) / static_cast<double>(A.window_trials);
    if (acc > 0.50) A.T *= 0.99;
    A.window_trials = 0;
    A.window_accepts = 0;
  }
  // the acceptance estimate last: the division only runs when the other conditions hold
  if (A.reheats < 5 && steps - A.last_improvement >= s.trigger &&
      steps - A.last_reheat >= s.cooldown &&
      (A.window_trials > 0u ? static_cast<double>(A.window_accepts) /
                                  static_cast<double>(A.window_trials)
                            : 1.0) < 0.05) {
    A.T *= 1.10;
    A.last_improvement = steps;
    A.last_reheat = steps;
    A.recent_best = energy;
    ++A.reheats;
  }
  A.T *= s.decay;
}

constexpr uint32_t kMaxPairAttempts = 1u << 20;

__device__ __forceinline__ int64_t ReadLane64(int64_t v, int lane) {
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), lane));
  const uint32_t hi =
      static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), lane));
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ __forceinline__ XYZ ReadLaneXYZ(const XYZ &v, int lane) {
  return {__builtin_amdgcn_readlane(v.X, lane), __builtin_amdgcn_readlane(v.Y, lane),
          __builtin_amdgcn_readlane(v.Z, lane)};
}

// ToXYZ with 32-bit divisions: every box has 4 n^3 <= 4e9 < 2^32 sites (n <= 1000).
__device__ __forceinline__ XYZ ToXYZ32(uint32_t id, uint32_t n) {
  const uint32_t plane = 2u * n * n;
  const uint32_t X = id / plane;
  const uint32_t rest = id - X * plane;
  const uint32_t Y = rest / n;
  const uint32_t Zh = rest - Y * n;
  return {static_cast<int>(X), static_cast<int>(Y), static_cast<int>(2 * Zh + ((X + Y) & 1))};
}

// Draw attempt att of step `steps` (Philox counter 2 att + 1): the two sites and their
// coordinates. CanonicalMcAbstract.cpp:43-101 redraws until the species differ; lane l of a
// wave holds attempt base + l, so the lowest lane of the first round with a hit is the
// sequential answer.
struct PairDraw {
  int64_t a, b;
  XYZ xa, xb;
};
__device__ __forceinline__ PairDraw DrawAttempt(const CmcDev &c, int r, uint64_t steps,
                                                uint32_t att, uint64_t N, int n) {
  const U4 u = Philox4x32_10(c.seed[r], steps, c.rid[r], 2u * att + 1u);
  PairDraw d;
  d.a = static_cast<int64_t>(MulHi64(Join64(u.x, u.y), N));
  d.b = static_cast<int64_t>(MulHi64(Join64(u.z, u.w), N));
  d.xa = ToXYZ32(static_cast<uint32_t>(d.a), static_cast<uint32_t>(n));
  d.xb = ToXYZ32(static_cast<uint32_t>(d.b), static_cast<uint32_t>(n));
  return d;
}

// Metropolis uniform of step `steps`.
__device__ __forceinline__ double TrialUniform(const CmcDev &c, int r, uint64_t steps) {
  const U4 u = Philox4x32_10(c.seed[r], steps, c.rid[r], 0xFFFFFFFEu);
  return UnitOpenRight(Join64(u.x, u.y));
}

// Cluster `tid` of the site table, decoded once: the weight index of centre species x with the
// other sites' species e1, e2 is base + x kx + e1 k1 + e2 k2 (SiteWeight's three cases).
struct SiteCluster {
  int base, kx, k1, k2;
  int o1, o2;      // environment slots read (0 where unused)
  int r1, r2;      // raw slots (63 = none), for the partner test
  bool live;
};
__device__ SiteCluster DecodeSiteCluster(const KTables &t, const SiteLds &L, int i) {
  SiteCluster q{};
  q.live = i < t.n_site;
  if (!q.live) return q;
  const uint32_t d = L.cl[i];
  const int S = t.S, label = d & 7;
  q.r1 = (d >> 4) & 63;
  q.r2 = (d >> 10) & 63;
  if (label == 0) {
    q.kx = 1;
  } else if (label < 4) {
    q.base = S + (label - 1) * S * S;
    q.kx = S;
    q.k1 = 1;
    q.o1 = q.r1;
  } else {
    q.base = S + 3 * S * S + (label - 4) * S * S * S;
    q.kx = S * S;
    q.k1 = S;
    q.k2 = 1;
    q.o1 = q.r1;
    q.o2 = q.r2;
  }
  return q;
}
// w(x -> en) - w(x -> eo) of the cluster in environment `which`, 0 when it contains slot skip.
__device__ __forceinline__ double ClusterDelta(const SiteLds &L, const SiteCluster &q, int which,
                                               int eo, int en, int skip) {
  if (!q.live || (skip >= 0 && (q.r1 == skip || q.r2 == skip))) return 0.0;
  const uint8_t *e = L.env[which];
  const int rest = q.base + e[q.o1] * q.k1 + e[q.o2] * q.k2;
  return L.w[rest + en * q.kx] - L.w[rest + eo * q.kx];
}

// One trial's pair as found by its wave: sites, coordinates, species, and the winning lane of
// the first draw round (-1: found in a later round; -2: none within kMaxPairAttempts).
struct TrialLds {
  int64_t p, q;
  XYZ xp, xq;
  int ep, eq, win;
};

// Finds the trial's pair among the wave's draw attempts (cand: attempts lane of step `step`,
// further rounds drawn here), gathers both environments into L.env[e0], L.env[e0 + 1] and both
// partner slots, and leaves the result in T (lane 0 writes). Called by a whole wave.
__device__ void FindTrial(const CmcDev &c, int r, uint64_t step, const uint8_t *occ, uint64_t N,
                          int n, int S, int lane, int ox, int oy, int oz, PairDraw &cand,
                          SiteLds &L, int e0, TrialLds &T) {
  uint64_t m = __ballot(occ[cand.a] != occ[cand.b]);
  const bool first = m != 0;
  for (uint32_t base = 64; !m && base < kMaxPairAttempts; base += 64) {
    cand = DrawAttempt(c, r, step, base + static_cast<uint32_t>(lane), N, n);
    m = __ballot(occ[cand.a] != occ[cand.b]);
  }
  if (!m) {
    if (lane == 0) T.win = -2;
    return;
  }
  const int w = __builtin_ctzll(m);
  const int64_t p = ReadLane64(cand.a, w), q = ReadLane64(cand.b, w);
  const XYZ xp = ReadLaneXYZ(cand.xa, w), xq = ReadLaneXYZ(cand.xb, w);
  bool hq = false, hp = false;
  if (lane < kSiteSlots) {
    const uint8_t vp = occ[SiteAt(xp, n, ox, oy, oz)];
    const uint8_t vq = occ[SiteAt(xq, n, ox, oy, oz)];
    L.env[e0][lane] = vp < S ? vp : 0;  // codes are validated at creation
    L.env[e0 + 1][lane] = vq < S ? vq : 0;
    if (lane >= 1) {  // partner slots: the first offset equal to the minimum image
      hq = ox == MinImage(xq.X - xp.X, n) && oy == MinImage(xq.Y - xp.Y, n) &&
           oz == MinImage(xq.Z - xp.Z, n);
      hp = ox == MinImage(xp.X - xq.X, n) && oy == MinImage(xp.Y - xq.Y, n) &&
           oz == MinImage(xp.Z - xq.Z, n);
    }
  }
  const uint64_t mq = __ballot(hq), mp = __ballot(hp);
  if (lane == 0) {
    L.slot[e0] = mq ? __builtin_ctzll(mq) : -1;
    L.slot[e0 + 1] = mp ? __builtin_ctzll(mp) : -1;
    T.p = p;
    T.q = q;
    T.xp = xp;
    T.xq = xq;
    T.ep = occ[p];
    T.eq = occ[q];
    T.win = first ? w : -1;
  }
}

// The dump cadence, evaluated for consecutive steps with one division per change of cadence.
// Canonical MC (CanonicalMcAbstract.cpp:69-81): steps % every == 0 with every =
// 10^(floor(log10(steps + 1)) - 1), at least 1 and at most L, while steps <= 10 L, and L past
// that. Annealing (`plain`, SimulatedAnnealing.cpp:90): steps % L == 0.
struct LogClock {
  uint64_t L = 0, every = 1, rem = 0, change = 0;  // change: the first step of another cadence
  bool plain = false;
  __device__ void Reset(uint64_t s) {
    every = L;
    change = ~0ull;
    if (!plain && s <= 10 * L) {
      uint64_t p = 1;
      for (uint64_t t = s + 1; t >= 100; t /= 10) p *= 10;
      every = p < L ? p : L;
      const uint64_t next = p <= (~0ull) / 100 ? 100 * p - 1 : ~0ull;  // s + 1 reaches 100 p
      change = next < 10 * L + 1 ? next : 10 * L + 1;
    }
    rem = s % every;
  }
  // Whether step s is due; s is the step Reset saw, then each following step once, in order.
  __device__ bool Due(uint64_t s) {
    if (s >= change) Reset(s);
    const bool due = rem == 0;
    if (++rem == every) rem = 0;
    return due;
  }
};

// CanonicalMcSerial (CanonicalMcSerial.cpp:40-51, CanonicalMcAbstract.cpp:43-101) with an
// optional linear temperature ramp t0 -> t1 over maximum_steps, or the adaptive annealing
// schedule. One workgroup per replica. kLdsOcc: the replica's occupancy (N bytes, N a multiple
// of 4) lives in dynamic LDS for the launch and is written back at its end.
// Two trials per round of three barriers: trial A (step s) and, speculatively, trial B (step
// s + 1) on the same occupancy.
//   1. wave 0 finds A's pair among its 64 precomputed draw attempts (ballot) and gathers both
//      environments and partner slots; wave 1 does the same for B;
//   2. every thread adds its cluster's change for both swaps (DPP wave sums); wave 1 checks
//      whether A's two sites lie among B's draw attempts up to its winning one or in B's
//      environments;
//   3. lane 0 of wave 3 accepts or rejects A, then B unless A was accepted and touched B (B is
//      then redone as the next round's A, so the trajectory is the sequential one); meanwhile
//      waves 0 and 1 draw the attempts of steps s + 2 and s + 3 and lane 0 of wave 2 their
//      Metropolis uniforms.
// kLog: the log records and the averaging ring are compiled in (taken when c.log_every is set).
template <bool kLdsOcc, bool kLog>
__global__ __launch_bounds__(kThreads) void CmcKernel(KTables t, CmcDev c, uint64_t ntrials,
                                                      uint64_t trace_base) {
  __shared__ SiteLds L;
  __shared__ TrialLds s_tr[2];
  __shared__ int s_done, s_adv, s_conf;
  __shared__ double s_u[4], s_beta[4];  // by step & 3; s_beta: the ramp's beta
  // annealing: T at the start of the round and the betas of A (that T) and B (after A's decay)
  __shared__ double s_T, s_bA, s_bB;
  extern __shared__ uint32_t occ_words[];
  constexpr int kAcceptThread = 3 * 64;
  const int r = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  LoadSiteTables(t, L);
  const int n = c.ncell[r];
  const uint64_t N = 4ull * n * n * n;
  uint8_t *occ = c.occ + c.off[r];
  if constexpr (kLdsOcc) {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(occ);
    for (uint64_t i = tid; i < N / 4; i += kThreads) occ_words[i] = src[i];
    occ = reinterpret_cast<uint8_t *>(occ_words);
  }
  uint64_t steps = c.steps[r];
  const uint64_t maxs = c.max_steps[r];
  uint64_t acc = 0;
  double energy = 0, abse = 0;
  CmcAnneal A{};
  AnnealSchedule sched{};
  LogClock clk;
  uint64_t log_i = 0, ring_i = 0, ring_at = 0;  // records, energies added, ring slot
  if (tid == kAcceptThread) {
    acc = c.accepted[r];
    energy = c.energy[r];
    abse = c.abs_energy[r];
    if (c.anneal) {
      A = c.anneal[r];
      sched = MakeAnnealSchedule(maxs);
      s_T = A.T;
    }
    if constexpr (kLog) {
      clk.L = c.log_every;
      clk.plain = c.anneal != nullptr;
      clk.Reset(steps);
      log_i = c.log_n[r];
      c.log_lo[r] = log_i;
      if (c.avg_m) {
        ring_i = c.ring_n[r];
        ring_at = ring_i % c.ring_cap;
      }
    }
  }
  const double decay = c.anneal ? MakeAnnealSchedule(maxs).decay : 1.0;
  if (tid == 0) s_done = steps > maxs ? 1 : 0;
  PairDraw cand{};
  int ox = 0, oy = 0, oz = 0;  // waves 0 and 1: the lane's environment offset
  if (wave < 2) {
    cand = DrawAttempt(c, r, steps + wave, static_cast<uint32_t>(lane), N, n);
    if (lane < kSiteSlots) Unpack(t.site_off[lane], ox, oy, oz);
  }
  // the ramp's temperature and beta of a step (CanonicalMcAbstract: beta = 1 / kB / T)
  auto ramp_T = [&](uint64_t step) {
    return c.t0[r] + (c.t1[r] - c.t0[r]) * (static_cast<double>(step) / static_cast<double>(maxs));
  };
  auto ramp_beta = [&](uint64_t step) { return 1.0 / kBoltzmann / ramp_T(step); };
  // a log record of the step about to be tried (accept thread)
  auto put_log = [&](uint64_t step, double T, double second) {
    if (log_i < c.log_cap) {
      CmcLogRecord &rec = c.log[static_cast<uint64_t>(r) * c.log_cap + log_i];
      rec.steps = step;
      rec.temperature = T;
      rec.energy = energy;
      rec.second = second;
      rec.abs_energy = abse;
      rec.count = ring_i;
    }
    ++log_i;  // counts past the capacity too, so the host can report the loss
  };
  // lane 0 of wave 2: the Metropolis uniforms (and ramp betas) of steps s and s + 1
  auto draw_ahead = [&](uint64_t step) {
    s_u[step & 3] = TrialUniform(c, r, step);
    if (!c.anneal) s_beta[step & 3] = ramp_beta(step);
  };
  if (tid == 128) {
    draw_ahead(steps);
    draw_ahead(steps + 1);
  }
  __syncthreads();
  const SiteCluster cl = DecodeSiteCluster(t, L, tid);
  // Metropolis step of the trial at `step` (accept thread): accept, trace, swap, anneal.
  // pre_T, pre_beta: a temperature and its beta computed ahead (used when A.T equals pre_T).
  auto metropolis = [&](uint64_t step, uint64_t trace_index, const TrialLds &T, double de,
                        double pre_T, double pre_beta) {
    if constexpr (kLog) {
      if (c.avg_m) {  // ThermodynamicAveraging::AddEnergy before Dump (CanonicalMcSerial.cpp:45)
        c.ring[static_cast<uint64_t>(r) * c.ring_cap + ring_at] = energy;
        if (++ring_at == c.ring_cap) ring_at = 0;
        ++ring_i;
      }
    }
    double beta;
    if (c.anneal) {
      const bool lower = energy < A.lowest - 1e-4;  // Dump (SimulatedAnnealing.cpp:80-97)
      if (lower) A.lowest = energy;
      if constexpr (kLog) {
        if (lower) put_log(step, A.T, A.lowest);
        if (clk.Due(step)) put_log(step, A.T, A.lowest);
      }
      beta = A.T == pre_T ? pre_beta : 1.0 / kBoltzmann / fmax(A.T, 1e-12);
    } else {
      beta = s_beta[step & 3];
      if constexpr (kLog) {
        if (clk.Due(step)) put_log(step, ramp_T(step), 0.0);
      }
    }
    const bool accept = de < 0 || s_u[step & 3] < exp(-de * beta);
    if (trace_index < c.trace_cap) {
      const uint64_t o = static_cast<uint64_t>(r) * c.trace_cap + trace_index;
      c.tr_p[o] = T.p;
      c.tr_q[o] = T.q;
      c.tr_de[o] = de;
      c.tr_acc[o] = accept ? 1 : 0;
    }
    if (accept) {
      if (c.atom) {
        const int64_t o = c.off[r];
        AtomJump(c.atom + o, c.shift + 3 * o, T.p, T.q, ImageChange(T.xq.X - T.xp.X, n),
                 ImageChange(T.xq.Y - T.xp.Y, n), ImageChange(T.xq.Z - T.xp.Z, n));
      }
      occ[T.p] = static_cast<uint8_t>(T.eq);
      occ[T.q] = static_cast<uint8_t>(T.ep);
      energy += de;
      abse += de;
      ++acc;
    }
    if (c.anneal) AnnealUpdate(A, sched, accept, step, energy);
    return accept;
  };
  bool fresh = true;  // waves 0 and 1 hold the attempts of steps and steps + 1
  uint64_t it = 0;
  while (it < ntrials) {
    if (s_done) break;
    if (wave < 2) {
      if (!fresh) cand = DrawAttempt(c, r, steps + wave, static_cast<uint32_t>(lane), N, n);
      FindTrial(c, r, steps + wave, occ, N, n, t.S, lane, ox, oy, oz, cand, L, 2 * wave,
                s_tr[wave]);
      if (wave == 0 && lane == 0 && s_tr[0].win == -2) {
        atomicOr(c.error, 2);
        s_done = 1;
      }
    }
    __syncthreads();
    if (s_done) break;
    {
      const TrialLds &TA = s_tr[0], &TB = s_tr[1];
      const bool hasB = TB.win != -2;
      double va = ClusterDelta(L, cl, 0, TA.ep, TA.eq, L.slot[0]) +
                  ClusterDelta(L, cl, 1, TA.eq, TA.ep, L.slot[1]);
      double vb = hasB ? ClusterDelta(L, cl, 2, TB.ep, TB.eq, L.slot[2]) +
                             ClusterDelta(L, cl, 3, TB.eq, TB.ep, L.slot[3])
                       : 0.0;
      va = WaveSum(va);
      vb = WaveSum(vb);
      if (lane == 0) {
        L.part[wave] = va;
        L.part2[wave] = vb;
      }
      if (c.anneal && tid == 128) {  // both betas of the round, off the Metropolis lane's path
        const double T = s_T, Tb = T * decay;
        s_bA = 1.0 / kBoltzmann / fmax(T, 1e-12);
        s_bB = 1.0 / kBoltzmann / fmax(Tb, 1e-12);
      }
      if (wave == 1) {  // does A's swap touch B's draw or environments?
        bool hit = false;
        if (TB.win >= 0 && lane <= TB.win)
          hit = cand.a == TA.p || cand.a == TA.q || cand.b == TA.p || cand.b == TA.q;
        if (hasB && lane < kSiteSlots) {
          const int64_t s1 = SiteAt(TB.xp, n, ox, oy, oz), s2 = SiteAt(TB.xq, n, ox, oy, oz);
          hit = hit || s1 == TA.p || s1 == TA.q || s2 == TA.p || s2 == TA.q;
        }
        const uint64_t mh = __ballot(hit);
        if (lane == 0) s_conf = (!hasB || TB.win < 0 || mh) ? 1 : 0;
      }
    }
    __syncthreads();
    if (tid == kAcceptThread) {
      double de = 0.0;
      for (int w = 0; w < kThreads / 64; ++w) de += L.part[w];
      const double T = c.anneal ? s_T : 0.0;
      const bool acc_a = metropolis(steps, trace_base + it, s_tr[0], de, T, c.anneal ? s_bA : 0.0);
      int adv = 1;
      const bool b_ok = s_tr[1].win != -2 && !(acc_a && s_conf);
      if (steps + 1 <= maxs && it + 1 < ntrials && b_ok) {
        double deb = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) deb += L.part2[w];
        (void)metropolis(steps + 1, trace_base + it + 1, s_tr[1], deb, T * decay,
                         c.anneal ? s_bB : 0.0);
        adv = 2;
      }
      s_adv = adv;
      if (c.anneal) s_T = A.T;
      if (steps + adv > maxs) s_done = 1;
    } else if (wave < 2) {
      cand = DrawAttempt(c, r, steps + 2 + wave, static_cast<uint32_t>(lane), N, n);
    } else if (tid == 128) {
      draw_ahead(steps + 2);
      draw_ahead(steps + 3);
    }
    __syncthreads();
    const int adv = s_adv;
    fresh = adv == 2;
    steps += adv;
    it += adv;
  }
  if constexpr (kLdsOcc) {
    __syncthreads();
    uint32_t *dst = reinterpret_cast<uint32_t *>(c.occ + c.off[r]);
    for (uint64_t i = tid; i < N / 4; i += kThreads) dst[i] = occ_words[i];
  }
  if (tid == kAcceptThread) {
    c.steps[r] = steps;
    c.accepted[r] = acc;
    c.energy[r] = energy;
    c.abs_energy[r] = abse;
    if (c.anneal) c.anneal[r] = A;
    if constexpr (kLog) {
      c.log_n[r] = log_i;
      if (c.avg_m) c.ring_n[r] = ring_i;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// CanonicalMcOmp (lmc/mc/src/CanonicalMcOmp.cpp:40-92): rounds of up to K swap trials. A round
// draws its pairs through one attempt stream (attempt a of the round starting at step s0 is
// Philox counter (s0, replica, 2a + 1), redrawn while the species agree); each event takes the
// first pair with neither site within 3NN of (or on) an earlier event's sites, and the 50th
// draw of an event ends the round. All dE are taken on the round's starting occupancy (the
// exclusion keeps them exact), the Metropolis tests of a round are independent of each other
// (uniform of step s: counter (s, replica, 0xFFFFFFFE)), and the accepted swaps touch disjoint
// sites. A new round starts only while steps <= maximum_steps; a started round runs to its end,
// across launches if need be (the round is kept in c.omp_*).
//
// One 1024-thread workgroup per replica (the canonical workloads have few replicas, so a CU
// each; 16 waves hide the phases' latencies). Per round:
//   draw   : passes of 1024 attempts, one per thread, species checked in parallel; the first 128
//            valid attempts of a pass ranked by ballots; their exclusion by the round's earlier
//            events and their pairwise conflicts checked by 8 threads each (LDS atomics); then
//            one lane walks them in order (the reference's do/while with its draw count);
//   dE     : both 43-site environments of every event gathered at once; thread (chunk, event)
//            adds the changes of 16 site clusters for both swapped sites, and the 16 chunk sums
//            of an event are added in chunk order;
//   trials : event lanes test their Metropolis steps in parallel and apply the accepted swaps;
//            one lane adds the energies in trial order (and takes the log records).
constexpr int kOmpThreads = 1024;
constexpr int kOmpMaxK = 64;
constexpr int kOmpPass = kOmpThreads;          // attempts per draw pass
constexpr int kOmpMaxValid = 128;              // valid attempts resolved per pass
constexpr int kOmpGroups = kOmpThreads / kOmpMaxValid;  // checking threads per valid attempt
constexpr int kOmpChunks = kOmpThreads / kOmpMaxK;      // cluster chunks per event
constexpr int kOmpChunkLen = kMaxSiteClusters / kOmpChunks;
static_assert(kOmpChunks * kOmpChunkLen == kMaxSiteClusters, "chunks cover the site clusters");

struct OmpDraw {  // one draw pass: its valid attempts in attempt order
  int64_t va[kOmpMaxValid], vb[kOmpMaxValid];
  XYZ vxa[kOmpMaxValid], vxb[kOmpMaxValid];
  uint32_t conf[kOmpMaxValid][kOmpMaxValid / 32];  // bit j: conflicts with valid attempt j
  uint32_t excl[kOmpMaxValid];                     // meets an earlier pass's event
  int vidx[kOmpMaxValid];
};

struct OmpLds {
  uint8_t env[2 * kOmpMaxK][48];
  int slot[2 * kOmpMaxK];
  XYZ ch[2 * kOmpMaxK];  // sites of the round's events so far
  int64_t p[kOmpMaxK], q[kOmpMaxK];
  XYZ xp[kOmpMaxK], xq[kOmpMaxK];
  int ep[kOmpMaxK], eq[kOmpMaxK];
  double de[kOmpMaxK];
  int8_t acc[kOmpMaxK];
  union {
    OmpDraw d;                          // draw phase
    double part[kOmpChunks][kOmpMaxK];  // dE phase
  } u;
  int wave_valid[kOmpThreads / 64];
  int8_t slot_lut[125];  // env slot of offset (dx, dy, dz) in [-2, 2]^3, -1 beyond the 3rd shell
  // round control
  int len, pos, nch, ct, finished, error;
  uint32_t att0;
};

// Whether site b lies on site a or in its 1st-3rd neighbour shells (squared minimum-image
// distance at most 6 in (a/2)^2; FCC distances are even).
__device__ __forceinline__ bool Near(const XYZ &a, const XYZ &b, int n) {
  const int dx = MinImage(b.X - a.X, n), dy = MinImage(b.Y - a.Y, n), dz = MinImage(b.Z - a.Z, n);
  return dx * dx + dy * dy + dz * dz <= 6;
}

// w(x -> en) - w(x -> eo) of cluster q in environment e, 0 when it contains slot skip.
__device__ __forceinline__ double ClusterDeltaAt(const SiteLds &L, const SiteCluster &q,
                                                 const uint8_t *e, int eo, int en, int skip) {
  if (!q.live || (skip >= 0 && (q.r1 == skip || q.r2 == skip))) return 0.0;
  const int rest = q.base + e[q.o1] * q.k1 + e[q.o2] * q.k2;
  return L.w[rest + en * q.kx] - L.w[rest + eo * q.kx];
}

// The round starting at step s0: pairs, environments, partner slots and dE into O (all threads).
__device__ void BuildOmpRound(const KTables &t, const CmcDev &c, SiteLds &L, OmpLds &O, int r,
                              uint64_t s0, const uint8_t *occ, uint64_t N, int n, int K) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  OmpDraw &D = O.u.d;
  if (tid == 0) {
    O.len = 0;
    O.nch = 0;
    O.ct = 0;
    O.finished = 0;
    O.att0 = 0;
  }
  __syncthreads();
  while (!O.finished) {
    const PairDraw d = DrawAttempt(c, r, s0, O.att0 + static_cast<uint32_t>(tid), N, n);
    const bool valid = occ[d.a] != occ[d.b];
    const uint64_t m = __ballot(valid);
    if (lane == 0) O.wave_valid[wave] = __builtin_popcountll(m);
    __syncthreads();
    int u = __builtin_popcountll(m & ((1ull << lane) - 1));
    for (int w = 0; w < wave; ++w) u += O.wave_valid[w];
    int total = 0;
    for (int w = 0; w < kOmpThreads / 64; ++w) total += O.wave_valid[w];
    const int nv = total < kOmpMaxValid ? total : kOmpMaxValid;
    if (valid && u < kOmpMaxValid) {
      D.va[u] = d.a;
      D.vb[u] = d.b;
      D.vxa[u] = d.xa;
      D.vxb[u] = d.xb;
      D.vidx[u] = tid;
    }
    if (tid < kOmpMaxValid) {
      D.excl[tid] = 0;
      for (int w = 0; w < kOmpMaxValid / 32; ++w) D.conf[tid][w] = 0;
    }
    __syncthreads();
    {  // thread (g, v): the g-th share of valid attempt v's checks
      const int v = tid % kOmpMaxValid, g = tid / kOmpMaxValid;
      if (v < nv) {
        const XYZ a = D.vxa[v], b = D.vxb[v];
        bool ex = false;
        for (int j = g; j < O.nch; j += kOmpGroups) ex = ex || Near(O.ch[j], a, n) || Near(O.ch[j], b, n);
        if (ex) atomicOr(&D.excl[v], 1u);
        static_assert(kOmpMaxValid == 128, "conflict masks are two 64-bit words");
        uint64_t lo = 0, hi = 0;  // registers, not a run-time indexed array
        for (int j = g; j < v; j += kOmpGroups) {
          const bool hit = Near(D.vxa[j], a, n) || Near(D.vxa[j], b, n) || Near(D.vxb[j], a, n) ||
                           Near(D.vxb[j], b, n);
          if (hit) {
            if (j < 64) lo |= 1ull << j;
            else hi |= 1ull << (j - 64);
          }
        }
        if (lo & 0xFFFFFFFFull) atomicOr(&D.conf[v][0], static_cast<uint32_t>(lo));
        if (lo >> 32) atomicOr(&D.conf[v][1], static_cast<uint32_t>(lo >> 32));
        if (hi & 0xFFFFFFFFull) atomicOr(&D.conf[v][2], static_cast<uint32_t>(hi));
        if (hi >> 32) atomicOr(&D.conf[v][3], static_cast<uint32_t>(hi >> 32));
      }
    }
    __syncthreads();
    if (wave == 0) {
      // The reference's draw loop over the pass's valid attempts, in order: an attempt is taken
      // unless it meets an earlier pass's event or conflicts with an attempt taken before it.
      // Lane l decides attempts l and l + 64 as soon as every earlier attempt it conflicts with
      // is decided (ballots; the chains of conflicts are short), then the draw count and K cut
      // the taken set (a wave-uniform scan), and the taken attempts become the next events.
      const int v0 = lane, v1 = lane + 64;
      const uint64_t c0 = v0 < nv ? D.conf[v0][0] | (static_cast<uint64_t>(D.conf[v0][1]) << 32) : 0;
      const uint64_t c1lo = v1 < nv ? D.conf[v1][0] | (static_cast<uint64_t>(D.conf[v1][1]) << 32) : 0;
      const uint64_t c1hi = v1 < nv ? D.conf[v1][2] | (static_cast<uint64_t>(D.conf[v1][3]) << 32) : 0;
      bool d0 = v0 >= nv || D.excl[v0] != 0, d1 = v1 >= nv || D.excl[v1] != 0;
      bool a0 = false, a1 = false;
      uint64_t dlo = __ballot(d0), dhi = __ballot(d1), alo = 0, ahi = 0;
      while (~dlo | ~dhi) {
        if (!d0 && !(c0 & ~dlo)) {
          d0 = true;
          a0 = !(c0 & alo);
        }
        if (!d1 && !(c1lo & ~dlo) && !(c1hi & ~dhi)) {
          d1 = true;
          a1 = !(c1lo & alo) && !(c1hi & ahi);
        }
        dlo = __ballot(d0);
        dhi = __ballot(d1);
        alo = __ballot(a0);
        ahi = __ballot(a1);
      }
      // draw count and K (wave-uniform): draw v of the current event has count base + v - last
      const int len0 = O.len, nch0 = O.nch;
      int base = O.ct, last = -1, len = len0;
      bool fin = false;
      uint64_t tlo = 0, thi = 0, rlo = alo, rhi = ahi;
      while (true) {
        const int p = rlo ? __builtin_ctzll(rlo) : (rhi ? 64 + __builtin_ctzll(rhi) : nv);
        if (p >= nv) {  // no further taken attempt in this pass
          if (base + (nv - 1 - last) >= 50) fin = true;
          else base += nv - 1 - last;
          break;
        }
        if (base + (p - last) >= 50) {  // the 50th draw of the event comes first
          fin = true;
          break;
        }
        if (p < 64) {
          tlo |= 1ull << p;
          rlo &= rlo - 1;
        } else {
          thi |= 1ull << (p - 64);
          rhi &= rhi - 1;
        }
        ++len;
        base = 0;
        last = p;
        if (len == K) {
          fin = true;
          break;
        }
      }
      const bool t0 = (tlo >> v0) & 1, t1 = (thi >> v0) & 1;
      if (t0 || t1) {
        const int v = t0 ? v0 : v1;
        const int e = len0 + (t0 ? __builtin_popcountll(tlo & ((1ull << v0) - 1))
                                 : __builtin_popcountll(tlo) +
                                       __builtin_popcountll(thi & ((1ull << v0) - 1)));
        O.p[e] = D.va[v];
        O.q[e] = D.vb[v];
        O.xp[e] = D.vxa[v];
        O.xq[e] = D.vxb[v];
        O.ch[nch0 + 2 * (e - len0)] = D.vxa[v];
        O.ch[nch0 + 2 * (e - len0) + 1] = D.vxb[v];
      }
      if (t0 && t1) {  // the lane's second taken attempt
        const int e = len0 + __builtin_popcountll(tlo) + __builtin_popcountll(thi & ((1ull << v0) - 1));
        O.p[e] = D.va[v1];
        O.q[e] = D.vb[v1];
        O.xp[e] = D.vxa[v1];
        O.xq[e] = D.vxb[v1];
        O.ch[nch0 + 2 * (e - len0)] = D.vxa[v1];
        O.ch[nch0 + 2 * (e - len0) + 1] = D.vxb[v1];
      }
      if (lane == 0) {
        O.ct = base;
        O.len = len;
        O.nch = nch0 + 2 * (len - len0);
        // the next pass starts after the last attempt looked at
        const uint32_t adv = total > kOmpMaxValid ? static_cast<uint32_t>(D.vidx[nv - 1] + 1)
                                                  : static_cast<uint32_t>(kOmpPass);
        O.att0 += adv;
        if (!fin && O.att0 >= kMaxPairAttempts) {
          O.error = 1;
          fin = true;
        }
        O.finished = fin ? 1 : 0;
      }
    }
    __syncthreads();
  }
  const int len = O.len;
  // species, both environments and partner slots of every event
  for (int e = tid; e < len; e += kOmpThreads) {
    O.ep[e] = occ[O.p[e]];
    O.eq[e] = occ[O.q[e]];
  }
  for (int i = tid; i < 2 * len * 64; i += kOmpThreads) {
    const int e2 = i >> 6, j = i & 63;
    if (j < kSiteSlots) {
      int dx, dy, dz;
      Unpack(L.off[j], dx, dy, dz);
      const XYZ &x = (e2 & 1) ? O.xq[e2 >> 1] : O.xp[e2 >> 1];
      const uint8_t v = occ[SiteAt(x, n, dx, dy, dz)];
      O.env[e2][j] = v < t.S ? v : 0;  // codes are validated at creation
    }
  }
  for (int e2 = tid; e2 < 2 * len; e2 += kOmpThreads) {  // partner slots
    const XYZ &a = (e2 & 1) ? O.xq[e2 >> 1] : O.xp[e2 >> 1];
    const XYZ &b = (e2 & 1) ? O.xp[e2 >> 1] : O.xq[e2 >> 1];
    const int dx = MinImage(b.X - a.X, n), dy = MinImage(b.Y - a.Y, n), dz = MinImage(b.Z - a.Z, n);
    const bool in = dx >= -2 && dx <= 2 && dy >= -2 && dy <= 2 && dz >= -2 && dz <= 2;
    O.slot[e2] = in ? O.slot_lut[(dx + 2) * 25 + (dy + 2) * 5 + (dz + 2)] : -1;
  }
  __syncthreads();
  {  // dE: thread (chunk, event) adds its chunk of site clusters for both swapped sites
    const int e = tid % kOmpMaxK, ch = tid / kOmpMaxK;
    if (e < len) {
      const int ep = O.ep[e], eq = O.eq[e];
      const uint8_t *ea = O.env[2 * e], *eb = O.env[2 * e + 1];
      const int sa = O.slot[2 * e], sb = O.slot[2 * e + 1];
      double v = 0.0;
#pragma unroll 4
      for (int i = 0; i < kOmpChunkLen; ++i) {
        const SiteCluster q = DecodeSiteCluster(t, L, ch * kOmpChunkLen + i);
        v += ClusterDeltaAt(L, q, ea, ep, eq, sa) + ClusterDeltaAt(L, q, eb, eq, ep, sb);
      }
      O.u.part[ch][e] = v;
    }
  }
  __syncthreads();
  for (int e = tid; e < len; e += kOmpThreads) {
    double sum = 0.0;
    for (int ch = 0; ch < kOmpChunks; ++ch) sum += O.u.part[ch][e];
    O.de[e] = sum;
    const int64_t o = static_cast<int64_t>(r) * c.omp_k + e;
    c.omp_p[o] = O.p[e];
    c.omp_q[o] = O.q[e];
    c.omp_de[o] = sum;
  }
  if (tid == 0) O.pos = 0;
  __syncthreads();
}

template <bool kLdsOcc, bool kLog>
__global__ __launch_bounds__(kOmpThreads) void CmcOmpKernel(KTables t, CmcDev c,
                                                            uint64_t ntrials,
                                                            uint64_t trace_base) {
  __shared__ SiteLds L;
  __shared__ OmpLds O;
  __shared__ uint64_t s_steps;
  extern __shared__ uint32_t occ_words[];
  const int r = blockIdx.x;
  const int tid = threadIdx.x;
  LoadSiteTables(t, L);
  const int n = c.ncell[r];
  const uint64_t N = 4ull * n * n * n;
  const int K = c.omp_k;
  uint8_t *occ = c.occ + c.off[r];
  if constexpr (kLdsOcc) {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(occ);
    for (uint64_t i = tid; i < N / 4; i += kOmpThreads) occ_words[i] = src[i];
    occ = reinterpret_cast<uint8_t *>(occ_words);
  }
  const uint64_t maxs = c.max_steps[r];
  uint64_t acc = 0, log_i = 0, ring_i = 0, ring_at = 0;
  double energy = 0, abse = 0;
  LogClock clk;
  if (tid < 125) O.slot_lut[tid] = -1;
  __syncthreads();
  if (tid >= 1 && tid < kSiteSlots) {  // the first slot of each offset (slot 0: the site itself)
    int x, y, z;
    Unpack(L.off[tid], x, y, z);
    O.slot_lut[(x + 2) * 25 + (y + 2) * 5 + (z + 2)] = static_cast<int8_t>(tid);
  }
  if (tid == 0) {
    s_steps = c.steps[r];
    acc = c.accepted[r];
    energy = c.energy[r];
    abse = c.abs_energy[r];
    O.len = c.omp_len[r];
    O.pos = c.omp_pos[r];
    O.error = 0;
    if constexpr (kLog) {
      clk.L = c.log_every;
      clk.Reset(s_steps);
      log_i = c.log_n[r];
      c.log_lo[r] = log_i;
      if (c.avg_m) {
        ring_i = c.ring_n[r];
        ring_at = ring_i % c.ring_cap;
      }
    }
  }
  __syncthreads();
  // a round left by the previous launch: its pairs and dE, species and coordinates again (its
  // untried sites are untouched by the swaps tried so far)
  if (O.pos < O.len) {
    for (int e = tid; e < O.len; e += kOmpThreads) {
      const int64_t o = static_cast<int64_t>(r) * K + e;
      O.p[e] = c.omp_p[o];
      O.q[e] = c.omp_q[o];
      O.de[e] = c.omp_de[o];
      O.xp[e] = ToXYZ(O.p[e], n);
      O.xq[e] = ToXYZ(O.q[e], n);
      O.ep[e] = occ[O.p[e]];
      O.eq[e] = occ[O.q[e]];
    }
    __syncthreads();
  }
  auto ramp_T = [&](uint64_t step) {
    return c.t0[r] + (c.t1[r] - c.t0[r]) * (static_cast<double>(step) / static_cast<double>(maxs));
  };
  uint64_t it = 0;
  while (it < ntrials) {
    // every thread reads the round state before BuildOmpRound's first lane rewrites it
    const bool need = O.pos == O.len;
    const bool stop = need && s_steps > maxs;
    __syncthreads();
    if (stop) break;
    if (need) {
      BuildOmpRound(t, c, L, O, r, s_steps, occ, N, n, K);
      if (O.error) break;
    }
    const int pos = O.pos;
    const uint64_t left = ntrials - it;
    const int m = static_cast<uint64_t>(O.len - pos) < left ? O.len - pos : static_cast<int>(left);
    const uint64_t s0 = s_steps;
    // Metropolis tests of the round's next m events, in parallel
    for (int e = pos + tid; e < pos + m; e += kOmpThreads) {
      const uint64_t step = s0 + static_cast<uint64_t>(e - pos);
      const double de = O.de[e];
      bool a = de < 0;
      if (!a) a = TrialUniform(c, r, step) < exp(-de * (1.0 / kBoltzmann / ramp_T(step)));
      O.acc[e] = a ? 1 : 0;
      const uint64_t ti = trace_base + it + static_cast<uint64_t>(e - pos);
      if (ti < c.trace_cap) {
        const uint64_t o = static_cast<uint64_t>(r) * c.trace_cap + ti;
        c.tr_p[o] = O.p[e];
        c.tr_q[o] = O.q[e];
        c.tr_de[o] = de;
        c.tr_acc[o] = a ? 1 : 0;
      }
    }
    __syncthreads();
    if (tid == 0) {  // energies in trial order, with the averaging ring and log records
#pragma unroll 4
      for (int e = pos; e < pos + m; ++e) {
        const uint64_t step = s0 + static_cast<uint64_t>(e - pos);
        if constexpr (kLog) {
          if (c.avg_m) {
            c.ring[static_cast<uint64_t>(r) * c.ring_cap + ring_at] = energy;
            if (++ring_at == c.ring_cap) ring_at = 0;
            ++ring_i;
          }
          if (clk.Due(step)) {
            if (log_i < c.log_cap) {
              CmcLogRecord &rec = c.log[static_cast<uint64_t>(r) * c.log_cap + log_i];
              rec.steps = step;
              rec.temperature = ramp_T(step);
              rec.energy = energy;
              rec.second = 0.0;
              rec.abs_energy = abse;
              rec.count = ring_i;
            }
            ++log_i;
          }
        }
        if (O.acc[e]) {
          energy += O.de[e];
          abse += O.de[e];
          ++acc;
        }
      }
      s_steps = s0 + static_cast<uint64_t>(m);
      O.pos = pos + m;
    }
    // the accepted swaps (disjoint sites)
    for (int e = pos + tid; e < pos + m; e += kOmpThreads) {
      if (!O.acc[e]) continue;
      if (c.atom) {
        const int64_t o = c.off[r];
        AtomJump(c.atom + o, c.shift + 3 * o, O.p[e], O.q[e], ImageChange(O.xq[e].X - O.xp[e].X, n),
                 ImageChange(O.xq[e].Y - O.xp[e].Y, n), ImageChange(O.xq[e].Z - O.xp[e].Z, n));
      }
      occ[O.p[e]] = static_cast<uint8_t>(O.eq[e]);
      occ[O.q[e]] = static_cast<uint8_t>(O.ep[e]);
    }
    it += static_cast<uint64_t>(m);
    __syncthreads();
  }
  if constexpr (kLdsOcc) {
    __syncthreads();
    uint32_t *dst = reinterpret_cast<uint32_t *>(c.occ + c.off[r]);
    for (uint64_t i = tid; i < N / 4; i += kOmpThreads) dst[i] = occ_words[i];
  }
  if (tid == 0) {
    if (O.error) atomicOr(c.error, 2);
    c.steps[r] = s_steps;
    c.accepted[r] = acc;
    c.energy[r] = energy;
    c.abs_energy[r] = abse;
    c.omp_len[r] = O.len;
    c.omp_pos[r] = O.pos;
    if constexpr (kLog) {
      c.log_n[r] = log_i;
      if (c.avg_m) c.ring_n[r] = ring_i;
    }
  }
}

// ThermodynamicAveraging::GetThermodynamicAverage (ThermodynamicAveraging.cpp:25-40) of the log
// records of the last launch: over the last min(avg_m, count) energies of the ring, the mean m
// and sum((e - m) exp(-(e - m) beta)) / sum(exp(-(e - m) beta)) + m, beta = 1 / kB / T of the
// record. Grid (R, G): workgroup (x, y) takes records lo + y, lo + y + G, ... of replica x.
__device__ __forceinline__ double BlockSum256(double v, double *red) {
  v = WaveSum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const double s = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return s;
}

__global__ __launch_bounds__(kThreads) void CmcAverageKernel(CmcDev c) {
  static_assert(kThreads == 256, "BlockSum256 reduces four waves");
  __shared__ double red[4];
  const int r = blockIdx.x;
  const uint64_t lo = c.log_lo[r];
  const uint64_t hi = c.log_n[r] < c.log_cap ? c.log_n[r] : c.log_cap;
  const double *ring = c.ring + static_cast<uint64_t>(r) * c.ring_cap;
  for (uint64_t i = lo + blockIdx.y; i < hi; i += gridDim.y) {
    CmcLogRecord &rec = c.log[static_cast<uint64_t>(r) * c.log_cap + i];
    const uint64_t cnt = rec.count;
    const uint64_t w = cnt < c.avg_m ? cnt : c.avg_m;  // >= 1: the record's own energy
    const uint64_t first = (cnt - w) % c.ring_cap;
    double sum = 0.0;
    for (uint64_t j = threadIdx.x; j < w; j += kThreads) {
      const uint64_t x = first + j;
      sum += ring[x < c.ring_cap ? x : x - c.ring_cap];
    }
    const double mean = BlockSum256(sum, red) / static_cast<double>(w);
    const double beta = 1.0 / kBoltzmann / rec.temperature;
    double num = 0.0, den = 0.0;
    for (uint64_t j = threadIdx.x; j < w; j += kThreads) {
      const uint64_t x = first + j;
      const double e = ring[x < c.ring_cap ? x : x - c.ring_cap] - mean;
      const double ex = exp(-e * beta);
      num += e * ex;
      den += ex;
    }
    num = BlockSum256(num, red);
    den = BlockSum256(den, red);
    if (threadIdx.x == 0) rec.second = num / den + mean;
  }
}

// ---------------------------------------------------------------------------------------------
// EnergyPredictor::GetEncode / GetEnergy (EnergyPredictor.cpp:40-96, :173-177): ordered cluster
// counts over the whole lattice, integer histogram per replica, then theta . (count / N).
__constant__ int8_t c_nn1[12][3] = {{1, 1, 0},  {1, -1, 0},  {-1, 1, 0}, {-1, -1, 0},
                                    {1, 0, 1},  {1, 0, -1},  {-1, 0, 1}, {-1, 0, -1},
                                    {0, 1, 1},  {0, 1, -1},  {0, -1, 1}, {0, -1, -1}};
__constant__ int8_t c_nn2[6][3] = {{2, 0, 0}, {-2, 0, 0}, {0, 2, 0}, {0, -2, 0}, {0, 0, 2}, {0, 0, -2}};
__constant__ int8_t c_nn3[24][3] = {
    {2, 1, 1},  {2, 1, -1},  {2, -1, 1},  {2, -1, -1},  {-2, 1, 1},  {-2, 1, -1},
    {-2, -1, 1}, {-2, -1, -1}, {1, 2, 1},  {1, 2, -1},   {-1, 2, 1},  {-1, 2, -1},
    {1, -2, 1},  {1, -2, -1},  {-1, -2, 1}, {-1, -2, -1}, {1, 1, 2},  {1, -1, 2},
    {-1, 1, 2},  {-1, -1, 2},  {1, 1, -2},  {1, -1, -2},  {-1, 1, -2}, {-1, -1, -2}};

__device__ __forceinline__ int Code(uint8_t v, int S) { return v < S ? v : 0; }

__device__ __forceinline__ int Shell(int x, int y, int z) {
  const int r2 = x * x + y * y + z * z;
  return r2 == 2 ? 1 : r2 == 4 ? 2 : r2 == 6 ? 3 : 0;
}

__global__ __launch_bounds__(kThreads) void TotalCountsKernel(KTables t, const uint8_t *occ,
                                                              const int64_t *off,
                                                              const int32_t *ncell,
                                                              unsigned long long *counts) {
  __shared__ unsigned int hist[512];
  const int r = blockIdx.x;  // grid (R, 64): x the replica, y a slice of its sites
  const int n = ncell[r];
  const int64_t N = 4ll * n * n * n;
  const uint8_t *o = occ + off[r];
  for (int i = threadIdx.x; i < t.n_eci; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  const int S = t.S;
  for (int64_t s = static_cast<int64_t>(blockIdx.y) * blockDim.x + threadIdx.x; s < N;
       s += static_cast<int64_t>(gridDim.y) * blockDim.x) {
    const XYZ c = ToXYZ(s, n);
    const int e1 = Code(o[s], S);
    atomicAdd(&hist[t.eci1[e1]], 1u);
    for (int a = 0; a < 12; ++a) {
      const int ax = c_nn1[a][0], ay = c_nn1[a][1], az = c_nn1[a][2];
      const int e2 = Code(o[SiteAt(c, n, ax, ay, az)], S);
      atomicAdd(&hist[t.eci2[(0 * S + e1) * S + e2]], 1u);
      for (int b = 0; b < 12; ++b) {
        const int x = ax + c_nn1[b][0], y = ay + c_nn1[b][1], z = az + c_nn1[b][2];
        const int sh = Shell(x, y, z);
        if (!sh) continue;
        const int e3 = Code(o[SiteAt(c, n, x, y, z)], S);
        atomicAdd(&hist[t.eci3[(((3 + sh) - 4) * S * S * S) + (e1 * S + e2) * S + e3]], 1u);
      }
      for (int b = 0; b < 6; ++b) {
        const int x = ax + c_nn2[b][0], y = ay + c_nn2[b][1], z = az + c_nn2[b][2];
        if (Shell(x, y, z) != 3) continue;
        const int e3 = Code(o[SiteAt(c, n, x, y, z)], S);
        atomicAdd(&hist[t.eci3[3 * S * S * S + (e1 * S + e2) * S + e3]], 1u);
      }
    }
    for (int a = 0; a < 6; ++a) {
      const int e2 = o[SiteAt(c, n, c_nn2[a][0], c_nn2[a][1], c_nn2[a][2])];
      atomicAdd(&hist[t.eci2[(1 * S + e1) * S + e2]], 1u);
    }
    for (int a = 0; a < 24; ++a) {
      const int e2 = o[SiteAt(c, n, c_nn3[a][0], c_nn3[a][1], c_nn3[a][2])];
      atomicAdd(&hist[t.eci2[(2 * S + e1) * S + e2]], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < t.n_eci; i += blockDim.x)
    if (hist[i]) atomicAdd(&counts[static_cast<int64_t>(r) * t.n_eci + i], hist[i]);
}

// Membership of site s in the sorted list set[m].
__device__ __forceinline__ bool InSet(const int64_t *set, int m, int64_t s) {
  int lo = 0, hi = m;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (set[mid] < s) lo = mid + 1;
    else hi = mid;
  }
  return lo < m && set[lo] == s;
}

// EnergyPredictor::GetEncodeOfCluster (EnergyPredictor.cpp:97-172): the ordered cluster counts
// of the sites in set[m] (sorted, unique) whose member sites all lie in the set, enumerated as in
// GetEncode: from every first site, its neighbours, and for 1NN neighbours the triplets closed
// through the second site's 1NN and 2NN shells.
__global__ __launch_bounds__(kThreads) void ClusterCountsKernel(KTables t, const uint8_t *occ, int n,
                                                                const int64_t *set, int m,
                                                                unsigned long long *counts) {
  __shared__ unsigned int hist[512];
  for (int i = threadIdx.x; i < t.n_eci; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  const int S = t.S;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    const int64_t s1 = set[i];
    const XYZ c = ToXYZ(s1, n);
    const int e1 = Code(occ[s1], S);
    atomicAdd(&hist[t.eci1[e1]], 1u);
    for (int a = 0; a < 12; ++a) {
      const int ax = c_nn1[a][0], ay = c_nn1[a][1], az = c_nn1[a][2];
      const int64_t s2 = SiteAt(c, n, ax, ay, az);
      if (!InSet(set, m, s2)) continue;
      const int e2 = Code(occ[s2], S);
      atomicAdd(&hist[t.eci2[(0 * S + e1) * S + e2]], 1u);
      for (int b = 0; b < 12; ++b) {
        const int x = ax + c_nn1[b][0], y = ay + c_nn1[b][1], z = az + c_nn1[b][2];
        const int sh = Shell(x, y, z);
        if (!sh) continue;
        const int64_t s3 = SiteAt(c, n, x, y, z);
        if (!InSet(set, m, s3)) continue;
        const int e3 = Code(occ[s3], S);
        atomicAdd(&hist[t.eci3[(sh - 1) * S * S * S + (e1 * S + e2) * S + e3]], 1u);
      }
      for (int b = 0; b < 6; ++b) {
        const int x = ax + c_nn2[b][0], y = ay + c_nn2[b][1], z = az + c_nn2[b][2];
        if (Shell(x, y, z) != 3) continue;
        const int64_t s3 = SiteAt(c, n, x, y, z);
        if (!InSet(set, m, s3)) continue;
        const int e3 = Code(occ[s3], S);
        atomicAdd(&hist[t.eci3[3 * S * S * S + (e1 * S + e2) * S + e3]], 1u);
      }
    }
    for (int a = 0; a < 6; ++a) {
      const int64_t s2 = SiteAt(c, n, c_nn2[a][0], c_nn2[a][1], c_nn2[a][2]);
      if (InSet(set, m, s2)) atomicAdd(&hist[t.eci2[(1 * S + e1) * S + Code(occ[s2], S)]], 1u);
    }
    for (int a = 0; a < 24; ++a) {
      const int64_t s2 = SiteAt(c, n, c_nn3[a][0], c_nn3[a][1], c_nn3[a][2]);
      if (InSet(set, m, s2)) atomicAdd(&hist[t.eci2[(2 * S + e1) * S + Code(occ[s2], S)]], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < t.n_eci; i += blockDim.x)
    if (hist[i]) atomicAdd(&counts[i], hist[i]);
}

// The encode count / N_label of each replica (EnergyPredictor.cpp:89-95) and theta . encode.
__global__ void EncodeKernel(KTables t, const unsigned long long *counts, int32_t R,
                             double *encode, double *energy) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  double e = 0.0;
  for (int i = 0; i < t.n_eci; ++i) {
    const double x = static_cast<double>(counts[static_cast<int64_t>(r) * t.n_eci + i]) / t.norm_tot[i];
    if (encode) encode[static_cast<int64_t>(r) * t.n_eci + i] = x;
    e += t.theta[i] * x;
  }
  if (energy) energy[r] = e;
}

__global__ void TotalEnergyKernel(KTables t, const unsigned long long *counts, int32_t R,
                                  double *energy) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  double e = 0.0;
  for (int i = 0; i < t.n_eci; ++i)
    e += t.theta[i] * (static_cast<double>(counts[static_cast<int64_t>(r) * t.n_eci + i]) /
                       t.norm_tot[i]);
  energy[r] = e;
}

__global__ void ZeroKernel(unsigned long long *p, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

hipError_t LaunchSwapDe(const KTables &t, const uint8_t *occ, const int64_t *off,
                        const int32_t *ncell, const int64_t *pairs, int32_t R, double *de,
                        hipStream_t s) {
  if (R <= 0) return hipSuccess;
  hipLaunchKernelGGL(SwapDeKernel, dim3(R), dim3(kThreads), 0, s, t, occ, off, ncell, pairs, de);
  return hipGetLastError();
}

hipError_t LaunchSiteDe(const KTables &t, const uint8_t *occ, const int64_t *off,
                        const int32_t *ncell, const int64_t *sites, const uint8_t *from,
                        const uint8_t *codes, int32_t R, double *de, hipStream_t s) {
  if (R <= 0) return hipSuccess;
  hipLaunchKernelGGL(SiteDeKernel, dim3(R), dim3(kThreads), 0, s, t, occ, off, ncell, sites, from,
                     codes, de);
  return hipGetLastError();
}

// The occupancy goes to LDS when the largest replica fits beside the static LDS of the kernel.
hipError_t LaunchCmc(const KTables &t, const CmcDev &c, uint64_t trials, uint64_t trace_base,
                     hipStream_t s) {
  if (c.R <= 0 || trials == 0) return hipSuccess;
  // static LDS of the occupancy-in-LDS kernels, the same on every device of this build
  static std::mutex mu;
  static size_t stat = 0, stat_omp = 0;  // 0 until a query succeeds
  size_t st_serial, st_omp;
  {
    std::lock_guard<std::mutex> lock(mu);
    if (!stat) {
      hipFuncAttributes fa, fb, fc, fd;
      if (hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(CmcKernel<true, false>)) !=
              hipSuccess ||
          hipFuncGetAttributes(&fb, reinterpret_cast<const void *>(CmcKernel<true, true>)) !=
              hipSuccess ||
          hipFuncGetAttributes(&fc, reinterpret_cast<const void *>(CmcOmpKernel<true, false>)) !=
              hipSuccess ||
          hipFuncGetAttributes(&fd, reinterpret_cast<const void *>(CmcOmpKernel<true, true>)) !=
              hipSuccess)
        return hipErrorInvalidDevice;
      stat_omp = std::max(fc.sharedSizeBytes, fd.sharedSizeBytes);
      stat = std::max(fa.sharedSizeBytes, fb.sharedSizeBytes);
    }
    st_serial = stat;
    st_omp = stat_omp;
  }
  const size_t lds_block = CurrentDeviceLimits().lds_block;
  const size_t dyn = static_cast<size_t>(c.max_sites);
  auto launch = [&](auto kernel_lds, auto kernel_hbm, size_t st, int threads) {
    if (c.lds_occupancy && dyn % 4 == 0 && st + dyn <= lds_block) {
      if (dyn > 65536 - st) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel_lds),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 static_cast<int>(dyn));
        if (e != hipSuccess) return e;
      }
      hipLaunchKernelGGL(kernel_lds, dim3(c.R), dim3(threads), dyn, s, t, c, trials, trace_base);
    } else {
      hipLaunchKernelGGL(kernel_hbm, dim3(c.R), dim3(threads), 0, s, t, c, trials, trace_base);
    }
    return hipGetLastError();
  };
  if (c.omp_k > 0)
    return c.log_every
               ? launch(CmcOmpKernel<true, true>, CmcOmpKernel<false, true>, st_omp, kOmpThreads)
               : launch(CmcOmpKernel<true, false>, CmcOmpKernel<false, false>, st_omp, kOmpThreads);
  return c.log_every ? launch(CmcKernel<true, true>, CmcKernel<false, true>, st_serial, kThreads)
                     : launch(CmcKernel<true, false>, CmcKernel<false, false>, st_serial, kThreads);
}

hipError_t LaunchCmcAverage(const CmcDev &c, hipStream_t s) {
  if (c.R <= 0 || !c.log_every || !c.avg_m) return hipSuccess;
  hipLaunchKernelGGL(CmcAverageKernel, dim3(c.R, 64), dim3(kThreads), 0, s, c);
  return hipGetLastError();
}

hipError_t LaunchTotalEncode(const KTables &t, const uint8_t *occ, const int64_t *off,
                             const int32_t *ncell, int32_t R, unsigned long long *counts,
                             double *encode, double *energy, hipStream_t s) {
  if (R <= 0) return hipSuccess;
  const int64_t nc = static_cast<int64_t>(R) * t.n_eci;
  hipLaunchKernelGGL(ZeroKernel, dim3((nc + 255) / 256), dim3(256), 0, s, counts, nc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(TotalCountsKernel, dim3(R, 64), dim3(kThreads), 0, s, t, occ, off, ncell,
                     counts);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(EncodeKernel, dim3((R + 63) / 64), dim3(64), 0, s, t, counts, R, encode, energy);
  return hipGetLastError();
}

hipError_t LaunchClusterEncode(const KTables &t, const uint8_t *occ, int32_t n, const int64_t *set,
                               int32_t m, unsigned long long *counts, double *encode,
                               double *energy, hipStream_t s) {
  hipLaunchKernelGGL(ZeroKernel, dim3((t.n_eci + 255) / 256), dim3(256), 0, s, counts,
                     static_cast<int64_t>(t.n_eci));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int blocks = std::max(1, std::min(64, (m + kThreads - 1) / kThreads));
  hipLaunchKernelGGL(ClusterCountsKernel, dim3(blocks), dim3(kThreads), 0, s, t, occ, n, set, m,
                     counts);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(EncodeKernel, dim3(1), dim3(64), 0, s, t, counts, 1, encode, energy);
  return hipGetLastError();
}

// The cluster-count scratch of lmc_total_energy_noalloc: one copy per device, held until the
// process ends. Host code never passes `scratch` itself (its host-side name is not a device
// address); it passes the address TotalEnergyScratch looked up on the current device.
constexpr size_t kScratchCounts = size_t{1} << 19;  // 4 MiB of 64-bit counts
static __device__ __align__(256) unsigned long long scratch[kScratchCounts];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned long long *g_scratch_ptr[kMaxScratchDevices] = {};

size_t TotalEnergyScratchCapacity() { return kScratchCounts; }

hipError_t TotalEnergyScratch(unsigned long long **ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {  // once per device, with that device current
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<unsigned long long *>(p);
  }
  *ptr = g_scratch_ptr[dev];
  return hipSuccess;
}

hipError_t LaunchTotalEnergy(const KTables &t, const uint8_t *occ, const int64_t *off,
                             const int32_t *ncell, int32_t R, unsigned long long *counts,
                             double *energy, hipStream_t s) {
  if (R <= 0) return hipSuccess;
  const int64_t nc = static_cast<int64_t>(R) * t.n_eci;
  hipLaunchKernelGGL(ZeroKernel, dim3((nc + 255) / 256), dim3(256), 0, s, counts, nc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(TotalCountsKernel, dim3(R, 64), dim3(kThreads), 0, s, t, occ, off, ncell,
                     counts);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(TotalEnergyKernel, dim3((R + 63) / 64), dim3(64), 0, s, t, counts, R, energy);
  return hipGetLastError();
}

}  // namespace lmc
