// gfx950 kernels of the event-evaluation path.
//
// Work mapping: one 256-thread workgroup per replica. Per KMC step
//   1. gather   : the 12 jump environments (12 x 60 sites) -> LDS (bytes, from HBM/L2);
//   2. clusters : 21 lanes per jump sum the per-cluster weights of
//                   dE    (one-endpoint state clusters, VacancyMigrationPredictorQuartic.cpp:112-166),
//                   logit (MMM orbit clusters, VacancyMigrationPredictorE0.cpp:128-151),
//                   logit (MM2 forward + backward, VacancyMigrationPredictorQuartic.cpp:167-215);
//   3. events   : 12 lanes reduce the partials in a fixed order and form Ea, rate = exp(-Ea beta);
//   4. select   : one lane sums the 12 rates in neighbour-id order (FP64), draws Philox u1/u2,
//                 advances the residence-time clock and applies the jump
//                 (KineticMcFirstOmp.cpp:52-82, KineticMcAbstract.cpp:106-182).
// Clusters containing both jump endpoints have the same species multiset before and after the
// jump and are skipped; every other cluster adds w(end) - w(start).
#include <hip/hip_runtime.h>

#include "lmc_device.h"
#include "lmc_rng.h"

namespace lmc {
namespace {

constexpr int kThreads = 256;
constexpr int kLanes = 21;  // lanes per jump event (12 * 21 = 252 <= 256)
constexpr double kBoltzmann = 8.617333262145e-5;  // Constants.hpp:31
constexpr double kPrefactor = 1e13;               // Constants.hpp:32
constexpr double kHalfLattice = 4.046 * 0.5;      // a/2 in Angstrom (Constants.hpp:6)

struct XYZ {
  int X, Y, Z;
};

__device__ __forceinline__ XYZ ToXYZ(int64_t id, int n) {
  const int64_t plane = 2ll * n * n;
  const int X = static_cast<int>(id / plane);
  const int64_t r = id - static_cast<int64_t>(X) * plane;
  const int Y = static_cast<int>(r / n);
  const int Zh = static_cast<int>(r - static_cast<int64_t>(Y) * n);
  return {X, Y, 2 * Zh + ((X + Y) & 1)};
}
__device__ __forceinline__ int Wrap(int v, int m) {  // |v| < 2m
  v += v < 0 ? m : 0;
  v -= v >= m ? m : 0;
  return v;
}
__device__ __forceinline__ int64_t SiteAt(const XYZ &c, int n, int dx, int dy, int dz) {
  const int m = 2 * n;
  const int X = Wrap(c.X + dx, m), Y = Wrap(c.Y + dy, m), Z = Wrap(c.Z + dz, m);
  return static_cast<int64_t>(X) * 2 * n * n + static_cast<int64_t>(Y) * n + (Z >> 1);
}
__device__ __forceinline__ void Unpack(int32_t v, int &x, int &y, int &z) {
  x = static_cast<int8_t>(v & 0xff);
  y = static_cast<int8_t>((v >> 8) & 0xff);
  z = static_cast<int8_t>((v >> 16) & 0xff);
}

struct JumpLds {
  uint32_t state[kMaxStateClusters];
  uint32_t mmm[kMaxOrbitClusters];
  uint32_t mm2[kMaxOrbitClusters];
  double w[kMaxW];
  int32_t off[12 * kEnvSlots];
  uint8_t mirror[64];
  int8_t tt[3 * kMaxSpecies * kMaxSpecies];
  uint8_t env[12][64];
  double part[12][kLanes][3];
  double ea[12], de[12], rate[12];
  int64_t nb[12];
  int rank[12];
};

__device__ void LoadJumpTables(const KTables &t, JumpLds &L) {
  for (int i = threadIdx.x; i < t.n_state; i += blockDim.x) L.state[i] = t.state[i];
  for (int i = threadIdx.x; i < t.n_mmm; i += blockDim.x) L.mmm[i] = t.mmm[i];
  if (t.kind == kQuartic)
    for (int i = threadIdx.x; i < t.n_mm2; i += blockDim.x) L.mm2[i] = t.mm2[i];
  for (int i = threadIdx.x; i < t.nw; i += blockDim.x) L.w[i] = t.w[i];
  for (int i = threadIdx.x; i < 12 * kEnvSlots; i += blockDim.x) L.off[i] = t.off[i];
  for (int i = threadIdx.x; i < kEnvSlots; i += blockDim.x) L.mirror[i] = t.mirror[i];
  for (int i = threadIdx.x; i < 3 * t.S * t.S; i += blockDim.x) L.tt[i] = t.tt[i];
}

// Phase 1: species of the 12 environments; neighbour ids of the 12 jumps.
__device__ __forceinline__ void GatherEnv(const KTables &t, JumpLds &L, const uint8_t *occ, int n,
                                          int64_t vac, int32_t *err) {
  const XYZ c = ToXYZ(vac, n);
  for (int i = threadIdx.x; i < 12 * kEnvSlots; i += blockDim.x) {
    const int k = i / kEnvSlots, j = i - k * kEnvSlots;
    int dx, dy, dz;
    Unpack(L.off[i], dx, dy, dz);
    const int64_t s = SiteAt(c, n, dx, dy, dz);
    uint8_t v = occ[s];
    if (v >= t.S) {  // corrupt occupancy: flag it and keep every table index in range
      atomicOr(err, 4);
      v = 0;
    }
    L.env[k][j] = v;
    if (j == t.n_slot) L.nb[k] = s;
  }
}

// A replica's lattice size and vacancy site are usable (uniform over the block).
__device__ __forceinline__ bool ValidSite(int n, int64_t site) {
  return n >= 4 && site >= 0 && site < 4ll * n * n * n;
}

__device__ __forceinline__ double OrbitWeight(const JumpLds &L, const double *bm, const uint8_t *e,
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

// Phase 2: per-lane partial sums of the cluster weights.
__device__ __forceinline__ void AccumulateClusters(const KTables &t, JumpLds &L) {
  const int tid = threadIdx.x;
  if (tid >= 12 * kLanes) return;
  const int k = tid / kLanes, lane = tid - k * kLanes;
  const uint8_t *e = L.env[k];
  const int S = t.S, X = t.x_code, M = e[t.n_slot];
  const int w2 = S, w3 = S + 3 * S * S;
  double de = 0.0;
  for (int c = lane; c < t.n_state; c += kLanes) {
    const uint32_t d = L.state[c];
    const int label = d & 7;
    const int o1 = (d >> 4) & 63, o2 = (d >> 10) & 63;
    int im, ix;
    if (label == 0) {
      im = M;
      ix = X;
    } else if (label < 4) {
      const int b = w2 + (label - 1) * S * S + e[o1];
      im = b + M * S;
      ix = b + X * S;
    } else {
      const int b = w3 + (label - 4) * S * S * S + e[o1] * S + e[o2];
      im = b + M * S * S;
      ix = b + X * S * S;
    }
    const double v = L.w[im] - L.w[ix];
    de += (d & 8) ? -v : v;
  }
  double la = 0.0, lb = 0.0;
  if (t.kind == kE0 || t.kind == kQuartic) {
    const double *bm = t.b1 + static_cast<size_t>(M) * t.F1;
    for (int c = lane; c < t.n_mmm; c += kLanes) la += OrbitWeight(L, bm, e, L.mmm[c], S, false);
  }
  if (t.kind == kQuartic) {
    const double *bm = t.b2 + static_cast<size_t>(M) * t.F2;
    for (int c = lane; c < t.n_mm2; c += kLanes) {
      const uint32_t d = L.mm2[c];
      lb += OrbitWeight(L, bm, e, d, S, false);
      lb += OrbitWeight(L, bm, e, d, S, true);
    }
  }
  L.part[k][lane][0] = de;
  L.part[k][lane][1] = la;
  L.part[k][lane][2] = lb;
}

// Phase 3: Ea, dE and rate per jump; rank of the jump among the 12 by neighbour id.
__device__ __forceinline__ void FinishEvents(const KTables &t, JumpLds &L, double beta,
                                             int32_t *err) {
  const int k = threadIdx.x;
  if (k >= 12) return;
  double de = 0.0, la = 0.0, lb = 0.0;
  for (int l = 0; l < kLanes; ++l) {
    de += L.part[k][l][0];
    la += L.part[k][l][1];
    lb += L.part[k][l][2];
  }
  const int M = L.env[k][t.n_slot];
  const double *p = t.par + 6 * M;
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
  if (!(ea == ea) || !(de == de)) atomicOr(err, 1);  // unknown cluster type / missing element
  L.ea[k] = ea;
  L.de[k] = de;
  L.rate[k] = exp(-ea * beta);  // JumpEvent forward rate (JumpEvent.cpp:6-13)
  int r = 0;
  for (int j = 0; j < 12; ++j) r += L.nb[j] < L.nb[k];
  L.rank[k] = r;
}

__global__ __launch_bounds__(kThreads) void EvalJumpsKernel(KTables t, const uint8_t *occ,
                                                            const int64_t *off,
                                                            const int32_t *ncell,
                                                            const int64_t *vac, int64_t *nb,
                                                            double *ea, double *de, int32_t *err) {
  __shared__ JumpLds L;
  const int r = blockIdx.x;
  if (!ValidSite(ncell[r], vac[r])) {
    if (threadIdx.x == 0) atomicOr(err, 8);
    return;
  }
  LoadJumpTables(t, L);
  __syncthreads();
  GatherEnv(t, L, occ + off[r], ncell[r], vac[r], err);
  __syncthreads();
  AccumulateClusters(t, L);
  __syncthreads();
  FinishEvents(t, L, 0.0, err);
  __syncthreads();
  const int k = threadIdx.x;
  if (k < 12) {
    const int64_t o = static_cast<int64_t>(r) * 12 + L.rank[k];
    nb[o] = L.nb[k];
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

__global__ __launch_bounds__(kThreads) void KmcKernel(KTables t, KmcDev k, uint64_t nsteps,
                                                      uint64_t trace_base) {
  __shared__ JumpLds L;
  __shared__ double s_beta;
  __shared__ int64_t s_vac;
  __shared__ int s_done;
  const int r = blockIdx.x;
  const int tid = threadIdx.x;
  LoadJumpTables(t, L);
  uint8_t *occ = k.occ + k.off[r];
  const int n = k.ncell[r];
  // Replica state lives in lane 0's registers for the whole launch.
  uint64_t steps = 0, maxs = 0;
  double time = 0, energy = 0, abse = 0, T = 0, com[3] = {0, 0, 0};
  int64_t vac = 0, disp[3] = {0, 0, 0};
  int flags = 0, esc = 0, solvent = 0;
  if (tid == 0) {
    steps = k.steps[r];
    maxs = k.max_steps[r];
    time = k.time[r];
    energy = k.energy[r];
    abse = k.abs_energy[r];
    T = k.temp[r];
    vac = k.vac[r];
    flags = k.flags[r];
    esc = k.escaped[r];
    solvent = k.solvent[r];
    for (int d = 0; d < 3; ++d) {
      disp[d] = k.disp[3 * r + d];
      com[d] = k.com[3 * r + d];
    }
    s_vac = vac;
    s_done = (steps > maxs || esc) ? 1 : 0;
  }
  __syncthreads();
  for (uint64_t it = 0; it < nsteps; ++it) {
    if (s_done) break;
    if (tid == 0) {
      if (k.n_tt > 0) T = ScheduleTemperature(k.tt_time, k.tt_temp, k.n_tt, time);
      s_beta = 1.0 / kBoltzmann / T;
    }
    GatherEnv(t, L, occ, n, s_vac, k.error);
    __syncthreads();
    AccumulateClusters(t, L);
    __syncthreads();
    FinishEvents(t, L, s_beta, k.error);
    __syncthreads();
    if (tid == 0) {
      double srate[12], sde[12], sea[12];
      int64_t snb[12];
      int sdir[12];
      for (int j = 0; j < 12; ++j) {
        const int q = L.rank[j];
        srate[q] = L.rate[j];
        sde[q] = L.de[j];
        sea[q] = L.ea[j];
        snb[q] = L.nb[j];
        sdir[q] = j;
      }
      double total = 0.0;
      for (int j = 0; j < 12; ++j) total += srate[j];
      const U4 u = Philox4x32_10(k.seed[r], steps, k.rid[r], 0u);
      const double u1 = UnitOpenLeft(Join64(u.x, u.y));
      const double u2 = UnitOpenRight(Join64(u.z, u.w));
      double dt = -log(u1) / total / kPrefactor;
      if (flags & 1)  // RateCorrector::GetTimeCorrectionFactor (RateCorrector.hpp:17-24)
        dt *= k.cv[r] / (1.64 * exp(-(0.66 / kBoltzmann / T - 0.7))) / (1 - 13 * k.cs[r]);
      int sel = 11;
      double cum = 0.0;
      for (int j = 0; j < 12; ++j) {
        cum += srate[j] / total;
        if (!(cum < u2)) {
          sel = j;
          break;
        }
      }
      if (flags & 2) {  // IsEscaped (KineticMcAbstract.cpp:192-223)
        bool all = true;
        for (int j = 0; j < kEnvSlots; ++j)
          if (((t.vac_shell_mask >> j) & 1ull) && L.env[0][j] != solvent) all = false;
        if (all) {
          esc = 1;
          steps = maxs + 1;
        }
      }
      const int dir = sdir[sel];
      const int64_t jump = snb[sel];
      const uint64_t ti = trace_base + it;
      if (ti < k.trace_cap) {
        const uint64_t o = static_cast<uint64_t>(r) * k.trace_cap + ti;
        k.tr_sel[o] = jump;
        k.tr_ea[o] = sea[sel];
        k.tr_de[o] = sde[sel];
        k.tr_time[o] = time + dt;
      }
      time += dt;
      energy += sde[sel];
      abse += sde[sel];
      int dx, dy, dz;
      Unpack(L.off[dir * kEnvSlots + t.n_slot], dx, dy, dz);
      const int mig = L.env[dir][t.n_slot];
      if ((flags & 4) && mig != solvent) {  // solute centre of mass (KineticMcAbstract.cpp:163-177)
        const double wgt = k.mass[mig] / k.solute_mass[r];
        com[0] -= dx * kHalfLattice * wgt;
        com[1] -= dy * kHalfLattice * wgt;
        com[2] -= dz * kHalfLattice * wgt;
      }
      disp[0] += dx;
      disp[1] += dy;
      disp[2] += dz;
      occ[vac] = static_cast<uint8_t>(mig);  // Config::LatticeJump (Config.cpp:431-456)
      occ[jump] = static_cast<uint8_t>(t.x_code);
      ++steps;
      vac = jump;
      s_vac = vac;
      if (steps > maxs || esc) s_done = 1;
    }
    __syncthreads();
  }
  if (tid == 0) {
    k.steps[r] = steps;
    k.time[r] = time;
    k.energy[r] = energy;
    k.abs_energy[r] = abse;
    k.temp[r] = T;
    k.vac[r] = vac;
    k.escaped[r] = esc;
    for (int d = 0; d < 3; ++d) {
      k.disp[3 * r + d] = disp[d];
      k.com[3 * r + d] = com[d];
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Swap / site energy change (EnergyChangePredictorPairSite.cpp:69-192).
struct SiteLds {
  uint32_t cl[kMaxSiteClusters];
  int32_t off[kSiteSlots];
  double w[kMaxW];
  uint8_t env[2][48];
  double red[kThreads];
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
// contain env slot `skip` (the partner of a coupled swap). Result in lane 0.
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

__device__ double BlockSum(SiteLds &L, double v) {  // fixed-order tree reduction
  L.red[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) L.red[threadIdx.x] += L.red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = L.red[0];
  __syncthreads();
  return r;
}

__device__ void GatherSite(const SiteLds &L, uint8_t *dst, const uint8_t *occ, int n, int64_t x,
                           int S) {
  const XYZ c = ToXYZ(x, n);
  for (int i = threadIdx.x; i < kSiteSlots; i += blockDim.x) {
    int dx, dy, dz;
    Unpack(L.off[i], dx, dy, dz);
    const uint8_t v = occ[SiteAt(c, n, dx, dy, dz)];
    dst[i] = v < S ? v : 0;  // codes are validated by the callers; keep LDS indices in range
  }
}

// Slot of q in the 43-site environment of p, or -1 when q is beyond the third shell.
__device__ int PartnerSlot(const SiteLds &L, int n, int64_t p, int64_t q) {
  const XYZ a = ToXYZ(p, n), b = ToXYZ(q, n);
  const int m = 2 * n;
  auto mi = [m, n](int v) {
    v = ((v % m) + m) % m;
    return v >= n ? v - m : v;
  };
  const int dx = mi(b.X - a.X), dy = mi(b.Y - a.Y), dz = mi(b.Z - a.Z);
  for (int i = 1; i < kSiteSlots; ++i) {
    int x, y, z;
    Unpack(L.off[i], x, y, z);
    if (x == dx && y == dy && z == dz) return i;
  }
  return -1;
}

// dE of swapping the species at p and q. Far pairs: dE(p) + dE(q) on the unswapped config
// (...WithoutCoupling :136-146). Near pairs: clusters containing both keep their multiset and
// are skipped (...WithCoupling :96-134). All threads call; result valid in every thread.
__device__ double SwapDeBlock(const KTables &t, SiteLds &L, const uint8_t *occ, int n, int64_t p,
                              int64_t q) {
  __shared__ int s_qs, s_ps;
  const int ep = occ[p], eq = occ[q];
  if (ep == eq) return 0.0;
  GatherSite(L, L.env[0], occ, n, p, t.S);
  GatherSite(L, L.env[1], occ, n, q, t.S);
  if (threadIdx.x == 0) {
    s_qs = PartnerSlot(L, n, p, q);
    s_ps = PartnerSlot(L, n, q, p);
  }
  __syncthreads();
  const int qs = s_qs, ps = s_ps;
  double v = SitePartial(t, L, 0, ep, eq, qs) + SitePartial(t, L, 1, eq, ep, ps);
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
  LoadSiteTables(t, L);
  __syncthreads();
  const double v = SwapDeBlock(t, L, o, ncell[r], p, q);
  if (threadIdx.x == 0) de[r] = v;
}

// GetDeFromLatticeIdSite (EnergyChangePredictorPairSite.cpp:154-192)
__global__ __launch_bounds__(kThreads) void SiteDeKernel(KTables t, const uint8_t *occ,
                                                         const int64_t *off, const int32_t *ncell,
                                                         const int64_t *sites,
                                                         const uint8_t *codes, double *de) {
  __shared__ SiteLds L;
  const int r = blockIdx.x;
  const uint8_t *o = occ + off[r];
  if (!ValidSite(ncell[r], sites[r]) || o[sites[r]] >= t.S || codes[r] >= t.S) {
    if (threadIdx.x == 0) de[r] = __builtin_nan("");
    return;
  }
  LoadSiteTables(t, L);
  __syncthreads();
  const int eo = o[sites[r]], en = codes[r];
  if (eo == en) {
    if (threadIdx.x == 0) de[r] = 0.0;
    return;
  }
  GatherSite(L, L.env[0], o, ncell[r], sites[r], t.S);
  __syncthreads();
  const double v = BlockSum(L, SitePartial(t, L, 0, eo, en, -1));
  if (threadIdx.x == 0) de[r] = v;
}

// CanonicalMcSerial (CanonicalMcSerial.cpp:40-51, CanonicalMcAbstract.cpp:43-101) with an
// optional linear temperature ramp t0 -> t1 over maximum_steps.
__global__ __launch_bounds__(kThreads) void CmcKernel(KTables t, CmcDev c, uint64_t ntrials,
                                                      uint64_t trace_base) {
  __shared__ SiteLds L;
  __shared__ int64_t s_p, s_q;
  __shared__ int s_done;
  const int r = blockIdx.x;
  const int tid = threadIdx.x;
  LoadSiteTables(t, L);
  uint8_t *occ = c.occ + c.off[r];
  const int n = c.ncell[r];
  const uint64_t N = 4ull * n * n * n;
  uint64_t steps = 0, maxs = 0, acc = 0;
  double energy = 0, abse = 0;
  if (tid == 0) {
    steps = c.steps[r];
    maxs = c.max_steps[r];
    acc = c.accepted[r];
    energy = c.energy[r];
    abse = c.abs_energy[r];
    s_done = steps > maxs ? 1 : 0;
  }
  __syncthreads();
  for (uint64_t it = 0; it < ntrials; ++it) {
    if (s_done) break;
    if (tid == 0) {
      int64_t p = 0, q = 0;
      uint32_t att = 0;
      for (;; ++att) {
        if (att >= (1u << 20)) {
          atomicOr(c.error, 2);
          s_done = 1;
          break;
        }
        const U4 u = Philox4x32_10(c.seed[r], steps, c.rid[r], 2u * att + 1u);
        p = static_cast<int64_t>(MulHi64(Join64(u.x, u.y), N));
        q = static_cast<int64_t>(MulHi64(Join64(u.z, u.w), N));
        if (occ[p] != occ[q]) break;
      }
      s_p = p;
      s_q = q;
    }
    __syncthreads();
    if (s_done) break;
    const int64_t p = s_p, q = s_q;
    const double de = SwapDeBlock(t, L, occ, n, p, q);
    if (tid == 0) {
      const double T = c.t0[r] + (c.t1[r] - c.t0[r]) *
                                     (static_cast<double>(steps) / static_cast<double>(maxs));
      const double beta = 1.0 / kBoltzmann / T;
      bool accept = de < 0;
      if (!accept) {
        const U4 u = Philox4x32_10(c.seed[r], steps, c.rid[r], 0xFFFFFFFEu);
        accept = UnitOpenRight(Join64(u.x, u.y)) < exp(-de * beta);
      }
      const uint64_t ti = trace_base + it;
      if (ti < c.trace_cap) {
        const uint64_t o = static_cast<uint64_t>(r) * c.trace_cap + ti;
        c.tr_p[o] = p;
        c.tr_q[o] = q;
        c.tr_de[o] = de;
        c.tr_acc[o] = accept ? 1 : 0;
      }
      if (accept) {
        const uint8_t a = occ[p];
        occ[p] = occ[q];
        occ[q] = a;
        energy += de;
        abse += de;
        ++acc;
      }
      ++steps;
      if (steps > maxs) s_done = 1;
    }
    __syncthreads();
  }
  if (tid == 0) {
    c.steps[r] = steps;
    c.accepted[r] = acc;
    c.energy[r] = energy;
    c.abs_energy[r] = abse;
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
  const int r = blockIdx.y;
  const int n = ncell[r];
  const int64_t N = 4ll * n * n * n;
  const uint8_t *o = occ + off[r];
  for (int i = threadIdx.x; i < t.n_eci; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  const int S = t.S;
  for (int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < N;
       s += static_cast<int64_t>(gridDim.x) * blockDim.x) {
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

hipError_t LaunchEvalJumps(const KTables &t, const uint8_t *occ, const int64_t *off,
                           const int32_t *ncell, const int64_t *vac, int32_t R, int64_t *nb,
                           double *ea, double *de, int32_t *err, hipStream_t s) {
  if (R <= 0) return hipSuccess;
  hipLaunchKernelGGL(EvalJumpsKernel, dim3(R), dim3(kThreads), 0, s, t, occ, off, ncell, vac, nb,
                     ea, de, err);
  return hipGetLastError();
}

hipError_t LaunchKmc(const KTables &t, const KmcDev &k, uint64_t steps, uint64_t trace_base,
                     hipStream_t s) {
  if (k.R <= 0 || steps == 0) return hipSuccess;
  hipLaunchKernelGGL(KmcKernel, dim3(k.R), dim3(kThreads), 0, s, t, k, steps, trace_base);
  return hipGetLastError();
}

hipError_t LaunchSwapDe(const KTables &t, const uint8_t *occ, const int64_t *off,
                        const int32_t *ncell, const int64_t *pairs, int32_t R, double *de,
                        hipStream_t s) {
  if (R <= 0) return hipSuccess;
  hipLaunchKernelGGL(SwapDeKernel, dim3(R), dim3(kThreads), 0, s, t, occ, off, ncell, pairs, de);
  return hipGetLastError();
}

hipError_t LaunchSiteDe(const KTables &t, const uint8_t *occ, const int64_t *off,
                        const int32_t *ncell, const int64_t *sites, const uint8_t *codes,
                        int32_t R, double *de, hipStream_t s) {
  if (R <= 0) return hipSuccess;
  hipLaunchKernelGGL(SiteDeKernel, dim3(R), dim3(kThreads), 0, s, t, occ, off, ncell, sites, codes,
                     de);
  return hipGetLastError();
}

hipError_t LaunchCmc(const KTables &t, const CmcDev &c, uint64_t trials, uint64_t trace_base,
                     hipStream_t s) {
  if (c.R <= 0 || trials == 0) return hipSuccess;
  hipLaunchKernelGGL(CmcKernel, dim3(c.R), dim3(kThreads), 0, s, t, c, trials, trace_base);
  return hipGetLastError();
}

hipError_t LaunchTotalEnergy(const KTables &t, const uint8_t *occ, const int64_t *off,
                             const int32_t *ncell, int32_t R, unsigned long long *counts,
                             double *energy, hipStream_t s) {
  if (R <= 0) return hipSuccess;
  const int64_t nc = static_cast<int64_t>(R) * t.n_eci;
  hipLaunchKernelGGL(ZeroKernel, dim3((nc + 255) / 256), dim3(256), 0, s, counts, nc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(TotalCountsKernel, dim3(64, R), dim3(kThreads), 0, s, t, occ, off, ncell,
                     counts);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(TotalEnergyKernel, dim3((R + 63) / 64), dim3(64), 0, s, t, counts, R, energy);
  return hipGetLastError();
}

}  // namespace lmc
