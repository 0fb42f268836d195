// gfx950 kernels of the swap / site energy change (EnergyChangePredictorPairSite), the
// CanonicalMc driver and the full-lattice cluster-expansion energy (EnergyPredictor).
#include <hip/hip_runtime.h>

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
// SimulatedAnnealing::UpdateTemperature (SimulatedAnnealing.cpp:99-133) after the trial of step
// `steps`, with maximum_steps N and the energy after the trial.
__device__ void AnnealUpdate(CmcAnneal &A, bool accepted, uint64_t steps, uint64_t N,
                             double energy) {
  const uint64_t window = max(1ull, static_cast<unsigned long long>(N * 0.001));
  const uint64_t trigger = max(1ull, static_cast<unsigned long long>(N * 0.05));
  const uint64_t cooldown = max(1ull, static_cast<unsigned long long>(N * 0.10));
  ++A.window_trials;
  if (accepted) {
    ++A.window_accepts;
    if (energy < A.recent_best - 1e-4) {
      A.recent_best = energy;
      A.last_improvement = steps;
    }
  }
  if (A.window_trials >= window) {
    const double acc = static_cast<double>(A.window_accepts) / static_cast<double>(A.window_trials);
    if (acc > 0.50) A.T *= 0.99;
    A.window_trials = 0;
    A.window_accepts = 0;
  }
  const double acc_est = A.window_trials > 0u ? static_cast<double>(A.window_accepts) /
                                                     static_cast<double>(A.window_trials)
                                               : 1.0;
  if (A.reheats < 5 && steps - A.last_improvement >= trigger &&
      steps - A.last_reheat >= cooldown && acc_est < 0.05) {
    A.T *= 1.10;
    A.last_improvement = steps;
    A.last_reheat = steps;
    A.recent_best = energy;
    ++A.reheats;
  }
  A.T *= exp(-3.0 / static_cast<double>(max(1ull, static_cast<unsigned long long>(N))));
}

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
  CmcAnneal A{};
  if (tid == 0) {
    steps = c.steps[r];
    maxs = c.max_steps[r];
    acc = c.accepted[r];
    energy = c.energy[r];
    abse = c.abs_energy[r];
    s_done = steps > maxs ? 1 : 0;
    if (c.anneal) A = c.anneal[r];
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
      double beta;
      if (c.anneal) {
        if (energy < A.lowest - 1e-4) A.lowest = energy;  // Dump (SimulatedAnnealing.cpp:80-97)
        beta = 1.0 / kBoltzmann / fmax(A.T, 1e-12);
      } else {
        const double T = c.t0[r] + (c.t1[r] - c.t0[r]) *
                                       (static_cast<double>(steps) / static_cast<double>(maxs));
        beta = 1.0 / kBoltzmann / T;
      }
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
      if (c.anneal) AnnealUpdate(A, accept, steps, maxs, energy);
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
    if (c.anneal) c.anneal[r] = A;
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
