// gfx950 persistent KineticMcFirst kernel with the surrogate ensemble standing in for the
// jump evaluation, and its stateless counterpart (one evaluation per replica).
//
// Work mapping: one 512-thread workgroup holds kG = 4 replicas (a 128-thread block each; two
// waves per SIMD leave 256 registers per lane for the ensemble). Per step:
//   1. gather    : each block gathers its replica's 12 jump environments into LDS;
//   2. ensemble  : the 8 waves evaluate the 4 members on the 48 event rows; wave (m, h) owns
//                  member m's columns 64h .. 64h+63 of both hidden layers, so every weight
//                  fragment is fetched once per workgroup step and used for 3 row tiles;
//   3. events    : per event, the member mean of (E_kra, dE) gives Ea = max(0, E_kra + dE/2);
//                  an event falls back to the cluster expansion when its environment is out of
//                  distribution (a vacancy off the endpoints) or when the members disagree by
//                  more than the threshold on an event whose rate is not negligible (lowest
//                  member barrier within `window` kT of the replica's lowest barrier);
//   4. fallback  : the flagged events of each replica are summed exactly, with all 128 lanes of
//                  its block spread over them;
//   5. select    : ordered cumulative probabilities, Philox draw, residence time (as the plain
//                  kernel);
//   6. exact dE  : the selected jump's dE is recomputed from the state clusters when the
//                  surrogate supplied it, so the FP64 energy bookkeeping stays exact;
//   7. apply.
#include <hip/hip_runtime.h>

#include "lmc_jump_device.h"
#include "lmc_mlp.h"
#include "lmc_surrogate.h"

namespace lmc {
namespace {

constexpr int kG = 4;                // replicas per workgroup
constexpr int kSub = 128;            // threads per replica
constexpr int kBlock = kG * kSub;
constexpr int kRowsS = 12 * kG;      // event rows per workgroup step
constexpr int kMT = kRowsS / 16;     // 16-row tiles
constexpr int kMembers = 4;

constexpr int kH1Stride = kHidden + 8;  // bf16 per h1 row: 272 B, so 16 rows hit distinct banks

struct SurLds {
  uint16_t h1[kMembers][kRowsS][kH1Stride];
  float yp[kMembers][2][kRowsS][kMlpOut];
  uint8_t none[kEnvRow];  // input row of an absent replica (no channel matches)
  double dsum[kG][2];     // exact dE: per-wave sums
  int fbmask[kG], sel[kG];
  double dt[kG];
};

// First-layer weight fragments of wave (m, h) for one of its 4 column tiles (KS K-steps). The
// first tile is loaded before the gather so that its latency overlaps it; each further tile is
// loaded while the previous one is multiplied (two tiles live instead of four).
template <int C>
struct TileWeights {
  static constexpr int KS = (kEnvUsed * C + 31) / 32;
  u32x4 b1[KS];
};

template <int C>
__device__ __forceinline__ void LoadTileWeights(const SurDev &sd, int nt, TileWeights<C> &tw) {
  constexpr int Kp = (kEnvUsed * C + 31) / 32 * 32, KS = Kp / 32;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, col = lane & 15, grp = lane >> 4;
  const int m = wv >> 1, q = wv & 1;
  const uint16_t *W1t = sd.w.w1t + static_cast<size_t>(m) * kHidden * Kp;
  const int n = 64 * q + 16 * nt + col;
#pragma unroll
  for (int s = 0; s < KS; ++s) tw.b1[s] = Load8(W1t + static_cast<size_t>(n) * Kp + 32 * s + 8 * grp);
}

// Phase 2: ensemble outputs (normalised) of the 48 rows into S.yp (all 8 waves).
template <int C>
__device__ void EnsembleForward(const SurDev &sd, const TileWeights<C> &first, const JumpLds *Ls,
                                const bool *act, SurLds &S) {
  constexpr int KS = TileWeights<C>::KS;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, col = lane & 15, grp = lane >> 4;
  const int m = wv >> 1, q = wv & 1;
  const MlpDev &w = sd.w;
  const float *par = w.master + static_cast<size_t>(m) * w.lay.P;
  const uint16_t *W2t = w.w2t + static_cast<size_t>(m) * kHidden * kHidden;
  const uint16_t *W3 = w.w3 + static_cast<size_t>(m) * kHidden * kMlpOut;
  const int chan[4] = {w.chan[0], w.chan[1], w.chan[2], w.chan[3]};
  u32x4 a[kMT][KS];  // one-hot fragments of the wave's rows, built once per step
#pragma unroll
  for (int mt = 0; mt < kMT; ++mt) {
    const int row = 16 * mt + col, g = row / 12, k = row - 12 * g;
    const uint8_t *e = act[g] ? Ls[g].env[k] : S.none;
#pragma unroll
    for (int s = 0; s < KS; ++s) a[mt][s] = OneHotFragment<C>(e, 32 * s + 8 * grp, chan);
  }
  TileWeights<C> cur = first, next;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const int n = 64 * q + 16 * nt + col;
    if (nt < 3) LoadTileWeights<C>(sd, nt + 1, next);
    f32x4 acc[kMT];
#pragma unroll
    for (int mt = 0; mt < kMT; ++mt) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int mt = 0; mt < kMT; ++mt) acc[mt] = Mfma(a[mt][s], cur.b1[s], acc[mt]);
    if (nt < 3) cur = next;
    const float bias = par[w.lay.ob1 + n];
#pragma unroll
    for (int mt = 0; mt < kMT; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float z = acc[mt][i] + bias;
        S.h1[m][16 * mt + 4 * grp + i][n] = static_cast<uint16_t>(Bf16Bits(z * Sigmoid(z)));
      }
  }
  u32x4 b2[4][4];  // second-layer fragments, in flight across the barrier
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int s = 0; s < 4; ++s)
      b2[nt][s] = Load8(W2t + static_cast<size_t>(64 * q + 16 * nt + col) * kHidden + 32 * s + 8 * grp);
  __syncthreads();
  // second layer and this wave's share of the output layer: per-lane partial sums over the
  // wave's 4 column tiles, then one reduction over the 16 lanes of each row group (each level's
  // 24 independent cross-lane moves are issued together)
  float p[kMT * 4 * kMlpOut] = {};
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const int n = 64 * q + 16 * nt + col;
    f32x4 acc[kMT];
#pragma unroll
    for (int mt = 0; mt < kMT; ++mt) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int mt = 0; mt < kMT; ++mt)
        acc[mt] = Mfma(Load8(&S.h1[m][16 * mt + col][32 * s + 8 * grp]), b2[nt][s], acc[mt]);
    const float bias = par[w.lay.ob2 + n];
    const float w30 = Bf16Float(W3[n * kMlpOut]);
    const float w31 = Bf16Float(W3[n * kMlpOut + 1]);
#pragma unroll
    for (int mt = 0; mt < kMT; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float z = acc[mt][i] + bias;
        const float h = Bf16Round(z * Sigmoid(z));
        p[(mt * 4 + i) * kMlpOut] += h * w30;
        p[(mt * 4 + i) * kMlpOut + 1] += h * w31;
      }
  }
#pragma unroll
  for (int off = 1; off < 16; off <<= 1) {
    float o[kMT * 4 * kMlpOut];
#pragma unroll
    for (int j = 0; j < kMT * 4 * kMlpOut; ++j) o[j] = __shfl_xor(p[j], off);
#pragma unroll
    for (int j = 0; j < kMT * 4 * kMlpOut; ++j) p[j] += o[j];
  }
  if (col == 0)
#pragma unroll
    for (int mt = 0; mt < kMT; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float *y = S.yp[m][q][16 * mt + 4 * grp + i];
        y[0] = p[(mt * 4 + i) * kMlpOut];
        y[1] = p[(mt * 4 + i) * kMlpOut + 1];
      }
}

// Phase 3 (lanes k < 12 of block g): surrogate events, fallback flags. Returns the flag.
__device__ bool EnsembleEvent(const KTables &t, const SurDev &sd, JumpLds &L, const SurLds &S,
                              int g, int k, double beta) {
  const int row = 12 * g + k;
  double ek[kMembers], dv[kMembers];
  double ek_mean = 0.0, de_mean = 0.0;
#pragma unroll
  for (int m = 0; m < kMembers; ++m) {
    const float *par = sd.w.master + static_cast<size_t>(m) * sd.w.lay.P + sd.w.lay.ob3;
    const float y0 = par[0] + (S.yp[m][0][row][0] + S.yp[m][1][row][0]);
    const float y1 = par[1] + (S.yp[m][0][row][1] + S.yp[m][1][row][1]);
    const double a = static_cast<double>(y0) * sd.sigma[0] + sd.mu[0];
    ek[m] = sd.log_ekra ? exp(a) : a;
    dv[m] = static_cast<double>(y1) * sd.sigma[1] + sd.mu[1];
    ek_mean += ek[m];
    de_mean += dv[m];
  }
  ek_mean *= 1.0 / kMembers;
  de_mean *= 1.0 / kMembers;
  const double ea = fmax(0.0, ek_mean + de_mean / 2);
  double lo = 1e300, hi = -1e300, dlo = 1e300, dhi = -1e300;
#pragma unroll
  for (int m = 0; m < kMembers; ++m) {
    const double e = fmax(0.0, ek[m] + dv[m] / 2);
    lo = fmin(lo, e);
    hi = fmax(hi, e);
    dlo = fmin(dlo, dv[m]);
    dhi = fmax(dhi, dv[m]);
  }
  double emin = ea;  // lowest ensemble barrier of the replica
#pragma unroll
  for (int j = 0; j < 12; ++j) emin = fmin(emin, __shfl(ea, j));
  int solutes = 0;  // environment slots holding neither the base species nor the vacancy
  {
    const uint32_t *words = reinterpret_cast<const uint32_t *>(L.env[k]);
    uint32_t wv[kEnvUsed / 4];
#pragma unroll
    for (int i = 0; i < kEnvUsed / 4; ++i) wv[i] = words[i];
#pragma unroll
    for (int i = 0; i < kEnvUsed / 4; ++i)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        solutes += (((wv[i] >> (8 * b)) & 0xFF) != static_cast<uint32_t>(sd.base_code)) ? 1 : 0;
    solutes -= (L.env[k][t.v_slot] != sd.base_code) ? 1 : 0;
  }
  const double spread = fmax(hi - lo, dhi - dlo);
  const bool relevant = sd.window <= 0.0 || lo < emin + sd.window / beta;
  const bool flag = ((L.bad >> k) & 1) || solutes > sd.max_solutes || sd.threshold < 0.0 ||
                    !(spread == spread) || (relevant && spread > sd.threshold);
  if (!flag) StoreEvent(L, k, ea, de_mean, ek_mean, beta);
  return flag;
}

// Phase 4a (all lanes of block g): exact cluster sums of the flagged events, kSub / nf lanes
// per event; partials in L.part.
__device__ void FallbackClusters(const KTables &t, JumpLds &L, const Tabs &T, const SurLds &S,
                                 int g, int lt) {
  const int mask = S.fbmask[g];
  const int nf = __popc(mask);
  const int le = kSub / nf;
  const int e = lt / le, lane = lt - e * le;
  double de = 0.0, la = 0.0, lb = 0.0;
  if (e < nf) {
    int mm = mask;
    for (int i = 0; i < e; ++i) mm &= mm - 1;
    const int k = __ffs(mm) - 1;
    const EventRanks &er = L.er[k];
    const int M = L.env[k][t.n_slot];
    de = StateRange(WRow(t, T, M), er, T.state, 0, t.n_state, lane, le);
    if (t.kind == kE0 || t.kind == kQuartic)
      la = OrbitRange(BRow(T.b1, t.F1, t, M), er, T.mmm, t.n_mmm, lane, le);
    if (t.kind == kQuartic)
      lb = OrbitRange(BRow(T.b2, t.F2, t, M), er, T.mm2, t.n_mm2, lane, le);
  }
  L.part[lt][0] = de;
  L.part[lt][1] = la;
  L.part[lt][2] = lb;
}

// Phase 4b (lane e < nf of block g): exact event e from its lanes' partials (in lane order).
__device__ void FallbackFinish(const KTables &t, JumpLds &L, const SurLds &S, int g, int e,
                               double beta, int32_t *err) {
  const int mask = S.fbmask[g];
  const int nf = __popc(mask);
  if (e >= nf) return;
  const int le = kSub / nf;
  int mm = mask;
  for (int i = 0; i < e; ++i) mm &= mm - 1;
  const int k = __ffs(mm) - 1;
  double de = 0.0, la = 0.0, lb = 0.0;
  for (int i = e * le; i < (e + 1) * le; ++i) {
    de += L.part[i][0];
    la += L.part[i][1];
    lb += L.part[i][2];
  }
  if ((t.kind == kE0 || t.kind == kQuartic) && ((L.bad >> k) & 1)) la = __builtin_nan("");
  double ea, ekra;
  Barrier(t, L, k, de, la, lb, ea, ekra);
  if (!(ea == ea) || !(de == de) || !(la == la)) atomicOr(err, 1);
  StoreEvent(L, k, ea, de, ekra, beta);
}

// Phase 6 (all lanes of block g): exact dE of direction k from the state clusters.
__device__ void ExactDe(const KTables &t, const JumpLds &L, const Tabs &T, SurLds &S, int g, int lt,
                        int k) {
  const int M = L.env[k][t.n_slot];
  double v = StateRange(WRow(t, T, M), L.er[k], T.state, 0, t.n_state, lt, kSub);
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) v += __shfl_xor(v, off);
  if ((lt & 63) == 0) S.dsum[g][lt >> 6] = v;
}

template <int C, bool kKmc, bool kBL>
__global__ __launch_bounds__(kBlock) void SurrogateKernel(KTables t, DynLayout lay, KmcDev k,
                                                          SurDev sd, uint64_t nsteps,
                                                          uint64_t trace_base, SurOut out) {
  __shared__ JumpLds Ls[kG];
  __shared__ ReplicaLds Qs[kG];
  __shared__ SurLds S;
  __shared__ bool act[kG];
  extern __shared__ char dyn[];
  const int tid = threadIdx.x, g = tid / kSub, lt = tid - g * kSub;
  const int r = blockIdx.x * kG + g;
  const bool live = r < k.R;
  JumpLds &L = Ls[g];
  ReplicaLds &Q = Qs[g];
  const Tabs T = LoadJumpTables<false, kBL>(t, L, dyn, lay, lt, kSub, tid, kBlock);
  uint8_t *occ = live ? k.occ + k.off[r] : nullptr;
  const int n = live ? k.ncell[r] : 0;
  if (tid < kEnvRow) S.none[tid] = 0xFF;
  if (lt == 0) {
    Q.done = 1;
    if (live && kKmc) {
      Q.steps = k.steps[r];
      Q.max_steps = k.max_steps[r];
      Q.seed = k.seed[r];
      Q.rid = k.rid[r];
      Q.time = k.time[r];
      Q.energy = k.energy[r];
      Q.abse = k.abs_energy[r];
      Q.cv = k.cv[r];
      Q.cs = k.cs[r];
      Q.smass = k.solute_mass[r];
      Q.flags = k.flags[r];
      Q.esc = k.escaped[r];
      Q.solvent = k.solvent[r];
      for (int a = 0; a < 3; ++a) {
        Q.disp[a] = k.disp[3 * r + a];
        Q.com[a] = k.com[3 * r + a];
      }
    }
    if (live) {  // the stateless mode reads only the vacancy and the temperature
      Q.vac = k.vac[r];
      Q.T = k.temp[r];
      Q.done = kKmc ? ((Q.steps > Q.max_steps || Q.esc || !ValidSite(n, Q.vac)) ? 1 : 0)
                    : (ValidSite(n, Q.vac) ? 0 : 1);
      if (!ValidSite(n, Q.vac)) atomicOr(k.error, 8);
      Q.beta = 1.0 / kBoltzmann / Q.T;
      SetVacancyXYZ(Q, n);
    }
  }
  __syncthreads();
  uint64_t n_sur = 0, n_fb = 0;
  for (uint64_t it = 0; it < nsteps; ++it) {
    if (!__syncthreads_or(live && !Q.done)) break;
    if (lt == 0) act[g] = live && !Q.done;
    if (kKmc && k.n_tt > 0 && lt == 0 && live && !Q.done) {
      Q.T = ScheduleTemperature(k.tt_time, k.tt_temp, k.n_tt, Q.time);
      Q.beta = 1.0 / kBoltzmann / Q.T;
    }
    __syncthreads();
    const bool on = act[g];
    Gathered<kSub> gat;
    if (on) GatherIssue<kSub>(L, occ, n, VacancyXYZ(Q), lt, gat);
    TileWeights<C> ww;  // issued after the occupancy loads, consumed after the gather
    LoadTileWeights<C>(sd, 0, ww);
    if (kKmc && on && lt == kSub - 1) DrawStep(L, Q);
    if (on) GatherStore<kSub>(t, L, k.error, lt, gat);
    __syncthreads();
    EnsembleForward<C>(sd, ww, Ls, act, S);
    __syncthreads();
    if (on && lt < 12) {
      const bool f = EnsembleEvent(t, sd, L, S, g, lt, Q.beta);
      const unsigned long long b = __ballot(f) & 0xFFFull;
      if (lt == 0) {
        S.fbmask[g] = static_cast<int>(b);
        n_fb += __popcll(b);
        n_sur += 12 - __popcll(b);
      }
    }
    __syncthreads();
    if (on && S.fbmask[g]) FallbackClusters(t, L, T, S, g, lt);
    __syncthreads();
    if (on && S.fbmask[g] && lt < 12) FallbackFinish(t, L, S, g, lt, Q.beta, k.error);
    __syncthreads();
    if (!kKmc) {
      if (on && lt < 12) {
        const int64_t o = static_cast<int64_t>(r) * 12 + lt;
        out.nb[o] = L.snb[lt];
        out.ea[o] = L.ea[lt];
        out.de[o] = L.de[lt];
        out.flag[o] = static_cast<int8_t>((S.fbmask[g] >> L.dir[lt]) & 1);
      }
      if (lt == 0 && live) Q.done = 1;
      continue;
    }
    if (on && lt < 12) {
      if (lt == 0) L.bad = 0;
      int sel;
      double dt;
      if (ChooseEvent(L, Q, lt, sel, dt)) {
        S.sel[g] = sel;
        S.dt[g] = dt;
      }
    }
    __syncthreads();
    const int sel = S.sel[g];
    const bool need = on && !((S.fbmask[g] >> L.dir[sel]) & 1);
    if (need) ExactDe(t, L, T, S, g, lt, L.dir[sel]);
    __syncthreads();
    if (on && lt == 0) {
      const double de = need ? S.dsum[g][0] + S.dsum[g][1] : L.de[sel];
      ApplyEvent(t, k, L, Q, occ, r, trace_base + it, sel, S.dt[g], de);
    }
    __syncthreads();
  }
  if (lt == 0 && live) {
    if (kKmc) {
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
    if (sd.n_sur) {
      sd.n_sur[r] += n_sur;
      sd.n_fb[r] += n_fb;
    }
  }
}

// LDS plan of the surrogate kernel: one workgroup per CU; the fallback's tables (dE weights,
// cluster lists) always in LDS, the barrier weights when they fit.
DynLayout PlanSurrogateLds(const KTables &t) {
  static size_t lds_cu = 0, lds_block = 0;
  if (!lds_cu) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) {
      lds_cu = p.maxSharedMemoryPerMultiProcessor ? p.maxSharedMemoryPerMultiProcessor : 65536;
      lds_block = p.sharedMemPerBlock;
    } else {
      lds_cu = lds_block = 65536;
    }
  }
  const size_t stat = kG * (sizeof(JumpLds) + sizeof(ReplicaLds)) + sizeof(SurLds) + 256;
  const size_t share = std::min(lds_block, lds_cu);
  auto al = [](size_t x) { return (x + 15) & ~size_t(15); };
  DynLayout lay{};
  size_t at = 0;
  lay.w = 0;
  at = al(sizeof(double) * t.nwd);
  lay.lists = static_cast<int>(at);
  at = al(at + sizeof(uint32_t) * (t.n_state + t.n_mmm + (t.kind == kQuartic ? t.n_mm2 : 0)));
  const size_t b1 = sizeof(double) * static_cast<size_t>(t.ne) * t.F1;
  const size_t b2 = sizeof(double) * static_cast<size_t>(t.ne) * t.F2;
  if (stat + at + b1 <= share) {
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
hipError_t AllowSurrogateLds(K kernel, size_t bytes) {
  static int allowed = -1;  // per kernel instance: the largest size already allowed
  if (static_cast<int>(bytes) <= allowed) return hipSuccess;
  const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           static_cast<int>(bytes));
  if (e == hipSuccess) allowed = static_cast<int>(bytes);
  return e;
}

template <bool kKmc>
hipError_t LaunchSurrogate(const KTables &t, const KmcDev &k, const SurDev &sd, uint64_t steps,
                           uint64_t trace_base, const SurOut &out, hipStream_t s) {
  if (k.R <= 0 || steps == 0) return hipSuccess;
  if (sd.w.M != kMembers) return hipErrorInvalidValue;
  const DynLayout lay = PlanSurrogateLds(t);
  const dim3 grid((k.R + kG - 1) / kG);
  auto go = [&](auto kernel) {
    hipError_t e = AllowSurrogateLds(kernel, lay.bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kernel, grid, dim3(kBlock), lay.bytes, s, t, lay, k, sd, steps,
                       trace_base, out);
    return hipGetLastError();
  };
  if (lay.b1_lds) {
    if (sd.w.C == 2) return go(SurrogateKernel<2, kKmc, true>);
    if (sd.w.C == 3) return go(SurrogateKernel<3, kKmc, true>);
  } else {
    if (sd.w.C == 2) return go(SurrogateKernel<2, kKmc, false>);
    if (sd.w.C == 3) return go(SurrogateKernel<3, kKmc, false>);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t PrepareSurrogateKmc(const KTables &t, int C) {
  const DynLayout lay = PlanSurrogateLds(t);
  if (lay.b1_lds) {
    if (C == 2) return AllowSurrogateLds(SurrogateKernel<2, true, true>, lay.bytes);
    if (C == 3) return AllowSurrogateLds(SurrogateKernel<3, true, true>, lay.bytes);
  } else {
    if (C == 2) return AllowSurrogateLds(SurrogateKernel<2, true, false>, lay.bytes);
    if (C == 3) return AllowSurrogateLds(SurrogateKernel<3, true, false>, lay.bytes);
  }
  return hipErrorInvalidValue;
}

hipError_t LaunchSurrogateKmc(const KTables &t, const KmcDev &k, const SurDev &sd, uint64_t steps,
                              uint64_t trace_base, hipStream_t s) {
  return LaunchSurrogate<true>(t, k, sd, steps, trace_base, SurOut{}, s);
}

hipError_t LaunchSurrogateEval(const KTables &t, const KmcDev &k, const SurDev &sd,
                               const SurOut &out, hipStream_t s) {
  return LaunchSurrogate<false>(t, k, sd, 1, 0, out, s);
}

}  // namespace lmc
