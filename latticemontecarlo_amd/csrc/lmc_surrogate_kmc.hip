// gfx950 persistent KineticMcFirst kernel with the surrogate ensemble standing in for the
// jump evaluation, and its stateless counterpart (one evaluation per replica).
//
// Work mapping: one workgroup holds G replicas, kSub threads each (a block): G = 4 with 128
// threads each when the batch fills the chip, G = 1 with 256 threads when there are no more
// replicas than CUs. Per step:
//   1. gather    : each block gathers its replica's 12 jump environments into LDS;
//   2. ensemble  : wave w evaluates member w % 4 on the 16-row tiles (12 events + 4 unused rows)
//                  of G / (waves / 4) replicas. Layer 1 is computed transposed (hidden units as
//                  rows, events as columns), so its accumulators, rounded to bf16, are already the
//                  second layer's B fragments in a permuted k order; the second layer's W2
//                  fragments are loaded in that order. Hidden activations never leave registers;
//                  W1 and W2 fragments are prefetched one 16-unit tile ahead;
//   3. events    : per event, the member mean of (E_kra, dE) gives Ea = max(0, E_kra + dE/2);
//                  an event falls back to the cluster expansion when its environment is out of
//                  distribution (a vacancy off the endpoints, more solutes than any training row)
//                  or when the members disagree by more than the threshold (with a relevance
//                  window: only on events whose lowest member barrier is within `window` kT of
//                  the replica's lowest barrier);
//   4. fallback  : the flagged events of all G replicas are spread over the whole workgroup,
//                  an aligned power-of-two lane group per event; each group reduces its cluster
//                  sums by lane exchanges and its first lane finishes the event;
//   5. select    : ordered cumulative probabilities, Philox draw, residence time (as the plain
//                  kernel);
//   6. exact dE  : the selected jump's dE is recomputed from the state clusters when the
//                  surrogate supplied it, so the FP64 energy bookkeeping stays exact;
//   7. apply.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "lmc_jump_device.h"
#include "lmc_mlp.h"
#include "lmc_surrogate.h"

namespace lmc {
namespace {

constexpr int kMembers = 4;
constexpr int kTile = 16;    // rows of a replica's tile: its 12 events and 4 unused rows

template <int G, int kSub>
struct Shape {
  static constexpr int kBlock = G * kSub;
  static constexpr int kWaves = kBlock / 64;
  static constexpr int kQ = kWaves / kMembers;  // waves per member
  static constexpr int kTPW = G / kQ;           // replica tiles per wave
  static constexpr int kSubWaves = kSub / 64;
  static_assert(kQ * kTPW == G && kSubWaves * 64 == kSub && kSubWaves <= 4, "whole tiles and waves");
};

template <int C>
struct MlpDims {
  static constexpr int Kp = (kEnvUsed * C + 31) / 32 * 32;
  static constexpr int KS = Kp / 32;
};

template <int G>
struct SurLds {
  float b1[kMembers][kHidden], b2[kMembers][kHidden];  // biases of the hidden layers
  uint32_t w3[kMembers][kHidden];                      // bf16 W3 row pairs (out 0 | out 1 << 16)
  float yp[kMembers][G * kTile][kMlpOut];   // member outputs, normalised, without b3
  uint8_t none[kEnvRow];                    // input row of an absent replica (no channel matches)
  double dsum[G][4];                        // exact dE: per-wave sums
  int fbmask[G], sel[G];
  double dt[G];
};

// Per-lane element offset of the weight fragments (8 bf16 per lane), made opaque once per step
// so that the compiler recomputes each fragment's address from a uniform base (cheap) instead of
// keeping every address of the persistent loop live in registers.
struct LaneOffsets {
  uint32_t f;
};
template <int C>
__device__ __forceinline__ LaneOffsets MakeLaneOffsets(int lane) {
  LaneOffsets o{static_cast<uint32_t>(8 * lane)};
  asm volatile("" : "+v"(o.f));
  return o;
}

// A fragments of W1 (member base W1m of the fragment-ordered copy) for hidden units 16h ..
// 16h+15 (rows): one contiguous 1 KB load per k-step.
template <int C>
__device__ __forceinline__ void LoadW1(const uint16_t *W1m, int h, uint32_t lo,
                                       u32x4 (&a)[MlpDims<C>::KS]) {
  const uint16_t *p = W1m + h * MlpDims<C>::KS * 512;
#pragma unroll
  for (int s = 0; s < MlpDims<C>::KS; ++s) a[s] = Load8(p + 512 * s + lo);
}

// A fragments of W2^T (member base W2m of the fragment-ordered copy) for output units 16h ..
// 16h+15 (rows), in the chained k order: at k-step j lane group grp holds the inputs
// 32j + 4 grp + (0..3) and 32j + 16 + 4 grp + (0..3), which is where layer 1's accumulators of
// unit tiles 2j and 2j + 1 sit.
__device__ __forceinline__ void LoadW2(const uint16_t *W2m, int h, uint32_t lo, u32x4 (&a)[4]) {
  const uint16_t *p = W2m + h * 4 * 512;
#pragma unroll
  for (int j = 0; j < 4; ++j) a[j] = Load8(p + 512 * j + lo);
}

// Weight ring of a wave: kDepth slots of KS fragments. Slot h % kDepth holds W1 unit tile h
// (h < 8), then W2 unit tile h - 8 in its first four fragments; a slot is refilled as soon as
// its tile has been multiplied, so up to kDepth tiles of weights are in flight.
template <int C>
struct Ring {
  static constexpr int KS = MlpDims<C>::KS;
  static constexpr int kDepth = KS <= 4 ? 8 : 4;  // 128 / 96 registers for C = 2 / 3
  u32x4 v[kDepth][KS];
};

// Issues weight tile i (0..7: W1, 8..15: W2) into its ring slot.
template <int C>
__device__ __forceinline__ void RingIssue(Ring<C> &R, int i, const uint16_t *W1m,
                                          const uint16_t *W2m, const LaneOffsets &lo) {
  constexpr int D = Ring<C>::kDepth;
  if (i < 8) {
    LoadW1<C>(W1m, i, lo.f, R.v[i % D]);
  } else {
    u32x4 a[4];
    LoadW2(W2m, i - 8, lo.f, a);
#pragma unroll
    for (int j = 0; j < 4; ++j) R.v[i % D][j] = a[j];
  }
}

// Member of this wave (uniform). Workgroups that run on the same XCD (every eighth block index)
// start on different members, so that their weight streams do not read the same lines at once.
__device__ __forceinline__ int WaveMember() {
  return (__builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6) +
          static_cast<int>(blockIdx.x >> 3)) & 3;
}

// The replica of tile slot g of this workgroup takes part in the step (uniform: LDS state
// written before the previous step's last barrier).
template <int G>
__device__ __forceinline__ bool Active(const KmcDev &k, const ReplicaLds *Qs, int g) {
  return static_cast<int>(blockIdx.x) * G + g < k.R && !Qs[g].done;
}

// Phase 2: member outputs (normalised, without b3) of every replica's 12 events into S.yp.
// The ring R holds the wave's first weight tiles, issued before the gather barrier.
template <int C, int G, int kSub>
__device__ void EnsembleForward(const SurDev &sd, const KmcDev &k, const JumpLds *Ls,
                                const ReplicaLds *Qs, const LaneOffsets &lo, Ring<C> &R,
                                SurLds<G> &S) {
  using Sh = Shape<G, kSub>;
  constexpr int KS = MlpDims<C>::KS, Kp = MlpDims<C>::Kp, TPW = Sh::kTPW, D = Ring<C>::kDepth;
  // with three channels the one-hot slot and channel indices of a lane are divisions by 3: the
  // lane comes from the per-step opaque offset, so that they are recomputed each step rather
  // than held through the persistent loop
  const int tid = threadIdx.x;
  const int lane = C == 3 ? static_cast<int>(lo.f >> 3) : (tid & 63), col = lane & 15, grp = lane >> 4;
  const int m = WaveMember();
  const int q = __builtin_amdgcn_readfirstlane(tid >> 6) >> 2;
  const MlpDev &w = sd.w;
  const uint16_t *W1m = w.w1f + static_cast<size_t>(m) * kHidden * Kp;
  const uint16_t *W2m = w.w2f + static_cast<size_t>(m) * kHidden * kHidden;
  const int chan[4] = {w.chan[0], w.chan[1], w.chan[2], w.chan[3]};
  // B fragments of the one-hot inputs: column col = event col of the tile's replica
  u32x4 x[TPW][KS];
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) {
    const int g = q * TPW + tt;
    const uint8_t *e = (col < 12 && Active<G>(k, Qs, g)) ? Ls[g].env[col] : S.none;
#pragma unroll
    for (int s = 0; s < KS; ++s) x[tt][s] = OneHotFragment<C>(e, 32 * s + 8 * grp, chan);
  }
  // layer 1, transposed: z1^T = W1^T x^T; bf16(silu) packed as layer 2's B fragments
  u32x4 hb[TPW][4];
#pragma unroll
  for (int h = 0; h < 8; ++h) {
    const f32x4 b = *reinterpret_cast<const f32x4 *>(&S.b1[m][16 * h + 4 * grp]);
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < KS; ++s) acc = Mfma(R.v[h % D][s], x[tt][s], acc);
      uint32_t v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float z = acc[i] + b[i];
        v[i] = Bf16Bits(z * Sigmoid(z));
      }
      hb[tt][h >> 1][(h & 1) * 2] = v[0] | (v[1] << 16);
      hb[tt][h >> 1][(h & 1) * 2 + 1] = v[2] | (v[3] << 16);
    }
    RingIssue<C>(R, h + D, W1m, W2m, lo);
  }
  // layer 2 (z2^T = W2^T h1^T) and the output layer's partial sums over this lane's 4 units
  float y0[TPW], y1[TPW];
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) y0[tt] = y1[tt] = 0.f;
#pragma unroll
  for (int h = 0; h < 8; ++h) {
    const f32x4 b = *reinterpret_cast<const f32x4 *>(&S.b2[m][16 * h + 4 * grp]);
    const u32x4 w3 = *reinterpret_cast<const u32x4 *>(&S.w3[m][16 * h + 4 * grp]);
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int j = 0; j < 4; ++j) acc = Mfma(R.v[(h + 8) % D][j], hb[tt][j], acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float z = acc[i] + b[i];
        const float hh = Bf16Round(z * Sigmoid(z));
        y0[tt] += hh * Bf16Float(w3[i] & 0xFFFFu);
        y1[tt] += hh * Bf16Float(w3[i] >> 16);
      }
    }
    if (h + 8 + D < 16) RingIssue<C>(R, h + 8 + D, W1m, W2m, lo);
  }
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) {
    y0[tt] += __shfl_xor(y0[tt], 16);
    y1[tt] += __shfl_xor(y1[tt], 16);
    y0[tt] += __shfl_xor(y0[tt], 32);
    y1[tt] += __shfl_xor(y1[tt], 32);
    if (grp == 0) {
      float *y = S.yp[m][kTile * (q * TPW + tt) + col];
      y[0] = y0[tt];
      y[1] = y1[tt];
    }
  }
}

// Phase 3 (lanes k < 12 of block g): surrogate events, fallback flags. Returns the flag.
template <int G>
__device__ bool EnsembleEvent(const KTables &t, const SurDev &sd, JumpLds &L, const SurLds<G> &S,
                              int g, int k, double beta) {
  const int row = kTile * g + k;
  double ek[kMembers], dv[kMembers];
  double ek_mean = 0.0, de_mean = 0.0;
#pragma unroll
  for (int m = 0; m < kMembers; ++m) {
    const float *par = sd.w.master + static_cast<size_t>(m) * sd.w.lay.P + sd.w.lay.ob3;
    const float y0 = par[0] + S.yp[m][row][0];
    const float y1 = par[1] + S.yp[m][row][1];
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
  bool relevant = true;
  if (sd.window > 0.0) {  // lowest ensemble barrier of the replica (lanes 0..11 of the block)
    double emin = ea;
#pragma unroll
    for (int j = 0; j < 12; ++j) emin = fmin(emin, __shfl(ea, j));
    relevant = lo < emin + sd.window / beta;
  }
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
  const bool flag = ((L.bad >> k) & 1) || solutes > sd.max_solutes || sd.threshold < 0.0 ||
                    !(spread == spread) || (relevant && spread > sd.threshold);
  if (!flag) StoreEvent(L, k, ea, de_mean, ek_mean, beta);
  return flag;
}

// Sums of an event's state list (dE, row wm) and orbit list (barrier logit, row bm) over the
// lanes lane, lane + kS, ...: four state and four orbit entries per lane at a time (eight
// independent load chains); entries past a list's end are read clamped and not added.
template <int kS>
__device__ __forceinline__ void StateOrbitGroup(const char *wm, const char *bm, const EventRanks &er,
                                                const uint32_t *sl, int ns, const uint32_t *ol,
                                                int no, int lane, double &de, double &la) {
  double a[4] = {0.0, 0.0, 0.0, 0.0}, b[4] = {0.0, 0.0, 0.0, 0.0};
  const int nmax = ns > no ? ns : no;
  for (int c = lane; c < nmax; c += 4 * kS) {
    uint32_t ds[4], dq[4];
    bool vs[4], vq[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = c + u * kS;
      vs[u] = i < ns;
      vq[u] = i < no;
      ds[u] = sl[vs[u] ? i : 0];
      dq[u] = ol[vq[u] ? i : 0];
    }
    double v[8];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t x1 = R8At(er, ds[u] & 0xFF), x2 = Rn8At(er, (ds[u] >> 8) & 0xFF);
      const uint32_t y0 = Rn8At(er, dq[u] & 0xFF), y1 = R8At(er, (dq[u] >> 8) & 0xFF);
      v[u] = *reinterpret_cast<const double *>(wm + (x1 + x2 + (ds[u] >> 16)));
      v[4 + u] = *reinterpret_cast<const double *>(bm + (y0 + y1 + (dq[u] >> 16)));
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] += vs[u] ? v[u] : 0.0;
      b[u] += vq[u] ? v[4 + u] : 0.0;
    }
  }
  de = (a[0] + a[1]) + (a[2] + a[3]);
  la = (b[0] + b[1]) + (b[2] + b[3]);
}

// Phase 4 (all threads): exact cluster expansion of the flagged events of every replica of the
// workgroup. The nf events get aligned lane groups of s threads (the largest power of two <= 64
// with s nf <= the workgroup size); a group's partial sums meet by butterfly exchanges and its
// first lane finishes the event.
template <int G, int kSub>
__device__ void FallbackEvents(const KTables &t, JumpLds *Ls, const ReplicaLds *Qs, const Tabs &T,
                               const SurLds<G> &S, int32_t *err) {
  constexpr int kBlock = Shape<G, kSub>::kBlock;
  const int tid = threadIdx.x;
  int cnt[G], nf = 0;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    cnt[g] = __popc(S.fbmask[g]);
    nf += cnt[g];
  }
  if (nf == 0) return;  // uniform
  int s = 64;
  while (s * nf > kBlock) s >>= 1;
  const int e = tid / s, lane = tid & (s - 1);
  const bool mine = e < nf;
  int g = 0, kk = 0;
  double de = 0.0, la = 0.0, lb = 0.0;
  if (mine) {
    int rest = e;
    while (rest >= cnt[g]) rest -= cnt[g++];
    int mm = S.fbmask[g];
    for (int i = 0; i < rest; ++i) mm &= mm - 1;
    kk = __ffs(mm) - 1;
    const JumpLds &L = Ls[g];
    const EventRanks &er = L.er[kk];
    const int M = L.env[kk][t.n_slot];
    if (t.kind == kE0 || t.kind == kQuartic) {
      const char *wm = WRow(t, T, M), *bm = BRow(T.b1, t.F1, t, M);
      if (s == 64)
        StateOrbitGroup<64>(wm, bm, er, T.state, t.n_state, T.mmm, t.n_mmm, lane, de, la);
      else if (s == 32)
        StateOrbitGroup<32>(wm, bm, er, T.state, t.n_state, T.mmm, t.n_mmm, lane, de, la);
      else if (s == 16)
        StateOrbitGroup<16>(wm, bm, er, T.state, t.n_state, T.mmm, t.n_mmm, lane, de, la);
      else
        StateOrbitGroup<8>(wm, bm, er, T.state, t.n_state, T.mmm, t.n_mmm, lane, de, la);
    } else {
      de = StateRange(WRow(t, T, M), er, T.state, 0, t.n_state, lane, s);
    }
    if (t.kind == kQuartic) lb = OrbitRange(BRow(T.b2, t.F2, t, M), er, T.mm2, t.n_mm2, lane, s);
  }
  for (int off = s >> 1; off > 0; off >>= 1) {  // uniform trip count
    de += __shfl_xor(de, off);
    la += __shfl_xor(la, off);
    lb += __shfl_xor(lb, off);
  }
  if (!mine || lane != 0) return;
  JumpLds &L = Ls[g];
  if ((t.kind == kE0 || t.kind == kQuartic) && ((L.bad >> kk) & 1)) la = __builtin_nan("");
  double ea, ekra;
  Barrier(t, L, kk, de, la, lb, ea, ekra);
  if (!(ea == ea) || !(de == de) || !(la == la)) atomicOr(err, 1);
  StoreEvent(L, kk, ea, de, ekra, Qs[g].beta);
}

// Phase 6 (all lanes of block g): exact dE of direction k from the state clusters; per-wave sums
// in S.dsum[g].
template <int G, int kSub>
__device__ void ExactDe(const KTables &t, const JumpLds &L, const Tabs &T, SurLds<G> &S, int g,
                        int lt, int k) {
  const int M = L.env[k][t.n_slot];
  const double v = WaveSum(StateRange(WRow(t, T, M), L.er[k], T.state, 0, t.n_state, lt, kSub));
  if ((lt & 63) == 0) S.dsum[g][lt >> 6] = v;
}

template <int C, int G, int kSub, bool kKmc, bool kBL>
__global__ __launch_bounds__(G * kSub) void SurrogateKernel(KTables t, DynLayout lay, KmcDev k,
                                                             SurDev sd, uint64_t nsteps,
                                                             uint64_t trace_base, SurOut out) {
  constexpr int kBlock = Shape<G, kSub>::kBlock, Kp = MlpDims<C>::Kp;
  __shared__ JumpLds Ls[G];
  __shared__ ReplicaLds Qs[G];
  __shared__ SurLds<G> S;
  extern __shared__ char dyn[];
  const int tid = threadIdx.x, g = tid / kSub, lt = tid - g * kSub;
  const int r = blockIdx.x * G + g;
  const bool live = r < k.R;
  JumpLds &L = Ls[g];
  ReplicaLds &Q = Qs[g];
  const Tabs T = LoadJumpTables<false, kBL>(t, L, dyn, lay, lt, kSub, tid, kBlock);
  uint8_t *occ = live ? k.occ + k.off[r] : nullptr;
  const int n = live ? k.ncell[r] : 0;
  if (tid < kEnvRow) S.none[tid] = 0xFF;
  for (int i = tid; i < kMembers * kHidden; i += kBlock) {  // constant through the launch
    const int m = i / kHidden, u = i - m * kHidden;
    const float *par = sd.w.master + static_cast<size_t>(m) * sd.w.lay.P;
    S.b1[m][u] = par[sd.w.lay.ob1 + u];
    S.b2[m][u] = par[sd.w.lay.ob2 + u];
    const uint16_t *w3 = sd.w.w3 + (static_cast<size_t>(m) * kHidden + u) * kMlpOut;
    S.w3[m][u] = static_cast<uint32_t>(w3[0]) | (static_cast<uint32_t>(w3[1]) << 16);
  }
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
  const uint16_t *W1m = sd.w.w1f + static_cast<size_t>(WaveMember()) * kHidden * Kp;
  const uint16_t *W2m = sd.w.w2f + static_cast<size_t>(WaveMember()) * kHidden * kHidden;
  uint64_t n_sur = 0, n_fb = 0;
  for (uint64_t it = 0; it < nsteps; ++it) {
    bool any = false;
#pragma unroll
    for (int j = 0; j < G; ++j) any = any || Active<G>(k, Qs, j);
    if (!any) break;  // uniform: every thread reads the same LDS state
    const bool on = live && !Q.done;
    if (kKmc && k.n_tt > 0) {  // before the draw lane reads T
      if (lt == 0 && on) {
        Q.T = ScheduleTemperature(k.tt_time, k.tt_temp, k.n_tt, Q.time);
        Q.beta = 1.0 / kBoltzmann / Q.T;
      }
      __syncthreads();
    }
    // lane index made opaque per step: the gather's per-lane offsets, slot indices and predicates
    // are recomputed each step instead of being held in registers through the ensemble phase
    int lg = lt;
    asm volatile("" : "+v"(lg));
    Gathered<kSub> gat;
    if (on) GatherIssue<kSub>(L, occ, n, VacancyXYZ(Q), lg, gat);
    const LaneOffsets lo = MakeLaneOffsets<C>(tid & 63);
    if (kKmc && on && lt == kSub - 1) DrawStep(L, Q);
    if (on) GatherStore<kSub>(t, L, k.error, lg, gat);
    Ring<C> ring;  // the first weight tiles, issued once the occupancy has arrived
#pragma unroll
    for (int i = 0; i < Ring<C>::kDepth; ++i) RingIssue<C>(ring, i, W1m, W2m, lo);
    __syncthreads();
    EnsembleForward<C, G, kSub>(sd, k, Ls, Qs, lo, ring, S);
    __syncthreads();
    if (lt < 12) {
      const bool f = on && EnsembleEvent(t, sd, L, S, g, lt, Q.beta);
      const unsigned long long b = __ballot(f) & 0xFFFull;
      if (lt == 0) {
        S.fbmask[g] = static_cast<int>(b);
        if (on) {
          n_fb += __popcll(b);
          n_sur += 12 - __popcll(b);
        }
      }
    }
    __syncthreads();
    FallbackEvents<G, kSub>(t, Ls, Qs, T, S, k.error);
    __syncthreads();
    if (!kKmc) {
      if (on && lt < 12) {
        const int64_t o = static_cast<int64_t>(r) * 12 + lt;
        out.nb[o] = L.snb[lt];
        out.ea[o] = L.ea[lt];
        out.de[o] = L.de[lt];
        out.flag[o] = static_cast<int8_t>((S.fbmask[g] >> L.dir[lt]) & 1);
      }
      break;  // one evaluation
    }
    if (on && lt < 12) {
      if (lt == 0) L.bad = 0;  // read by phases 3 and 4 only
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
    if (need) ExactDe<G, kSub>(t, L, T, S, g, lt, L.dir[sel]);
    __syncthreads();
    if (on && lt == 0) {
      double de = L.de[sel];
      if (need) {
        de = S.dsum[g][0];
#pragma unroll
        for (int j = 1; j < Shape<G, kSub>::kSubWaves; ++j) de += S.dsum[g][j];
      }
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
DynLayout PlanSurrogateLds(const KTables &t, int G) {
  const DeviceLimits &dl = CurrentDeviceLimits();
  const size_t lds_cu = dl.lds_cu, lds_block = dl.lds_block;
  const size_t stat = G * (sizeof(JumpLds) + sizeof(ReplicaLds)) +
                      (G == 1 ? sizeof(SurLds<1>) : sizeof(SurLds<4>)) + 256;
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

// Raises a kernel's dynamic LDS limit to `bytes` once (the largest size allowed so far is kept
// per kernel function; every instance of the surrogate kernel has the same pointer type).
template <typename K>
hipError_t AllowSurrogateLds(K kernel, size_t bytes) {
  static std::mutex mu;
  static std::map<const void *, int> allowed;
  const void *f = reinterpret_cast<const void *>(kernel);
  std::lock_guard<std::mutex> lock(mu);
  auto it = allowed.find(f);
  if (it != allowed.end() && static_cast<int>(bytes) <= it->second) return hipSuccess;
  const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           static_cast<int>(bytes));
  if (e == hipSuccess) allowed[f] = static_cast<int>(bytes);
  return e;
}

// Workgroup shapes: 4 replicas x 128 threads when the batch has more replicas than the device
// has CUs, else 1 replica x 256 threads (every replica gets a CU of its own).
int SurrogateGroup(int32_t R) { return R > CurrentDeviceLimits().cus ? 4 : 1; }

template <bool kKmc, int G, int kSub>
hipError_t LaunchSurrogateG(const KTables &t, const KmcDev &k, const SurDev &sd, uint64_t steps,
                            uint64_t trace_base, const SurOut &out, hipStream_t s) {
  const DynLayout lay = PlanSurrogateLds(t, G);
  const dim3 grid((k.R + G - 1) / G);
  auto go = [&](auto kernel) {
    hipError_t e = AllowSurrogateLds(kernel, lay.bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kernel, grid, dim3(G * kSub), lay.bytes, s, t, lay, k, sd, steps,
                       trace_base, out);
    return hipGetLastError();
  };
  if (lay.b1_lds) {
    if (sd.w.C == 2) return go(SurrogateKernel<2, G, kSub, kKmc, true>);
    if (sd.w.C == 3) return go(SurrogateKernel<3, G, kSub, kKmc, true>);
  } else {
    if (sd.w.C == 2) return go(SurrogateKernel<2, G, kSub, kKmc, false>);
    if (sd.w.C == 3) return go(SurrogateKernel<3, G, kSub, kKmc, false>);
  }
  return hipErrorInvalidValue;
}

template <bool kKmc>
hipError_t LaunchSurrogate(const KTables &t, const KmcDev &k, const SurDev &sd, uint64_t steps,
                           uint64_t trace_base, const SurOut &out, hipStream_t s) {
  if (k.R <= 0 || steps == 0) return hipSuccess;
  if (sd.w.M != kMembers) return hipErrorInvalidValue;
  if (SurrogateGroup(k.R) == 4) return LaunchSurrogateG<kKmc, 4, 128>(t, k, sd, steps, trace_base, out, s);
  return LaunchSurrogateG<kKmc, 1, 256>(t, k, sd, steps, trace_base, out, s);
}

template <int G, int kSub>
hipError_t PrepareSurrogateG(const KTables &t, int C) {
  const DynLayout lay = PlanSurrogateLds(t, G);
  if (lay.b1_lds) {
    if (C == 2) return AllowSurrogateLds(SurrogateKernel<2, G, kSub, true, true>, lay.bytes);
    if (C == 3) return AllowSurrogateLds(SurrogateKernel<3, G, kSub, true, true>, lay.bytes);
  } else {
    if (C == 2) return AllowSurrogateLds(SurrogateKernel<2, G, kSub, true, false>, lay.bytes);
    if (C == 3) return AllowSurrogateLds(SurrogateKernel<3, G, kSub, true, false>, lay.bytes);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t PrepareSurrogateKmc(const KTables &t, int C) {
  const hipError_t e = PrepareSurrogateG<4, 128>(t, C);
  return e != hipSuccess ? e : PrepareSurrogateG<1, 256>(t, C);
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
