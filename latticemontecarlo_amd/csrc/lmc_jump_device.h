// Device-side building blocks of the vacancy-jump evaluation and the KineticMcFirst step, shared
// by the plain kernels (lmc_jump_kernels.hip) and the surrogate kernels (lmc_surrogate.hip).
// Every function works on one replica's 256-thread block; `tid` is the thread's index in it.
#pragma once
#include <hip/hip_runtime.h>

#include "lmc_dev_common.h"
#include "lmc_device.h"
#include "lmc_rng.h"

namespace lmc {
namespace {

constexpr int kLanes = 21;      // dense: lanes per jump event (3 events per wave)
static_assert(kClusterPad == 4 * kLanes, "cluster lists are padded to whole dense batches");
constexpr int kSLanes = 16;     // solvent baseline: lanes per jump event (4 events per wave)

// Element ranks of one event's environment slots, pre-scaled for the cluster tables: rank * 8
// (bytes) and rank * 8 * ne (16-bit); slots 60..63 hold 0 (slot 63 is the absent site).
struct EventRanks {
  uint8_t r8[64];
  uint16_t rn8[64];
};

struct JumpLds {
  uint16_t off[12 * kEnvSlots];  // relative offsets, 4 bits per component (+8)
  double par[kMaxSpecies * 6];
  uint8_t env[12][64];
  EventRanks er[12];
  int64_t nb[12];   // by direction
  // by rank (neighbour-id order)
  double ea[12], de[12], rate[12], ekra[12];
  int64_t snb[12];
  int dir[12];
  // random numbers of the current step (drawn by an idle lane during the gather)
  double u2, log_u1, corr;
  // Philox u2 and log u1 of steps rnd_base .. rnd_base + kDrawAhead - 1, drawn together by the
  // lanes of one wave (rnd_ok: the buffer holds them)
  double rnd_u2[64], rnd_lu1[64];
  uint64_t rnd_base;
  int rnd_ok;
  // selection probabilities by rank; per-event sums (dE, MMM logit, MM2 logit)
  double prob[12], sums[12][3];
  // events whose environment holds a vacancy off the jump endpoints (no one-hot type)
  int bad;
  // per-lane partial sums (dE, MMM logit, MM2 logit): dense [event * kLanes + lane], solvent
  // baseline [event * kSLanes + lane], surrogate fallback per thread
  double part[kThreads][3];
};

// Replica state of the persistent KMC kernel (lane 0 reads and writes it in phase 4).
struct ReplicaLds {
  uint64_t steps, max_steps, seed;
  double time, energy, abse, T, beta, cv, cs, smass;
  double com[3];
  int64_t vac, disp[3];
  uint32_t rid;
  int flags, esc, solvent, done;
  int vx, vy, vz, m2;  // vacancy coordinates (a/2 units) and the box edge 2n, advanced per jump
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

// Copies n elements from global memory to LDS with threads gtid of gsize, eight loads per thread
// in flight before their stores (the copy at a launch's start is latency-bound otherwise).
template <typename E>
__device__ __forceinline__ void CopyToLds(E *dst, const E *src, int n, int gtid, int gsize) {
  int i = gtid;
  for (; i + 7 * gsize < n; i += 8 * gsize) {
    E v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = src[i + u * gsize];
#pragma unroll
    for (int u = 0; u < 8; ++u) dst[i + u * gsize] = v[u];
  }
  for (; i < n; i += gsize) dst[i] = src[i];
}

// Per-replica LDS tables (threads tid of this replica's block) and the workgroup's shared
// dynamic tables (threads gtid of gsize); returns the resolved table pointers.
// kBL: the barrier weight tables live in LDS (lay.b1_lds, and lay.b2_lds for the quartic model),
// known at compile time so that their reads are LDS instructions.
template <bool kSolvent, bool kBL, bool kWide = false>
__device__ Tabs LoadJumpTables(const KTables &t, JumpLds &L, char *dyn, const DynLayout &lay,
                               int tid, int lsize, int gtid, int gsize) {
  for (int i = tid; i < 12 * kEnvSlots; i += lsize) {
    int dx, dy, dz;
    Unpack(t.off[i], dx, dy, dz);
    L.off[i] = static_cast<uint16_t>((dx + 8) | ((dy + 8) << 4) | ((dz + 8) << 8));
  }
  for (int i = tid; i < 6 * t.S; i += lsize) L.par[i] = t.par[i];
  if (tid < 12 * 4) {  // slots 60..63; slot 63 is the target of every packed cluster's absent sites
    const int k = tid >> 2, j = kEnvSlots + (tid & 3);
    L.env[k][j] = 0;
    L.er[k].r8[j] = 0;
    L.er[k].rn8[j] = 0;
  }
  if (tid == 0) {
    L.bad = 0;
    L.rnd_ok = 0;  // the random-number buffer is refilled at the first step
  }
  Tabs T{};
  double *w = reinterpret_cast<double *>(dyn + lay.w);
  CopyToLds(w, t.wd, t.nwd, gtid, gsize);
  T.w = w;
  if (kBL) {
    double *p = reinterpret_cast<double *>(dyn + lay.b1);
    CopyToLds(p, t.b1, t.ne * t.F1, gtid, gsize);
    T.b1 = p;
    double *q = reinterpret_cast<double *>(dyn + lay.b2);
    if (t.kind == kQuartic)
      CopyToLds(q, t.b2, t.ne * t.F2, gtid, gsize);
    T.b2 = q;
  } else {
    T.b1 = t.b1;
    T.b2 = t.b2;
  }
  if (kSolvent) {
    double *bs = reinterpret_cast<double *>(dyn + lay.base);
    for (int i = gtid; i < 3 * t.S * t.S; i += gsize) bs[i] = t.base[i];
    T.base = bs;
    uint32_t *in = reinterpret_cast<uint32_t *>(dyn + lay.lists);
    CopyToLds(in, t.inc, t.n_inc, gtid, gsize);
    T.inc = in;
    uint16_t *io = reinterpret_cast<uint16_t *>(dyn + lay.offs);
    for (int i = gtid; i < 3 * (kEnvSlots + 1); i += gsize) io[i] = t.inc_off[i];
    T.inc_off = io;
  } else {
    // the wide shape reads the copies padded to its batch size
    const int ns = kWide ? t.n_state_w : t.n_state, no = kWide ? t.n_mmm_w : t.n_mmm;
    const int nq = kWide ? t.n_mm2_w : t.n_mm2;
    const uint32_t *gs = kWide ? t.state_w : t.state, *go = kWide ? t.mmm_w : t.mmm;
    const uint32_t *gq = kWide ? t.mm2_w : t.mm2;
    uint32_t *p = reinterpret_cast<uint32_t *>(dyn + lay.lists);
    CopyToLds(p, gs, ns, gtid, gsize);
    T.state = p;
    p += ns;
    CopyToLds(p, go, no, gtid, gsize);
    T.mmm = p;
    p += no;
    if (t.kind == kQuartic)
      CopyToLds(p, gq, nq, gtid, gsize);
    T.mm2 = p;
  }
  return T;
}

// Element rank of a species code (a vacancy gives 0).
__device__ __forceinline__ uint32_t RankOf(const KTables &t, int code) {
  return (t.rank_lut >> (4 * code)) & 15u;
}

// Row of the dE difference table (bytes) and of an expanded barrier table for migrating
// species code M.
__device__ __forceinline__ const char *WRow(const KTables &t, const Tabs &T, int M) {
  return reinterpret_cast<const char *>(T.w) + __umul24(RankOf(t, M), 8 * t.wrow);
}
__device__ __forceinline__ const char *BRow(const double *b, int F, const KTables &t, int M) {
  return reinterpret_cast<const char *>(b) + __umul24(RankOf(t, M), 8 * F);
}

// Vacancy coordinates of the replica state from its site id (once per launch).
__device__ __forceinline__ void SetVacancyXYZ(ReplicaLds &Q, int n) {
  const XYZ c = ToXYZ(Q.vac, n);
  Q.vx = c.X;
  Q.vy = c.Y;
  Q.vz = c.Z;
  Q.m2 = 2 * n;
}
__device__ __forceinline__ XYZ VacancyXYZ(const ReplicaLds &Q) { return {Q.vx, Q.vy, Q.vz}; }

// Phase 1: species of the 12 environments; neighbour ids of the 12 jumps. 720 sites over the
// kBT threads of the replica's block: independent loads per thread, issued together. Split in
// two so that a caller can issue other loads between them.
template <int kBT = kThreads>
struct Gathered {
  static constexpr int kU = (12 * kEnvSlots + kBT - 1) / kBT;
  int64_t s[kU];
  uint8_t v[kU];
};

template <int kBT = kThreads>
__device__ __forceinline__ void GatherIssue(const JumpLds &L, const uint8_t *occ, int n,
                                            const XYZ &c, int tid, Gathered<kBT> &G) {
  constexpr int kTotal = 12 * kEnvSlots;
#pragma unroll
  for (int u = 0; u < Gathered<kBT>::kU; ++u) {
    const int i = tid + u * kBT;
    if (i < kTotal) {
      int dx, dy, dz;
      Unpack16(L.off[i], dx, dy, dz);
      G.s[u] = SiteAt(c, n, dx, dy, dz);
      G.v[u] = occ[G.s[u]];
    }
  }
}

// Vacancy scan (ExitTime analysis): the stored configuration with site `site` holding species
// `code` (the real vacancy turned into solvent), and the vacancy placed at the jump origin.
struct GatherSub {
  int64_t site;
  int code;
};

template <int kBT = kThreads, bool kScan = false>
__device__ __forceinline__ void GatherStore(const KTables &t, JumpLds &L, int32_t *err, int tid,
                                            const Gathered<kBT> &G, const GatherSub &sub = {-1, 0}) {
  constexpr int kTotal = 12 * kEnvSlots;
#pragma unroll
  for (int u = 0; u < Gathered<kBT>::kU; ++u) {
    const int i = tid + u * kBT;
    if (i < kTotal) {
      const int k = i / kEnvSlots, j = i - k * kEnvSlots;
      uint8_t x = G.v[u];
      if (kScan) {
        if (G.s[u] == sub.site) x = static_cast<uint8_t>(sub.code);
        if (j == t.v_slot) x = static_cast<uint8_t>(t.x_code);
      }
      if (x >= t.S) {  // corrupt occupancy: flag it and keep every table index in range
        atomicOr(err, 4);
        x = 0;
      }
      if (x == t.x_code && j != t.v_slot) atomicOr(&L.bad, 1 << k);
      L.env[k][j] = x;
      const uint32_t rk = RankOf(t, x);
      L.er[k].r8[j] = static_cast<uint8_t>(rk << 3);
      L.er[k].rn8[j] = static_cast<uint16_t>(__umul24(rk, 8 * t.ne));
      if (j == t.n_slot) L.nb[k] = G.s[u];
    }
  }
}

template <int kBT = kThreads>
__device__ __forceinline__ void GatherEnv(const KTables &t, JumpLds &L, const uint8_t *occ, int n,
                                          const XYZ &vac, int32_t *err, int tid) {
  Gathered<kBT> G;
  GatherIssue<kBT>(L, occ, n, vac, tid, G);
  GatherStore<kBT>(t, L, err, tid, G);
}

// Byte offset of a packed entry's rank*8 slot (byte 1 of a state entry is the other one).
__device__ __forceinline__ uint32_t R8At(const EventRanks &er, uint32_t byte_off) {
  return er.r8[byte_off];
}
__device__ __forceinline__ uint32_t Rn8At(const EventRanks &er, uint32_t byte_off) {
  return *reinterpret_cast<const uint16_t *>(reinterpret_cast<const char *>(er.rn8) + byte_off);
}

// Sums of kU packed state clusters per lane: w(cluster | endpoint = M) - w(cluster | endpoint =
// X) from wm, the difference table's row of M (the entry's offset selects the negated copy for
// the neighbour endpoint's clusters). The loads of the kU clusters are issued level by level
// (entries, ranks, weights).
template <int kU>
__device__ __forceinline__ void StateBatch(const char *wm, const EventRanks &er,
                                           const uint32_t *list, int c, int stride, double *acc) {
  uint32_t d[kU], x1[kU], x2[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) d[u] = list[c + u * stride];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    x1[u] = R8At(er, d[u] & 0xFF);
    x2[u] = Rn8At(er, (d[u] >> 8) & 0xFF);
  }
#pragma unroll
  for (int u = 0; u < kU; ++u)
    acc[u] += *reinterpret_cast<const double *>(wm + (x1[u] + x2[u] + (d[u] >> 16)));
}

// Sum of the state clusters [c0, c1) of an event, lanes lane, lane + stride, ...
__device__ __forceinline__ double StateRange(const char *wm, const EventRanks &er,
                                             const uint32_t *list, int c0, int c1, int lane,
                                             int stride) {
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  int c = c0 + lane;
  for (; c + 3 * stride < c1; c += 4 * stride) StateBatch<4>(wm, er, list, c, stride, a);
  for (; c < c1; c += stride) StateBatch<1>(wm, er, list, c, stride, a);
  return (a[0] + a[1]) + (a[2] + a[3]);
}

// Sums of kU packed orbit clusters per lane: bm[rank0 * ne + rank1] of the cluster's expanded
// weights (bm: the migrating species' row, bytes).
template <int kU>
__device__ __forceinline__ void OrbitBatch(const char *bm, const EventRanks &er,
                                           const uint32_t *list, int c, int stride, double *acc) {
  uint32_t d[kU], x0[kU], x1[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) d[u] = list[c + u * stride];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    x0[u] = Rn8At(er, d[u] & 0xFF);
    x1[u] = R8At(er, (d[u] >> 8) & 0xFF);
  }
#pragma unroll
  for (int u = 0; u < kU; ++u)
    acc[u] += *reinterpret_cast<const double *>(bm + (x0[u] + x1[u] + (d[u] >> 16)));
}

// Four state and four orbit clusters per lane at once (entries c + u kStride, u < 4): eight
// independent load chains, issued level by level.
template <int kStride>
__device__ __forceinline__ void StateOrbitBatch(const char *wm, const char *bm, const EventRanks &er,
                                                const uint32_t *sl, const uint32_t *ol, int c,
                                                double *a, double *b) {
  uint32_t ds[4], dq[4], x1[4], x2[4], y0[4], y1[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    ds[u] = sl[c + u * kStride];
    dq[u] = ol[c + u * kStride];
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    x1[u] = R8At(er, ds[u] & 0xFF);
    x2[u] = Rn8At(er, (ds[u] >> 8) & 0xFF);
    y0[u] = Rn8At(er, dq[u] & 0xFF);
    y1[u] = R8At(er, (dq[u] >> 8) & 0xFF);
  }
  const char *p[8];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    p[u] = wm + (x1[u] + x2[u] + (ds[u] >> 16));
    p[4 + u] = bm + (y0[u] + y1[u] + (dq[u] >> 16));
  }
  double v[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) v[u] = *reinterpret_cast<const double *>(p[u]);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    a[u] += v[u];
    b[u] += v[4 + u];
  }
}

// Dense sums of an event's state list (dE, row wm) and orbit list (logit, row bm), lanes lane,
// lane + kStride, ...: both lists are padded to whole batches (multiples of 4 kStride entries)
// and advance together.
template <int kStride = kLanes>
__device__ __forceinline__ void StateOrbitRange(const char *wm, const char *bm, const EventRanks &er,
                                                const uint32_t *sl, int ns, const uint32_t *ol,
                                                int no, int lane, double &de, double &la) {
  double a[4] = {0.0, 0.0, 0.0, 0.0}, b[4] = {0.0, 0.0, 0.0, 0.0};
  const int both = ns < no ? ns : no;
  int c = lane;
  for (; c < both; c += 4 * kStride) StateOrbitBatch<kStride>(wm, bm, er, sl, ol, c, a, b);
  for (; c < ns; c += 4 * kStride) StateBatch<4>(wm, er, sl, c, kStride, a);
  for (; c < no; c += 4 * kStride) OrbitBatch<4>(bm, er, ol, c, kStride, b);
  de = (a[0] + a[1]) + (a[2] + a[3]);
  la = (b[0] + b[1]) + (b[2] + b[3]);
}

// Sum of the orbit clusters [0, n) of an event, lanes lane, lane + stride, ...
__device__ __forceinline__ double OrbitRange(const char *bm, const EventRanks &er,
                                             const uint32_t *list, int n, int lane, int stride) {
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  int c = lane;
  for (; c + 3 * stride < n; c += 4 * stride) OrbitBatch<4>(bm, er, list, c, stride, a);
  for (; c < n; c += stride) OrbitBatch<1>(bm, er, list, c, stride, a);
  return (a[0] + a[1]) + (a[2] + a[3]);
}

// Phase 2, dense: per-lane partial sums of every cluster's weight (dE, MMM logit, MM2 logit).
// Wave w evaluates events 3w .. 3w + 2 (lanes 0..62), so that it can finish them alone.
__device__ void DenseClusters(const KTables &t, JumpLds &L, const Tabs &T, int tid) {
  const int w = tid >> 6, l = tid & 63;
  if (l >= 3 * kLanes) return;
  const int g = l / kLanes, k = 3 * w + g, lane = l - g * kLanes;
  const EventRanks &er = L.er[k];
  const int M = L.env[k][t.n_slot];
  double de, la = 0.0, lb = 0.0;
  if (t.kind == kE0 || t.kind == kQuartic)
    StateOrbitRange(WRow(t, T, M), BRow(T.b1, t.F1, t, M), er, T.state, t.n_state, T.mmm, t.n_mmm,
                    lane, de, la);
  else
    de = StateRange(WRow(t, T, M), er, T.state, 0, t.n_state, lane, kLanes);
  if (t.kind == kQuartic)  // forward and backward MM2 entries in one list
    lb = OrbitRange(BRow(T.b2, t.F2, t, M), er, T.mm2, t.n_mm2, lane, kLanes);
  L.part[k * kLanes + lane][0] = de;
  L.part[k * kLanes + lane][1] = la;
  L.part[k * kLanes + lane][2] = lb;
}

// Solvent baseline: change of one state cluster touching slot j relative to the all-solvent
// environment; zero when the cluster's other site o is a lower non-solvent slot (the cluster is
// counted there). rj, rnj: slot j's rank entries; sr, srn: the solvent's.
__device__ __forceinline__ double SolventState(const char *wm, const EventRanks &er,
                                               const uint8_t *e, uint32_t d, int j, uint32_t rj,
                                               uint32_t rnj, uint32_t sr, uint32_t srn, int sv) {
  const int o = (d >> 14) & 63, co = e[o];
  const bool second = (d >> 13) & 1, none = o == 63;
  const uint32_t blk = d & 8191;
  const uint32_t oa = blk + (second ? rnj + er.r8[o] : rj + er.rn8[o]);
  const uint32_t ob = blk + (second ? srn + (none ? 0u : sr) : sr + (none ? 0u : srn));
  const double v = *reinterpret_cast<const double *>(wm + oa) - *reinterpret_cast<const double *>(wm + ob);
  const double sv_ = __hiloint2double(__double2hiint(v) ^ static_cast<int>(d & 0x80000000u), __double2loint(v));
  return (!none && o < j && co != sv) ? 0.0 : sv_;
}

// Solvent baseline: change of one orbit cluster touching slot j.
__device__ __forceinline__ double SolventOrbit(const char *bm, const EventRanks &er,
                                               const uint8_t *e, uint32_t d, int j, uint32_t rj,
                                               uint32_t rnj, uint32_t sr, uint32_t srn, int sv) {
  const int o = (d >> 17) & 63, co = e[o];
  const bool role1 = (d >> 16) & 1, none = o == 63;
  const uint32_t base = d & 0xFFFF;
  const uint32_t ia = base + (role1 ? rj + er.rn8[o] : rnj + er.r8[o]);
  const uint32_t ib = base + (role1 ? sr + (none ? 0u : srn) : srn + (none ? 0u : sr));
  const double v = *reinterpret_cast<const double *>(bm + ia) - *reinterpret_cast<const double *>(bm + ib);
  return (!none && o < j && co != sv) ? 0.0 : v;
}

// Phase 2, solvent baseline: 16 lanes per event (threads 0..191). The event's non-solvent
// environment slots (endpoints excluded) come from wave ballots; the lanes stride over each such
// slot's incidence lists. Partial sums go to L.part[event * 16 + lane].
__device__ void SolventClusters(const KTables &t, JumpLds &L, const Tabs &T, int sv, int tid) {
  const int k = tid >> 4, lane = tid & 15;
  const bool active = tid < 12 * kSLanes;
  const uint8_t *e = L.env[active ? k : 0];
  uint64_t mask = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int j = lane + 16 * i;
    const bool f = active && j < kEnvSlots && j != t.v_slot && j != t.n_slot && e[j] != sv;
    const unsigned long long b = __ballot(f);
    mask |= ((b >> (16 * (k & 3))) & 0xFFFFull) << (16 * i);
  }
  if (!active) return;
  const EventRanks &er = L.er[k];
  const int M = e[t.n_slot];
  const char *wm = WRow(t, T, M);
  const char *b1 = BRow(T.b1, t.F1, t, M), *b2 = BRow(T.b2, t.F2, t, M);
  const uint32_t sr = RankOf(t, sv) << 3, srn = __umul24(RankOf(t, sv), 8 * t.ne);
  const bool orbits = t.kind == kE0 || t.kind == kQuartic, mm2 = t.kind == kQuartic;
  double de = 0.0, la = 0.0, lb = 0.0;
  const uint16_t *offS = T.inc_off, *offM = T.inc_off + 61, *offQ = T.inc_off + 122;
  // two slots at a time; their state, MMM (and MM2) lists advance together, branch-free (an
  // exhausted list reads entry 0 and contributes zero), so several load chains are in flight
  for (uint64_t mm = mask; mm;) {
    const int j0 = __ffsll(static_cast<unsigned long long>(mm)) - 1;
    mm &= mm - 1;
    const bool two = mm != 0;
    const int j1 = two ? __ffsll(static_cast<unsigned long long>(mm)) - 1 : j0;
    if (two) mm &= mm - 1;
    const uint32_t r0 = er.r8[j0], rn0 = er.rn8[j0], r1 = er.r8[j1], rn1 = er.rn8[j1];
    int s0 = offS[j0] + lane, s1 = offS[j1] + lane, m0 = offM[j0] + lane, m1 = offM[j1] + lane;
    const int s0e = offS[j0 + 1], m0e = orbits ? offM[j0 + 1] : 0;
    const int s1e = two ? offS[j1 + 1] : 0, m1e = two && orbits ? offM[j1 + 1] : 0;
    while (s0 < s0e || s1 < s1e || m0 < m0e || m1 < m1e) {
      const bool bs0 = s0 < s0e, bs1 = s1 < s1e, bm0 = m0 < m0e, bm1 = m1 < m1e;
      const double vs0 = SolventState(wm, er, e, T.inc[bs0 ? s0 : 0], j0, r0, rn0, sr, srn, sv);
      const double vs1 = SolventState(wm, er, e, T.inc[bs1 ? s1 : 0], j1, r1, rn1, sr, srn, sv);
      const double vm0 = SolventOrbit(b1, er, e, T.inc[bm0 ? m0 : offM[0]], j0, r0, rn0, sr, srn, sv);
      const double vm1 = SolventOrbit(b1, er, e, T.inc[bm1 ? m1 : offM[0]], j1, r1, rn1, sr, srn, sv);
      de += (bs0 ? vs0 : 0.0) + (bs1 ? vs1 : 0.0);
      la += (bm0 ? vm0 : 0.0) + (bm1 ? vm1 : 0.0);
      s0 += kSLanes;
      s1 += kSLanes;
      m0 += kSLanes;
      m1 += kSLanes;
    }
    if (mm2) {
      int q0 = offQ[j0] + lane, q1 = offQ[j1] + lane;
      const int q0e = offQ[j0 + 1], q1e = two ? offQ[j1 + 1] : 0;
      while (q0 < q0e || q1 < q1e) {
        const bool bq0 = q0 < q0e, bq1 = q1 < q1e;
        const double vq0 = SolventOrbit(b2, er, e, T.inc[bq0 ? q0 : offQ[0]], j0, r0, rn0, sr, srn, sv);
        const double vq1 = SolventOrbit(b2, er, e, T.inc[bq1 ? q1 : offQ[0]], j1, r1, rn1, sr, srn, sv);
        lb += (bq0 ? vq0 : 0.0) + (bq1 ? vq1 : 0.0);
        q0 += kSLanes;
        q1 += kSLanes;
      }
    }
  }
  L.part[tid][0] = de;
  L.part[tid][1] = la;
  L.part[tid][2] = lb;
}

// Ea and E_kra of event k from its cluster sums: dE, MMM logit la and MM2 logit lb.
__device__ __forceinline__ void Barrier(const KTables &t, const JumpLds &L, int k, double de,
                                        double la, double lb, double &ea, double &ekra) {
  const int M = L.env[k][t.n_slot];
  const double *p = L.par + 6 * M;
  ea = 0.0;
  ekra = 0.0;
  if (t.kind == kE0) {
    // E0 = exp(logit * sigma_e0 + mu_e0); Ea = max(0, E0 + dE/2) (...E0.cpp:147-158)
    const double e0 = exp((p[0] + la) * p[2] + p[1]);
    ea = fmax(0.0, e0 + de / 2);
    ekra = e0;
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
    ekra = ea - de / 2;  // kinetically resolved barrier: Ea(i->j) - Ea(j->i) = dE
  }
}

// Stores event k (direction) at its rank among the 12 neighbour ids, with its rate
// exp(-Ea beta) (JumpEvent forward rate, JumpEvent.cpp:6-13).
__device__ __forceinline__ void StoreEvent(JumpLds &L, int k, double ea, double de, double ekra,
                                           double beta) {
  int r = 0;
  const int64_t mine = L.nb[k];
#pragma unroll
  for (int j = 0; j < 12; ++j) r += L.nb[j] < mine;
  L.ea[r] = ea;
  L.de[r] = de;
  L.ekra[r] = ekra;
  L.rate[r] = exp(-ea * beta);
  L.snb[r] = mine;
  L.dir[r] = k;
}

// Phase 3: Ea, dE and rate per jump, stored at the jump's rank among the 12 neighbour ids. Each
// wave finishes the events whose partials its own lanes wrote (3 dense, 4 solvent-baseline), so
// no workgroup barrier separates this phase from the cluster sums.
template <bool kSolvent>
__device__ __forceinline__ void FinishEvents(const KTables &t, JumpLds &L, const Tabs &T, int sv,
                                             double beta, int32_t *err, int tid) {
  constexpr int kL = kSolvent ? kSLanes : kLanes;
  constexpr int kE = 64 / kL;  // events per wave
  const int w = tid >> 6, l = tid & 63;
  if (w * kE >= 12) return;
  // lane 3q + c adds the kL partials of component c of the wave's event q (LDS traffic of one
  // wave completes in program order)
  if (l < 3 * kE) {
    const int q = l / 3, c = l - 3 * q;
    const int e = w * kE + q;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kL; ++i) s += L.part[e * kL + i][c];
    L.sums[e][c] = s;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (l >= kE) return;
  const int k = w * kE + l;
  double de = L.sums[k][0], la = L.sums[k][1], lb = L.sums[k][2];
  if (kSolvent) {  // all-solvent sums of this (solvent, migrating species)
    const double *b = T.base + (static_cast<size_t>(sv) * t.S + L.env[k][t.n_slot]) * 3;
    de += b[0];
    la += b[1];
    lb += b[2];
  }
  if ((t.kind == kE0 || t.kind == kQuartic) && ((L.bad >> k) & 1))
    la = __builtin_nan("");  // no one-hot type for a vacancy (the reference throws)
  double ea, ekra;
  Barrier(t, L, k, de, la, lb, ea, ekra);
  if (!(ea == ea) || !(de == de) || !(la == la)) atomicOr(err, 1);  // unknown cluster type
  StoreEvent(L, k, ea, de, ekra, beta);
}

// Orders the LDS traffic of one wave around a point (its lanes then see each other's writes).
__device__ __forceinline__ void WaveFence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane k < 12 of wave 0: event k from its cluster sums in L.sums[k] (dE, MMM logit, MM2 logit):
// Ea, dE and the rate, stored at the jump's rank among the 12 neighbour ids.
__device__ __forceinline__ void FinishFromSums(const KTables &t, JumpLds &L, double beta,
                                               int32_t *err, int k) {
  if (k >= 12) return;
  const double de = L.sums[k][0], lb = L.sums[k][2];
  double la = L.sums[k][1];
  if ((t.kind == kE0 || t.kind == kQuartic) && ((L.bad >> k) & 1))
    la = __builtin_nan("");  // no one-hot type for a vacancy (the reference throws)
  double ea, ekra;
  Barrier(t, L, k, de, la, lb, ea, ekra);
  if (!(ea == ea) || !(de == de) || !(la == la)) atomicOr(err, 1);  // unknown cluster type
  StoreEvent(L, k, ea, de, ekra, beta);
}

// Phase 3, dense: wave 0 finishes all 12 events (after a workgroup barrier), so that the two FP64
// exponentials, the rank search and the partial sums of the 12 events are issued once, in 12
// lanes of one wave, instead of in 3 lanes of each of four waves. Lane 3e + c first adds the
// kLanes partials of component c of event e (in lane order, as before); lane e then finishes
// event e.
__device__ __forceinline__ void FinishDense(const KTables &t, JumpLds &L, double beta,
                                            int32_t *err, int tid) {
  if (tid >= 64) return;
  const int l = tid;
  if (l < 36) {
    const int e = l / 3, c = l - 3 * e;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kLanes; ++i) s += L.part[e * kLanes + i][c];
    L.sums[e][c] = s;
  }
  WaveFence();
  FinishFromSums(t, L, beta, err, l);
}

// Wide shape (kWideThreads threads per replica): wave k evaluates event k with its 64 lanes (the
// lists padded to kClusterPadWide, 4 entries of each per lane and batch), sums the lane partials
// by DPP wave sums, and lane 0 writes the event's sums to L.sums[k]; after a workgroup barrier,
// wave 0 finishes all 12 events (FinishFromSums) in the same wave that selects the jump.
constexpr int kWideThreads = 12 * 64;
static_assert(kClusterPadWide == 4 * 64, "wide lists are padded to whole batches of 64 lanes");

__device__ void WideSums(const KTables &t, JumpLds &L, const Tabs &T, int tid) {
  const int k = tid >> 6, lane = tid & 63;
  const EventRanks &er = L.er[k];
  const int M = L.env[k][t.n_slot];
  double de, la = 0.0, lb = 0.0;
  if (t.kind == kE0 || t.kind == kQuartic)
    StateOrbitRange<64>(WRow(t, T, M), BRow(T.b1, t.F1, t, M), er, T.state, t.n_state_w, T.mmm,
                        t.n_mmm_w, lane, de, la);
  else
    de = StateRange(WRow(t, T, M), er, T.state, 0, t.n_state_w, lane, 64);
  if (t.kind == kQuartic) lb = OrbitRange(BRow(T.b2, t.F2, t, M), er, T.mm2, t.n_mm2_w, lane, 64);
  de = WaveSum(de);
  la = WaveSum(la);
  lb = WaveSum(lb);
  if (lane != 0) return;
  L.sums[k][0] = de;
  L.sums[k][1] = la;
  L.sums[k][2] = lb;
}

// Random numbers of step Q.steps (one lane, overlapping the gather): Philox u1, u2, log(u1) and
// the RateCorrector factor (KineticMcFirstOmp.cpp:79-82, RateCorrector.hpp:17-24).
__device__ void DrawStep(JumpLds &L, const ReplicaLds &Q) {
  const U4 u = Philox4x32_10(Q.seed, Q.steps, Q.rid, 0u);
  L.u2 = UnitOpenRight(Join64(u.z, u.w));
  L.log_u1 = log(UnitOpenLeft(Join64(u.x, u.y)));
  L.corr = (Q.flags & 1) ? Q.cv / (1.64 * exp(-(0.66 / kBoltzmann / Q.T - 0.7))) / (1 - 13 * Q.cs)
                         : 1.0;
}

// Random numbers of step Q.steps from a buffer of kDrawAhead steps that the last wave of the block
// (threads kBT - 64 .. kBT - 1) refills when the step leaves it, one step per lane: Philox, the
// logarithm and their issue slots are spent once per 64 steps instead of in one lane per step.
// Same values as DrawStep. The buffer's state (L.rnd_ok, L.rnd_base) is read and written by
// that wave only; L.rnd_ok must be 0 at the start of a launch.
constexpr int kDrawAhead = 64;
template <int kBT>
__device__ void DrawSteps(JumpLds &L, const ReplicaLds &Q, int tid) {
  if (tid < kBT - 64) return;
  const int lane = tid - (kBT - 64);
  if (!L.rnd_ok || Q.steps - L.rnd_base >= static_cast<uint64_t>(kDrawAhead)) {  // uniform
    const U4 u = Philox4x32_10(Q.seed, Q.steps + lane, Q.rid, 0u);
    L.rnd_u2[lane] = UnitOpenRight(Join64(u.z, u.w));
    L.rnd_lu1[lane] = log(UnitOpenLeft(Join64(u.x, u.y)));
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane == 63) {
      L.rnd_base = Q.steps;
      L.rnd_ok = 1;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (lane != 63) return;
  const int i = static_cast<int>(Q.steps - L.rnd_base);
  L.u2 = L.rnd_u2[i];
  L.log_u1 = L.rnd_lu1[i];
  L.corr = (Q.flags & 1) ? Q.cv / (1.64 * exp(-(0.66 / kBoltzmann / Q.T - 0.7))) / (1 - 13 * Q.cs)
                         : 1.0;
}

// `draw`: the replica whose next random numbers the last wave draws while the environment loads
// are in flight (null: none). kScan: the gather applies `sub` (see GatherSub).
template <bool kSolvent, bool kScan = false, bool kTailSync = true>
__device__ __forceinline__ void EvaluateJumps(const KTables &t, JumpLds &L, const Tabs &T,
                                              const uint8_t *occ, int n, const XYZ &vac, int sv,
                                              double beta, int32_t *err, int tid,
                                              const ReplicaLds *draw = nullptr,
                                              const GatherSub &sub = {-1, 0}) {
  Gathered<> G;
  GatherIssue(L, occ, n, vac, tid, G);
  if (draw) DrawSteps<kThreads>(L, *draw, tid);
  GatherStore<kThreads, kScan>(t, L, err, tid, G, sub);
  __syncthreads();
  if (kSolvent) {
    SolventClusters(t, L, T, sv, tid);
    FinishEvents<true>(t, L, T, sv, beta, err, tid);
  } else {
    DenseClusters(t, L, T, tid);
    __syncthreads();
    FinishDense(t, L, beta, err, tid);
  }
  // without kTailSync (dense only), the caller's next reader of the events must be wave 0,
  // which finished them
  if (kTailSync || kSolvent) __syncthreads();
  else if (tid < 64) WaveFence();
}

// TimeTemperatureInterpolator::GetTemperature (TimeTemperatureInterpolator.cpp:44-65)
__device__ double ScheduleTemperature(const double *tt, const double *tv, int n, double time) {
  int i = 0;
  while (i < n && tt[i] < time) ++i;
  if (i == n) return tv[n - 1];
  if (i == 0 && time <= tt[0]) return tv[0];
  return tv[i - 1] + ((time - tt[i - 1]) / (tt[i] - tt[i - 1])) * (tv[i] - tv[i - 1]);
}


// Phase 4 (lanes i < 12): lane i divides rate i by the total (KineticMcFirstOmp.cpp:69-77);
// lane 0 then forms the ordered cumulative probabilities and takes the first one >= u2, or event
// 11 when there is none (KineticMcAbstract.cpp:106-123), and the residence time
// (KineticMcFirstOmp.cpp:79-82). Returns true on lane 0 with the event's rank.
__device__ __forceinline__ bool ChooseEvent(JumpLds &L, const ReplicaLds &Q, int i, int &sel,
                                            double &dt) {
  double total = 0.0;
#pragma unroll
  for (int j = 0; j < 12; ++j) total += L.rate[j];
  L.prob[i] = L.rate[i] / total;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (i != 0) return false;
  const double u2 = L.u2;
  double cum = 0.0;
  sel = 11;
#pragma unroll
  for (int j = 0; j < 11; ++j) {
    cum += L.prob[j];
    if (sel == 11 && !(cum < u2)) sel = j;
  }
  dt = -L.log_u1 / total / kPrefactor;
  if (Q.flags & 1) dt *= L.corr;
  return true;
}

// The dump cadence of KineticMcFirstAbstract::Dump (KineticMcAbstract.cpp:83-98): every
// log_every steps past 10 log_every; below that every 10^(floor(log10(steps + 1)) - 1) steps
// (every step below 99), at most log_every.
__device__ bool LogDue(uint64_t steps, uint64_t log_every) {
  uint64_t every = log_every;
  if (steps <= 10 * log_every) {
    uint64_t p = 1;
    for (uint64_t t = steps + 1; t >= 100; t /= 10) p *= 10;
    every = p < log_every ? p : log_every;
  }
  return steps % every == 0;
}

// The kmc_log.txt record of the step about to jump the vacancy to `jump` (one lane).
__device__ void LogStep(const KmcDev &k, const ReplicaLds &Q, int r, int64_t jump, double ea,
                        double de) {
  const uint64_t i = k.log_n[r];
  k.log_n[r] = i + 1;  // counts past the capacity too, so the host can report the loss
  if (i >= k.log_cap) return;
  KmcLogRecord &L = k.log[static_cast<uint64_t>(r) * k.log_cap + i];
  L.steps = Q.steps;
  L.time = Q.time;
  L.temperature = Q.T;
  L.energy = Q.energy;
  L.ea = ea;
  L.de = de;
  const int64_t o = k.off[r];
  const int32_t *sv = nullptr;
  L.atom = -1;
  if (k.atom) {
    L.atom = k.atom[o + jump];
    sv = k.shift + 3 * (o + k.atom[o + Q.vac]);
  }
  // Config::GetUnwrappedCartesianPositionOfLattice (Config.cpp:280-291) on the cubic box
  const double m = Q.m2, edge = k.log_edge[r];
  L.vac[0] = (Q.vx / m + (sv ? sv[0] : 0)) * edge;
  L.vac[1] = (Q.vy / m + (sv ? sv[1] : 0)) * edge;
  L.vac[2] = (Q.vz / m + (sv ? sv[2] : 0)) * edge;
  L.com[0] = Q.com[0];
  L.com[1] = Q.com[1];
  L.com[2] = Q.com[2];
}

// One lane: advances the clock and applies a jump of the vacancy to site `jump` (offset
// dx, dy, dz; migrating species mig) with barrier ea and energy change de
// (KineticMcAbstract.cpp:147-181). `escaped`: IsEscaped of the step.
// kRecord: atom tracking and log records are compiled in (taken when k.atom / k.log_every are
// set); kernels launched without either use the variant without them.
template <bool kRecord = true>
__device__ void ApplyJump(const KTables &t, const KmcDev &k, ReplicaLds &Q, uint8_t *occ, int r,
                          uint64_t trace_index, int64_t jump, int dx, int dy, int dz, int mig,
                          double ea, double de, double dt, bool escaped) {
  if (escaped) {
    if (k.esc_time) {
      k.esc_time[r] = Q.time;
      k.esc_step[r] = Q.steps;
      k.esc_site[r] = Q.vac;
    }
    Q.esc = 1;
    Q.steps = Q.max_steps + 1;
  }
  if (kRecord && k.log_every && LogDue(Q.steps, k.log_every)) LogStep(k, Q, r, jump, ea, de);
  if (trace_index < k.trace_cap) {
    const uint64_t o = static_cast<uint64_t>(r) * k.trace_cap + trace_index;
    k.tr_sel[o] = jump;
    k.tr_ea[o] = ea;
    k.tr_de[o] = de;
    k.tr_time[o] = Q.time + dt;
  }
  Q.time += dt;
  Q.energy += de;
  Q.abse += de;
  if ((Q.flags & 4) && mig != Q.solvent) {  // solute COM (KineticMcAbstract.cpp:163-177)
    const double wgt = k.mass[mig] / Q.smass;
    Q.com[0] -= dx * kHalfLattice * wgt;
    Q.com[1] -= dy * kHalfLattice * wgt;
    Q.com[2] -= dz * kHalfLattice * wgt;
  }
  if (kRecord && k.atom) {  // the vacancy at Q.vac is the lhs of the reference's jump pair
    const int64_t o = k.off[r];
    const int m = Q.m2, h = m >> 1;
    AtomJump(k.atom + o, k.shift + 3 * o, Q.vac, jump,
             ImageChange(Wrap(Q.vx + dx, m) - Q.vx, h), ImageChange(Wrap(Q.vy + dy, m) - Q.vy, h),
             ImageChange(Wrap(Q.vz + dz, m) - Q.vz, h));
  }
  Q.disp[0] += dx;
  Q.disp[1] += dy;
  Q.disp[2] += dz;
  Q.vx = Wrap(Q.vx + dx, Q.m2);
  Q.vy = Wrap(Q.vy + dy, Q.m2);
  Q.vz = Wrap(Q.vz + dz, Q.m2);
  occ[Q.vac] = static_cast<uint8_t>(mig);  // Config::LatticeJump (Config.cpp:431-456)
  occ[jump] = static_cast<uint8_t>(t.x_code);
  Q.steps += 1;
  Q.vac = jump;
  if (Q.steps > Q.max_steps || Q.esc) Q.done = 1;
}

// One lane: applies the chosen event `sel` of the first-order step with energy change de.
template <bool kRecord = true>
__device__ void ApplyEvent(const KTables &t, const KmcDev &k, const JumpLds &L, ReplicaLds &Q,
                           uint8_t *occ, int r, uint64_t trace_index, int sel, double dt,
                           double de) {
  bool all = false;
  if (Q.flags & 2) {  // IsEscaped (KineticMcAbstract.cpp:192-223)
    // walk the set bits of the shell mask without unrolling: an unrolled loop keeps one scalar
    // predicate per slot live through the whole kernel
    all = true;
    uint64_t m = t.vac_shell_mask;
#pragma unroll 1
    while (m && all) {
      const int j = __ffsll(static_cast<unsigned long long>(m)) - 1;
      m &= m - 1;
      all = L.env[0][j] == Q.solvent;
    }
  }
  const int dir = L.dir[sel];
  int dx, dy, dz;
  Unpack16(L.off[dir * kEnvSlots + t.n_slot], dx, dy, dz);
  ApplyJump<kRecord>(t, k, Q, occ, r, trace_index, L.snb[sel], dx, dy, dz, L.env[dir][t.n_slot], L.ea[sel],
            de, dt, all);
}

template <bool kRecord = true>
__device__ void SelectAndApply(const KTables &t, const KmcDev &k, JumpLds &L, ReplicaLds &Q,
                               uint8_t *occ, int r, uint64_t trace_index, int tid) {
  const int i = tid;
  if (i >= 12) return;
  if (i == 0) L.bad = 0;  // every event lane has read it in phase 3
  int sel;
  double dt;
  if (!ChooseEvent(L, Q, i, sel, dt)) return;
  ApplyEvent<kRecord>(t, k, L, Q, occ, r, trace_index, sel, dt, L.de[sel]);
}

// ---- Second-order (chain) steps: KineticMcChainOmpi (lmc/mc/src/KineticMcChainOmpi.cpp) ----

// Per first neighbour i of the vacancy k (in id order): its site, offset and species, and the
// event k -> i taken as the reverse of i -> k after the jump, with the total rate out of i.
struct ChainLds {
  int64_t nid[12];  // by direction
  int64_t sid[12];  // by rank
  uint16_t soff[12];
  uint8_t sm[12];
  double barrier[12], de[12], rate[12], back[12], total_i[12];
  double b[12];     // p_ki (1 - p_ik) of the selection
  int64_t prev;     // the site the vacancy came from
};

// Wave 0: the vacancy's 12 first neighbours ranked by id, with their offsets and species.
__device__ void ChainNeighbours(const KTables &t, const JumpLds &L, ChainLds &C,
                                const ReplicaLds &Q, const uint8_t *occ, int n, int tid) {
  if (tid >= 64) return;
  const bool on = tid < 12;
  int64_t id = 0;
  uint16_t o = 0;
  if (on) {
    o = L.off[tid * kEnvSlots + t.n_slot];
    int dx, dy, dz;
    Unpack16(o, dx, dy, dz);
    id = SiteAt(VacancyXYZ(Q), n, dx, dy, dz);
    C.nid[tid] = id;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (!on) return;
  int rk = 0;
#pragma unroll
  for (int j = 0; j < 12; ++j) rk += C.nid[j] < id;
  C.sid[rk] = id;
  C.soff[rk] = o;
  C.sm[rk] = occ[id];
}

// Lane 0 after the pass with the vacancy at neighbour i (rank order): the total rate out of i
// and the event k -> i as the reverse of i -> k (KineticMcChainOmpi.cpp:56-91, JumpEvent.cpp:6-13,
// 25-30,51-54).
__device__ void ChainPass(JumpLds &L, ChainLds &C, const ReplicaLds &Q, int i) {
  double tot = 0.0, ea = 0.0, de = 0.0;
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    tot += L.rate[j];
    if (L.snb[j] == Q.vac) {
      ea = L.ea[j];
      de = L.de[j];
    }
  }
  C.total_i[i] = tot;
  C.barrier[i] = ea - de;
  C.de[i] = -de;
  C.rate[i] = exp(-C.barrier[i] * Q.beta);
  C.back[i] = exp((C.de[i] - C.barrier[i]) * Q.beta);
  L.bad = 0;  // every event lane has read it
}

// Lane 0: second-order probabilities and time (KineticMcChainOmpi.cpp:93-151; the rank sums in
// neighbour order), the selection with u2 (KineticMcAbstract.cpp:106-123) and the jump.
template <bool kRecord = true>
__device__ void ChainSelectAndApply(const KTables &t, const KmcDev &k, const JumpLds &L,
                                    ChainLds &C, ReplicaLds &Q, uint8_t *occ, int n, int r,
                                    uint64_t trace_index) {
  double total = 0.0;
  for (int i = 0; i < 12; ++i) total += C.rate[i];
  const double t1 = 1 / total / kPrefactor;
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double *b = C.b;
  int jprev = -1;
  for (int i = 0; i < 12; ++i) {
    const double p_ki = C.rate[i] / total;
    const double p_ik = C.back[i] / C.total_i[i];
    const double bb = p_ki * p_ik;
    b[i] = p_ki * (1 - p_ik);
    const bool prev = C.sid[i] == C.prev;
    if (prev) jprev = i;
    const double ti = 1 / C.total_i[i] / kPrefactor;
    s[0] += bb;
    s[1] += b[i];
    s[2] += prev ? 0.0 : bb;
    s[3] += prev ? 0.0 : b[i];
    s[4] += prev ? b[i] : 0.0;
    s[5] += prev ? p_ki : 0.0;
    s[6] += (t1 + ti) * bb;
    s[7] += prev ? 0.0 : (t1 + ti) * bb;
  }
  const double ts = s[6] / s[0], ts_j = s[7] / s[2];
  const double inv = 1 / (1 - s[5]);
  double dt = inv * (s[3] * t1 + s[2] * (ts_j + t1 + s[0] / s[1] * ts));
  if (Q.flags & 1) dt *= L.corr;
  const double u2 = L.u2;
  double cum = 0.0;
  int sel = 11;
  for (int i = 0; i < 11; ++i) {
    cum += i == jprev ? inv * (s[2] / s[1]) * s[4] : inv * (1 + s[2] / s[1]) * b[i];
    if (!(cum < u2)) {
      sel = i;
      break;
    }
  }
  bool all = false;
  if (Q.flags & 2) {  // IsEscaped (KineticMcAbstract.cpp:192-223), read from the occupancy
    all = true;
    uint64_t m = t.vac_shell_mask;
#pragma unroll 1
    while (m && all) {
      const int j = __ffsll(static_cast<unsigned long long>(m)) - 1;
      m &= m - 1;
      int dx, dy, dz;
      Unpack16(L.off[j], dx, dy, dz);
      all = occ[SiteAt(VacancyXYZ(Q), n, dx, dy, dz)] == Q.solvent;
    }
  }
  int dx, dy, dz;
  Unpack16(C.soff[sel], dx, dy, dz);
  const int64_t from = Q.vac;
  ApplyJump<kRecord>(t, k, Q, occ, r, trace_index, C.sid[sel], dx, dy, dz, C.sm[sel], C.barrier[sel],
            C.de[sel], dt, all);
  C.prev = from;  // KineticMcChainAbstract::OneStepSimulation (KineticMcAbstract.cpp:260-263)
}

}  // namespace
}  // namespace lmc
