// C ABI of the engine: model loading, table upload, stateless evaluation and the batched
// KineticMcFirst / CanonicalMc drivers. Every compute entry point launches a HIP kernel; there
// is no host fallback.
#include "lmc_capi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "lmc_device.h"
#include "lmc_element.h"
#include "lmc_error.h"
#include "lmc_model.h"
#include "lmc_rng.h"
#include "lmc_surrogate.h"
#include "lmc_tables.h"

struct lmc_model {
  lmc::Model *m;
};

namespace {

using lmc::Check;
using lmc::Fail;
using lmc::Guard;
using lmc::HipError;
using lmc::ModelError;

int32_t PackOff(const lmc::I3 &o) {
  return (o.x & 0xff) | ((o.y & 0xff) << 8) | ((o.z & 0xff) << 16);
}

// Site-table entry: label | other1 << 4 | other2 << 10 (63 = none).
uint32_t PackSite(const lmc::StateCluster &c, int centre) {
  int other[2] = {63, 63}, k = 0;
  for (int i = 0; i < c.size; ++i)
    if (c.slots[i] != centre) other[k++] = c.slots[i];
  return static_cast<uint32_t>(c.label) | (static_cast<uint32_t>(other[0]) << 4) |
         (static_cast<uint32_t>(other[1]) << 10);
}

// Reduced dE-weight block of a cluster label over element ranks (n = number of elements):
// singlet 0, pairs 1 + (l - 1) n, triplets 1 + 3 n + (l - 4) n^2; entry r = block + r1 + n r2
// (ranks of the two other sites).
int ReducedBlock(int l, int n) {
  return l == 0 ? 0 : l < 4 ? 1 + (l - 1) * n : 1 + 3 * n + (l - 4) * n * n;
}

// Jump state-cluster entry: other1 [0:8] | 2 * other2 [8:16] | byte offset of its weight
// [16:32]. The kernel reads other1 from the event's rank*8 array (bytes) and other2 from its
// rank*8*ne array (16-bit words), so the two bytes are the array offsets themselves. The
// neighbour endpoint's clusters (opposite sign) index the negated copy of the table, at
// neg_bytes. Absent sites point at slot 63, whose rank entries are 0, so singlets and pairs
// index their blocks like triplets.
uint32_t PackState(const lmc::StateCluster &c, int endpoint_slot, bool n_side, int ne,
                   int neg_bytes) {
  int other[2] = {63, 63}, k = 0;
  for (int i = 0; i < c.size; ++i)
    if (c.slots[i] != endpoint_slot) other[k++] = c.slots[i];
  const int at = ReducedBlock(c.label, ne) * 8 + (n_side ? neg_bytes : 0);
  if (at > 65535) throw std::invalid_argument("dE weight table too large for the packed tables");
  return static_cast<uint32_t>(other[0]) | (static_cast<uint32_t>(2 * other[1]) << 8) |
         (static_cast<uint32_t>(at) << 16);
}

// Expanded barrier-feature layout: orbit o owns ne (singlet) or ne^2 (pair) consecutive
// weights, indexed by the element ranks (r0 ne + r1 for a pair, r for a singlet), so that the
// kernel needs no one-hot type lookup. Returns the bases and sets the total length.
std::vector<int> ExpandedBases(const lmc::OrbitTable &ot, int ne, int &total) {
  std::vector<int> b(ot.orbit_kind.size());
  total = 0;
  for (size_t o = 0; o < b.size(); ++o) {
    b[o] = total;
    total += ot.orbit_kind[o] == 0 ? ne : ne * ne;
  }
  return b;
}

// Expanded weights [ne][total] of one model's features b[S][F] (rows by element rank).
std::vector<double> ExpandWeights(const lmc::Model &m, const lmc::OrbitTable &ot,
                                  const std::vector<int> &base, const std::vector<double> &b,
                                  int F, const std::vector<int> &ebase, int total) {
  const int S = m.S, ne = m.ne;
  std::vector<int> code;  // element rank -> species code
  for (int c = 0; c < S; ++c)
    if (c != m.x_code) code.push_back(c);
  std::vector<double> v(static_cast<size_t>(ne) * total, 0.0);
  for (int r = 0; r < ne; ++r) {
    const double *row = b.data() + static_cast<size_t>(code[r]) * F;
    double *out = v.data() + static_cast<size_t>(r) * total;
    for (size_t o = 0; o < ebase.size(); ++o) {
      const int kind = ot.orbit_kind[o];
      const int n1 = kind == 0 ? 1 : ne;
      for (int r0 = 0; r0 < n1; ++r0)
        for (int r1 = 0; r1 < ne; ++r1) {
          const int a = kind == 0 ? code[r1] : code[r0], bb = kind == 0 ? code[0] : code[r1];
          const int ty = m.type_tab[(kind * S + a) * S + bb];
          if (ty < 0) throw std::logic_error("one-hot type of an element pair");
          out[ebase[o] + (kind == 0 ? r1 : r0 * ne + r1)] = row[base[o] + ty];
        }
    }
  }
  return v;
}

// Orbit-cluster entry: 2 * slot0 [0:8] | slot1 [8:16] | byte offset of its expanded weights
// [16:32]. slot0 is read from the event's rank*8*ne array, slot1 from its rank*8 array; a
// singlet's site is slot1 and its slot0 is 63 (rank entries 0). Backward entries read the
// species of canonical slot x from slot mirror[x].
std::vector<uint32_t> PackOrbits(const lmc::OrbitTable &ot, const std::vector<int> &ebase,
                                 const std::vector<int> *mirror) {
  std::vector<uint32_t> v;
  for (const auto &c : ot.clusters) {
    const uint32_t b = static_cast<uint32_t>(ebase[c.orbit]) * 8;
    if (b > 65535) throw std::invalid_argument("feature vector too long for the packed tables");
    int s0 = c.s0, s1 = c.s1;
    if (mirror) {
      s0 = (*mirror)[s0];
      s1 = s1 < 0 ? -1 : (*mirror)[s1];
    }
    if (s1 < 0) {  // singlet: its site takes the rank*8 role
      s1 = s0;
      s0 = 63;
    }
    v.push_back(static_cast<uint32_t>(2 * s0) | (static_cast<uint32_t>(s1) << 8) | (b << 16));
  }
  return v;
}

struct BlobBuilder {
  std::vector<char> bytes;
  std::vector<std::pair<size_t, const void **>> fix;
  template <typename T>
  void Add(const std::vector<T> &v, const T **dst) {
    const size_t at = (bytes.size() + 255) & ~size_t(255);
    bytes.resize(at + std::max<size_t>(v.size() * sizeof(T), 8), 0);
    if (!v.empty()) std::memcpy(bytes.data() + at, v.data(), v.size() * sizeof(T));
    fix.emplace_back(at, reinterpret_cast<const void **>(dst));
  }
};

int CurrentDevice() {
  int d = -1;
  Check(hipGetDevice(&d), "hipGetDevice");
  return d;
}

}  // namespace

namespace lmc {

Model::~Model() {
  if (dev) {
    if (dev->blob) (void)hipFree(dev->blob);
    delete dev;
  }
}

// Solvent-baseline tables. With every non-endpoint site of a jump environment holding the
// solvent s, the cluster sums of dE and of the barrier logits depend only on (s, M); a real
// environment adds, for each cluster with a non-solvent site, value(cluster) - value(all s).
// A cluster with several non-solvent sites is listed under each of them and evaluated only at
// the lowest-numbered one (the kernel checks the entry's other slot).
static void BuildSolventTables(const Model &m, const JumpTables &jt, const std::vector<int> &eb1,
                               const std::vector<int> &eb2, std::vector<uint16_t> &off,
                               std::vector<uint32_t> &inc, std::vector<double> &base) {
  const int S = m.S;
  // per slot j: the clusters containing j (j not a jump endpoint), by list type
  std::vector<std::vector<uint32_t>> per[3];
  for (auto &v : per) v.assign(kEnvSlots, {});
  // state entry: reduced weight-block byte offset [0:13] (rank layout, positive rows) | j is the
  // second other site [13] | the other site o [14:20] (63 = none) | neighbour-endpoint side [31]
  auto add_state = [&](const std::vector<StateCluster> &cl, int endpoint, bool nside) {
    for (const auto &c : cl) {
      int others[2] = {-1, -1}, n = 0;
      for (int i = 0; i < c.size; ++i)
        if (c.slots[i] != endpoint) others[n++] = c.slots[i];
      const int block = ReducedBlock(c.label, m.ne);
      for (int i = 0; i < n; ++i) {
        const int other = n == 2 ? others[1 - i] : 63;
        per[0][others[i]].push_back(static_cast<uint32_t>(block * 8) | (static_cast<uint32_t>(i) << 13) |
                                    (static_cast<uint32_t>(other) << 14) | (nside ? 0x80000000u : 0u));
      }
    }
  };
  add_state(jt.state_v_only, jt.v_slot, false);
  add_state(jt.state_n_only, jt.n_slot, true);
  // orbit entry: byte offset of the orbit's expanded weights [0:16] | j takes the rank*8 role
  // (second site, or a singlet's site) [16] | the other site o [17:23] (63 = none); backward
  // entries read canonical slot x at mirror[x]
  auto add_orbits = [&](const OrbitTable &ot, const std::vector<int> &eb, int list, bool backward) {
    for (const auto &c : ot.clusters) {
      const uint32_t bse = static_cast<uint32_t>(eb[c.orbit]) * 8;
      if (bse > 65535) throw std::invalid_argument("feature vector too long for the packed tables");
      const int nsl = c.s1 < 0 ? 1 : 2;
      for (int pos = 0; pos < nsl; ++pos) {
        const int canon = pos == 0 ? c.s0 : c.s1;
        const int canon_other = nsl == 1 ? -1 : (pos == 0 ? c.s1 : c.s0);
        const int slot = backward ? jt.mirror[canon] : canon;
        const int other = canon_other < 0 ? 63 : (backward ? jt.mirror[canon_other] : canon_other);
        const uint32_t role = (nsl == 1 || pos == 1) ? 1u : 0u;
        per[list][slot].push_back(bse | (role << 16) | (static_cast<uint32_t>(other) << 17));
      }
    }
  };
  if (m.kind == kE0 || m.kind == kQuartic) add_orbits(jt.mmm, eb1, 1, false);
  if (m.kind == kQuartic) {
    add_orbits(jt.mm2, eb2, 2, false);
    add_orbits(jt.mm2, eb2, 2, true);
  }
  off.assign(3 * (kEnvSlots + 1), 0);
  for (int t = 0; t < 3; ++t) {
    for (int j = 0; j < kEnvSlots; ++j) {
      off[t * (kEnvSlots + 1) + j] = static_cast<uint16_t>(inc.size());
      inc.insert(inc.end(), per[t][j].begin(), per[t][j].end());
    }
    off[t * (kEnvSlots + 1) + kEnvSlots] = static_cast<uint16_t>(inc.size());
  }
  if (inc.size() > 65535) throw std::invalid_argument("incidence table too large");

  const double nan = std::numeric_limits<double>::quiet_NaN();
  base.assign(static_cast<size_t>(S) * S * 3, nan);
  auto w_of = [&](int label, int a, int s) -> long double {
    if (label == 0) return m.w[a];
    if (label < 4) return m.w[S + (label - 1) * S * S + a * S + s];
    return m.w[S + 3 * S * S + (label - 4) * S * S * S + (a * S + s) * S + s];
  };
  auto tt = [&](int kind, int a, int b) { return m.type_tab[(kind * S + a) * S + b]; };
  for (int s = 0; s < S; ++s) {
    if (s == m.x_code) continue;
    for (int M = 0; M < S; ++M) {
      if (M == m.x_code) continue;
      long double de = 0, la = 0, lb = 0;
      for (const auto &c : jt.state_v_only) de += w_of(c.label, M, s) - w_of(c.label, m.x_code, s);
      for (const auto &c : jt.state_n_only) de -= w_of(c.label, M, s) - w_of(c.label, m.x_code, s);
      if (m.kind == kE0 || m.kind == kQuartic)
        for (const auto &c : jt.mmm.clusters) {
          const int kind = jt.mmm.orbit_kind[c.orbit];
          la += m.b_mmm[static_cast<size_t>(M) * m.F_mmm + m.base_mmm[c.orbit] + tt(kind, s, s)];
        }
      if (m.kind == kQuartic)
        for (const auto &c : jt.mm2.clusters) {
          const int kind = jt.mm2.orbit_kind[c.orbit];
          lb += 2.0L * m.b_mm2[static_cast<size_t>(M) * m.F_mm2 + m.base_mm2[c.orbit] +
                               tt(kind, s, s)];
        }
      double *o = &base[(static_cast<size_t>(s) * S + M) * 3];
      o[0] = static_cast<double>(de);
      o[1] = static_cast<double>(la);
      o[2] = static_cast<double>(lb);
    }
  }
}

// Models are shared by host threads (the reference's predictors are const objects called from
// OpenMP threads), so the lazy upload and every read of m.dev take this lock.
static std::mutex &UploadMutex() {
  static std::mutex mu;
  return mu;
}

bool IsUploaded(const Model &m) {
  std::lock_guard<std::mutex> lock(UploadMutex());
  return m.dev != nullptr;
}

const KTables *UploadedOn(const Model &m, int device) {
  std::lock_guard<std::mutex> lock(UploadMutex());
  return m.dev && m.dev->device == device ? &m.dev->k : nullptr;
}

const KTables &Upload(Model &m) {
  const int device = CurrentDevice();
  std::lock_guard<std::mutex> lock(UploadMutex());
  if (m.dev) {
    if (m.dev->device != device)
      throw std::invalid_argument("model was uploaded to another device");
    return m.dev->k;
  }
  const JumpTables &jt = GetJumpTables();
  const SiteTables &st = GetSiteTables();
  KTables k{};
  std::vector<int32_t> off;
  for (int d = 0; d < 12; ++d)
    for (int s = 0; s < kEnvSlots; ++s) off.push_back(PackOff(jt.off[d][s]));
  std::vector<uint8_t> mirror(jt.mirror.begin(), jt.mirror.end());
  // Jump dE weights by migrating species M and reduced index r over element ranks (see
  // ReducedBlock): wd[rank M][r] = w(cluster | endpoint M) - w(cluster | endpoint X); rows
  // ne .. 2 ne - 1 hold the same values negated (the neighbour endpoint's clusters). Each row
  // ends with one zero, the weight of the padding entries.
  const int S = m.S, ne = m.ne;
  const int nr = 1 + 3 * ne + 4 * ne * ne, wrow = nr + 1;
  std::vector<int> code;  // element rank -> species code
  for (int c = 0; c < S; ++c)
    if (c != m.x_code) code.push_back(c);
  auto wval = [&](int l, int a, int c1, int c2) {
    if (l == 0) return m.w[a];
    if (l < 4) return m.w[S + (l - 1) * S * S + a * S + c1];
    return m.w[S + 3 * S * S + (l - 4) * S * S * S + (a * S + c1) * S + c2];
  };
  std::vector<double> wd(static_cast<size_t>(2 * ne) * wrow, 0.0);
  for (int rm = 0; rm < ne; ++rm)
    for (int l = 0; l < 8; ++l) {
      const int n1 = l == 0 ? 1 : ne, n2 = l < 4 ? 1 : ne;
      for (int r1 = 0; r1 < n1; ++r1)
        for (int r2 = 0; r2 < n2; ++r2) {
          const size_t i = static_cast<size_t>(rm) * wrow + ReducedBlock(l, ne) + r1 + ne * r2;
          wd[i] = wval(l, code[rm], code[r1], code[r2]) - wval(l, m.x_code, code[r1], code[r2]);
          wd[i + static_cast<size_t>(ne) * wrow] = -wd[i];
        }
    }
  const int neg_bytes = ne * wrow * 8;
  std::vector<uint32_t> state;
  for (const auto &c : jt.state_v_only) state.push_back(PackState(c, jt.v_slot, false, ne, neg_bytes));
  for (const auto &c : jt.state_n_only) state.push_back(PackState(c, jt.n_slot, true, ne, neg_bytes));
  // expanded barrier rows: F1e / F2e weights, then one zero (the padding entries' weight)
  int F1e = 0, F2e = 0;
  const std::vector<int> eb1 = ExpandedBases(jt.mmm, ne, F1e), eb2 = ExpandedBases(jt.mm2, ne, F2e);
  auto mmm = PackOrbits(jt.mmm, eb1, nullptr);
  // MM2 clusters: forward entries, then the backward (mirrored) ones
  auto mm2 = PackOrbits(jt.mm2, eb2, nullptr);
  const std::vector<int> mir(jt.mirror.begin(), jt.mirror.end());
  const auto mm2b = PackOrbits(jt.mm2, eb2, &mir);
  mm2.insert(mm2.end(), mm2b.begin(), mm2b.end());
  // Lists padded to whole batches of the dense kernel (kClusterPad entries) with entries that
  // read the zero at the end of every weight row. Their sites are slot 63 (rank entries 0), in
  // each list's own byte roles: a state entry's byte 0 indexes the rank*8 bytes and byte 1 the
  // 16-bit rank*8*ne words (offset 2*63); an orbit entry the other way round.
  // The wide kernel shape (64 lanes per event, 4 entries each) reads copies padded to multiples
  // of kClusterPadWide.
  auto pad = [](std::vector<uint32_t> &v, int to, uint32_t roles, uint32_t zero_at) {
    if (zero_at > 65535) throw std::invalid_argument("weight table too large for the packed tables");
    while (v.size() % static_cast<size_t>(to)) v.push_back(roles | (zero_at << 16));
  };
  const uint32_t state_roles = 63u | (126u << 8), orbit_roles = 126u | (63u << 8);
  std::vector<uint32_t> state_w = state, mmm_w = mmm, mm2_w = mm2;
  pad(state, kClusterPad, state_roles, static_cast<uint32_t>(nr * 8));
  pad(mmm, kClusterPad, orbit_roles, static_cast<uint32_t>(F1e * 8));
  pad(mm2, kClusterPad, orbit_roles, static_cast<uint32_t>(F2e * 8));
  pad(state_w, kClusterPadWide, state_roles, static_cast<uint32_t>(nr * 8));
  pad(mmm_w, kClusterPadWide, orbit_roles, static_cast<uint32_t>(F1e * 8));
  pad(mm2_w, kClusterPadWide, orbit_roles, static_cast<uint32_t>(F2e * 8));
  // the 16-bit rank words are read at even byte offsets only
  for (const auto *v : {&state, &state_w})
    for (uint32_t e : *v)
      if ((e >> 8) & 1) throw std::logic_error("state entry with an odd rank-word offset");
  for (const auto *v : {&mmm, &mm2, &mmm_w, &mm2_w})
    for (uint32_t e : *v)
      if (e & 1) throw std::logic_error("orbit entry with an odd rank-word offset");
  std::vector<uint32_t> site_cl;
  for (const auto &c : st.clusters) site_cl.push_back(PackSite(c, 0));
  // species code -> element rank, 4 bits each (a vacancy maps to rank 0: it is flagged when
  // the environment is gathered, and every table index stays in range)
  uint32_t rank_lut = 0;
  for (int c = 0; c < S; ++c)
    if (c != m.x_code) rank_lut |= static_cast<uint32_t>(m.onehot[c]) << (4 * c);
  std::vector<int32_t> site_off;
  for (int s = 0; s < kSiteSlots; ++s) site_off.push_back(PackOff(st.off[s]));
  std::vector<double> par = m.par;
  for (int c = 0; c < m.S; ++c)
    if (!m.bar[c].present)
      for (int i = 0; i < 6; ++i) par[6 * c + i] = std::numeric_limits<double>::quiet_NaN();
  std::vector<int32_t> e1(m.eci.t1.begin(), m.eci.t1.end()), e2(m.eci.t2.begin(), m.eci.t2.end()),
      e3(m.eci.t3.begin(), m.eci.t3.end());
  for (auto *v : {&e1, &e2, &e3})
    for (auto &x : *v)
      if (x < 0) x = 0;  // invalid keys only occur with two vacancies in one cluster
  if (state.size() > kMaxStateClusters || mmm.size() > kMaxOrbitClusters ||
      site_cl.size() > kMaxSiteClusters || m.w.size() > kMaxW)
    throw std::invalid_argument("tables exceed the kernel's LDS capacity (too many species)");

  BlobBuilder b;
  b.Add(off, &k.off);
  b.Add(mirror, &k.mirror);
  b.Add(state, &k.state);
  b.Add(mmm, &k.mmm);
  b.Add(mm2, &k.mm2);
  b.Add(state_w, &k.state_w);
  b.Add(mmm_w, &k.mmm_w);
  b.Add(mm2_w, &k.mm2_w);
  b.Add(site_cl, &k.site_cl);
  b.Add(site_off, &k.site_off);
  b.Add(m.w, &k.w);
  b.Add(wd, &k.wd);
  const std::vector<double> b1c = ExpandWeights(m, jt.mmm, m.base_mmm, m.b_mmm, m.F_mmm, eb1, F1e + 1);
  const std::vector<double> b2c = ExpandWeights(m, jt.mm2, m.base_mm2, m.b_mm2, m.F_mm2, eb2, F2e + 1);
  b.Add(b1c, &k.b1);
  b.Add(b2c, &k.b2);
  b.Add(par, &k.par);
  b.Add(m.theta, &k.theta);
  b.Add(m.norm_total, &k.norm_tot);
  b.Add(e1, &k.eci1);
  b.Add(e2, &k.eci2);
  b.Add(e3, &k.eci3);
  std::vector<uint16_t> inc_off;
  std::vector<uint32_t> inc;
  std::vector<double> base;
  BuildSolventTables(m, jt, eb1, eb2, inc_off, inc, base);
  b.Add(inc_off, &k.inc_off);
  b.Add(inc, &k.inc);
  b.Add(base, &k.base);
  std::vector<int32_t> errflag(1, 0);
  const int32_t *err_ptr = nullptr;
  b.Add(errflag, &err_ptr);

  auto dev = std::make_unique<DeviceTables>();
  dev->device = device;
  Check(hipMalloc(&dev->blob, b.bytes.size()), "hipMalloc(model tables)");
  Check(hipMemcpy(dev->blob, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice),
        "hipMemcpy(model tables)");
  for (auto &f : b.fix) *f.second = static_cast<char *>(dev->blob) + f.first;
  k.vac_shell_mask = 0;
  for (int s : jt.vac_shell_slots) k.vac_shell_mask |= 1ull << s;
  k.n_state = static_cast<int>(state.size());
  k.n_mmm = static_cast<int>(mmm.size());
  k.n_mm2 = static_cast<int>(mm2.size());
  k.n_state_w = static_cast<int>(state_w.size());
  k.n_mmm_w = static_cast<int>(mmm_w.size());
  k.n_mm2_w = static_cast<int>(mm2_w.size());
  k.n_site = static_cast<int>(site_cl.size());
  k.n_eci = static_cast<int>(m.theta.size());
  k.v_slot = jt.v_slot;
  k.n_slot = jt.n_slot;
  k.kind = m.kind;
  k.S = m.S;
  k.ne = m.ne;
  k.x_code = m.x_code;
  k.F1 = F1e + 1;
  k.F2 = F2e + 1;
  k.nw = static_cast<int>(m.w.size());
  k.nwd = static_cast<int>(wd.size());
  k.wrow = wrow;
  k.rank_lut = rank_lut;
  k.n_state_v = static_cast<int>(jt.state_v_only.size());
  k.n_inc = static_cast<int>(inc.size());
  k.solvent_default = 0;
  for (int c = 0; c < m.S; ++c)
    if (m.eci.species[c] == "Al") k.solvent_default = c;
  if (k.solvent_default == m.x_code) k.solvent_default = m.x_code == 0 ? 1 : 0;
  dev->k = k;
  m.dev = dev.release();
  m.dev_err = const_cast<int32_t *>(err_ptr);
  return m.dev->k;
}

}  // namespace lmc

using lmc::Model;

namespace {

hipStream_t S(void *s) { return static_cast<hipStream_t>(s); }

void RequireModel(const lmc_model *m) {
  if (!m || !m->m) throw std::invalid_argument("null model");
}

// The buffers of a call over n items: a negative count, or a null buffer when n > 0, is an
// argument error (a null device pointer would otherwise reach a kernel).
void RequireBuffers(const char *fn, int64_t n, std::initializer_list<const void *> buffers) {
  if (n < 0) throw std::invalid_argument(std::string(fn) + ": negative count");
  if (n == 0) return;
  for (const void *b : buffers)
    if (!b) throw std::invalid_argument(std::string(fn) + ": null buffer");
}

template <typename T>
T *DevAlloc(size_t n) {
  void *p = nullptr;
  Check(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
  return static_cast<T *>(p);
}
template <typename T>
T *DevCopy(const std::vector<T> &v) {
  T *p = DevAlloc<T>(v.size());
  if (!v.empty())
    Check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
  return p;
}
template <typename T>
void DevFree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}
template <typename T>
void ToHost(T *dst, const T *src, size_t n) {
  if (dst && n) Check(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

using lmc::ElementMass;

struct Batch {
  std::vector<int64_t> off;
  int64_t total = 0;
};
Batch Layout(int32_t R, const int32_t *ncell) {
  if (R <= 0) throw std::invalid_argument("n_replicas must be positive");
  Batch b;
  for (int r = 0; r < R; ++r) {
    if (ncell[r] < 4 || ncell[r] > 1000)
      throw std::invalid_argument("lattice size n must be in [4, 1000]");
    b.off.push_back(b.total);
    b.total += 4ll * ncell[r] * ncell[r] * ncell[r];
  }
  return b;
}

void ComputeTotalEnergy(const lmc_model *em, const uint8_t *d_occ, const int64_t *d_off,
                        const int32_t *d_ncell, int32_t R, double *d_energy, hipStream_t s) {
  const auto &t = lmc::Upload(*em->m);
  unsigned long long *counts = DevAlloc<unsigned long long>(static_cast<size_t>(R) * t.n_eci);
  const hipError_t e = lmc::LaunchTotalEnergy(t, d_occ, d_off, d_ncell, R, counts, d_energy, s);
  (void)hipStreamSynchronize(s);
  (void)hipFree(counts);
  Check(e, "total energy kernel");
}

// Host copies of a batch's offsets and the total site count.
std::vector<int64_t> HostOffsets(int32_t R, const int64_t *d_off, const int32_t *d_ncell,
                                 int64_t &total) {
  std::vector<int64_t> off(R);
  std::vector<int32_t> nc(R);
  ToHost(off.data(), d_off, off.size());
  ToHost(nc.data(), d_ncell, nc.size());
  total = off.back() + 4ll * nc.back() * nc.back() * nc.back();
  off.push_back(total);
  return off;
}

// Device copies of the atom on every site (ids local to each replica, int32) and map_shift by
// atom; a null lattice_to_atom turns tracking off.
void UploadAtoms(int32_t R, const int64_t *d_off, const int32_t *d_ncell, const int64_t *l2a,
                 const int32_t *map_shift, int32_t *&d_atom, int32_t *&d_shift) {
  if (!l2a) {
    DevFree(d_atom);
    DevFree(d_shift);
    return;
  }
  int64_t total = 0;
  const std::vector<int64_t> off = HostOffsets(R, d_off, d_ncell, total);
  std::vector<int32_t> a(static_cast<size_t>(total));
  std::vector<char> seen;
  for (int r = 0; r < R; ++r) {
    const int64_t N = off[r + 1] - off[r];
    if (N > std::numeric_limits<int32_t>::max())
      throw std::invalid_argument("atom tracking needs fewer than 2^31 sites per replica");
    seen.assign(static_cast<size_t>(N), 0);
    for (int64_t s = 0; s < N; ++s) {
      const int64_t v = l2a[off[r] + s];
      if (v < 0 || v >= N || seen[static_cast<size_t>(v)])
        throw std::invalid_argument("lattice_to_atom of replica " + std::to_string(r) +
                                    " is not a permutation of its atom ids");
      seen[static_cast<size_t>(v)] = 1;
      a[static_cast<size_t>(off[r] + s)] = static_cast<int32_t>(v);
    }
  }
  std::vector<int32_t> sh = map_shift ? std::vector<int32_t>(map_shift, map_shift + 3 * total)
                                      : std::vector<int32_t>(static_cast<size_t>(3 * total), 0);
  DevFree(d_atom);
  DevFree(d_shift);
  d_atom = DevCopy(a);
  d_shift = DevCopy(sh);
}

void DownloadAtoms(int32_t R, const int64_t *d_off, const int32_t *d_ncell, const int32_t *d_atom,
                   const int32_t *d_shift, int64_t *l2a, int32_t *map_shift) {
  if (!d_atom) throw std::invalid_argument("atom tracking is off");
  Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  int64_t total = 0;
  (void)HostOffsets(R, d_off, d_ncell, total);
  if (l2a) {
    std::vector<int32_t> a(static_cast<size_t>(total));
    ToHost(a.data(), d_atom, a.size());
    std::copy(a.begin(), a.end(), l2a);
  }
  if (map_shift) ToHost(map_shift, d_shift, static_cast<size_t>(3 * total));
}

}  // namespace

struct lmc_kmc {
  const lmc_model *bm = nullptr;
  lmc::KmcDev d;
  uint64_t trace_pos = 0;
  // kmc_log.txt writing: per replica the starting step, the solute column switch, and the
  // stream state (file begun, std::fixed set by the first vector written)
  std::vector<uint64_t> start_steps;
  std::vector<char> solute_disp, log_begun, log_fixed, log_skip_first;
  int64_t total_sites = 0;
  int formulation = 0;
  // surrogate evaluation (optional)
  const lmc_surrogate *sur = nullptr;
  double sur_threshold = 0.01, sur_window = 20.0;
  uint64_t *n_sur = nullptr, *n_fb = nullptr;
  // captured launch sequence (lmc_kmc_run_graph)
  hipGraphExec_t graph = nullptr;
  hipStream_t capture = nullptr;
  uint64_t graph_steps = 0;
  int32_t graph_launches = 0, graph_form = -1;
  const void *graph_sur = nullptr;
  void DropGraph() {
    if (graph) (void)hipGraphExecDestroy(graph);
    graph = nullptr;
  }
  ~lmc_kmc() {
    DropGraph();
    if (capture) (void)hipStreamDestroy(capture);
    DevFree(n_sur);
    DevFree(n_fb);
    DevFree(d.occ);
    DevFree(d.off);
    DevFree(d.ncell);
    DevFree(d.vac);
    DevFree(d.steps);
    DevFree(d.max_steps);
    DevFree(d.seed);
    DevFree(d.rid);
    DevFree(d.time);
    DevFree(d.energy);
    DevFree(d.abs_energy);
    DevFree(d.temp);
    DevFree(d.cv);
    DevFree(d.cs);
    DevFree(d.solute_mass);
    DevFree(d.flags);
    DevFree(d.solvent);
    DevFree(d.escaped);
    DevFree(d.esc_time);
    DevFree(d.esc_step);
    DevFree(d.esc_site);
    DevFree(d.disp);
    DevFree(d.com);
    DevFree(d.error);
    DevFree(d.prev_j);
    DevFree(d.atom);
    DevFree(d.shift);
    DevFree(d.log_n);
    DevFree(d.log);
    DevFree(d.log_edge);
    DevFree(d.tt_time);
    DevFree(d.tt_temp);
    DevFree(d.tr_sel);
    DevFree(d.tr_ea);
    DevFree(d.tr_de);
    DevFree(d.tr_time);
  }
};

struct lmc_cmc {
  const lmc_model *pm = nullptr;
  lmc::CmcDev d;
  uint64_t trace_pos = 0;
  std::vector<uint64_t> start_steps;         // restart_steps of each replica
  std::vector<char> log_begun, log_skip_first;
  void FreeLog() {
    DevFree(d.log_n);
    DevFree(d.log_lo);
    DevFree(d.log);
    DevFree(d.ring);
    DevFree(d.ring_n);
    d.log_every = d.log_cap = d.avg_m = d.ring_cap = 0;
  }
  ~lmc_cmc() {
    FreeLog();
    DevFree(d.omp_p);
    DevFree(d.omp_q);
    DevFree(d.omp_de);
    DevFree(d.omp_len);
    DevFree(d.omp_pos);
    DevFree(d.anneal);
    DevFree(d.atom);
    DevFree(d.shift);
    DevFree(d.occ);
    DevFree(d.off);
    DevFree(d.ncell);
    DevFree(d.steps);
    DevFree(d.max_steps);
    DevFree(d.seed);
    DevFree(d.accepted);
    DevFree(d.rid);
    DevFree(d.t0);
    DevFree(d.t1);
    DevFree(d.energy);
    DevFree(d.abs_energy);
    DevFree(d.error);
    DevFree(d.tr_p);
    DevFree(d.tr_q);
    DevFree(d.tr_de);
    DevFree(d.tr_acc);
  }
};

extern "C" {

const char *lmc_last_error(void) { return lmc::g_last_error.c_str(); }
int lmc_version(void) { return 1; }

int lmc_device_alloc(uint64_t bytes, void **d_ptr) {
  return Guard([&] {
    if (!d_ptr) throw std::invalid_argument("lmc_device_alloc: null argument");
    *d_ptr = nullptr;
    if (bytes) Check(hipMalloc(d_ptr, bytes), "hipMalloc");
    return LMC_OK;
  });
}
int lmc_device_free(void *d_ptr) {
  return Guard([&] {
    if (d_ptr) Check(hipFree(d_ptr), "hipFree");
    return LMC_OK;
  });
}
int lmc_copy_to_device(void *d_dst, const void *src, uint64_t bytes) {
  return Guard([&] {
    if (bytes && (!d_dst || !src)) throw std::invalid_argument("lmc_copy_to_device: null argument");
    if (bytes) Check(hipMemcpy(d_dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    return LMC_OK;
  });
}
int lmc_copy_to_host(void *dst, const void *d_src, uint64_t bytes) {
  return Guard([&] {
    if (bytes && (!dst || !d_src)) throw std::invalid_argument("lmc_copy_to_host: null argument");
    if (bytes) Check(hipMemcpy(dst, d_src, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    return LMC_OK;
  });
}
int lmc_synchronize(void *stream) {
  return Guard([&] {
    Check(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
    return LMC_OK;
  });
}

int lmc_model_load(const char *json_path, const char *const *elements, int32_t n_elements,
                   int32_t kind, lmc_model **out) {
  return Guard([&] {
    if (!json_path || !elements || n_elements <= 0 || !out)
      throw std::invalid_argument("lmc_model_load: null argument");
    if (kind < LMC_MODEL_ENERGY || kind > LMC_MODEL_PAIRSITE)
      throw std::invalid_argument("lmc_model_load: unknown kind");
    std::vector<std::string> el;
    for (int i = 0; i < n_elements; ++i) el.emplace_back(elements[i] ? elements[i] : "");
    std::unique_ptr<Model> m(lmc::LoadModel(json_path, el, kind));
    *out = new lmc_model{m.release()};
    return LMC_OK;
  });
}
void lmc_model_free(lmc_model *m) {
  if (m) {
    delete m->m;
    delete m;
  }
}
int32_t lmc_model_num_species(const lmc_model *m) { return m && m->m ? m->m->S : -1; }
int lmc_model_set_formulation(lmc_model *m, int32_t formulation) {
  if (!m || !m->m || formulation < 0 || formulation > 1) return Fail(LMC_ERR_ARG, "bad formulation");
  m->m->formulation = formulation;
  return LMC_OK;
}
int32_t lmc_model_species_code(const lmc_model *m, const char *element) {
  if (!m || !m->m || !element) return -1;
  for (int i = 0; i < m->m->S; ++i)
    if (m->m->eci.species[i] == element) return i;
  return -1;
}
const char *lmc_model_species_name(const lmc_model *m, int32_t code) {
  if (!m || !m->m || code < 0 || code >= m->m->S) return nullptr;
  return m->m->eci.species[code].c_str();
}
int lmc_model_layout(const lmc_model *m, int32_t *n_eci, int32_t *n_mmm, int32_t *n_mm2) {
  return Guard([&] {
    RequireModel(m);
    if (n_eci) *n_eci = static_cast<int32_t>(m->m->eci.keys.size());
    if (n_mmm) *n_mmm = m->m->F_mmm;
    if (n_mm2) *n_mm2 = m->m->F_mm2;
    return LMC_OK;
  });
}
int lmc_model_eci_key(const lmc_model *m, int32_t index, int32_t *label, int32_t *codes3) {
  return Guard([&] {
    RequireModel(m);
    const auto &keys = m->m->eci.keys;
    if (index < 0 || index >= static_cast<int32_t>(keys.size()))
      throw std::invalid_argument("ECI index out of range");
    *label = keys[index].label;
    for (int i = 0; i < 3; ++i)
      codes3[i] = i < static_cast<int>(keys[index].codes.size()) ? keys[index].codes[i] : -1;
    return LMC_OK;
  });
}
int lmc_model_orbits(const lmc_model *m, int32_t which, int32_t cap, int32_t *n_orbits,
                     int32_t *size, int32_t *kind, int32_t *shell, int32_t *base, int32_t *len) {
  return Guard([&] {
    RequireModel(m);
    const auto &jt = lmc::GetJumpTables();
    const auto &ot = which == 0 ? jt.mmm : jt.mm2;
    const auto &bs = which == 0 ? m->m->base_mmm : m->m->base_mm2;
    const auto &ln = which == 0 ? m->m->len_mmm : m->m->len_mm2;
    const int no = static_cast<int>(ot.orbit_size.size());
    if (no > cap) throw std::invalid_argument("orbit array capacity too small");
    *n_orbits = no;
    for (int o = 0; o < no; ++o) {
      size[o] = ot.orbit_size[o];
      kind[o] = ot.orbit_kind[o];
      shell[o] = ot.orbit_shell[o];
      base[o] = bs[o];
      len[o] = ln[o];
    }
    return LMC_OK;
  });
}

void lmc_philox4x32(uint64_t seed, uint64_t counter, uint32_t replica, uint32_t tag,
                    uint32_t *out4) {
  const lmc::U4 u = lmc::Philox4x32_10(seed, counter, replica, tag);
  out4[0] = u.x;
  out4[1] = u.y;
  out4[2] = u.z;
  out4[3] = u.w;
}

int64_t lmc_fcc_num_sites(int32_t n) { return n >= 4 ? 4ll * n * n * n : -1; }

int32_t lmc_fcc_neighbors(int32_t n, int64_t site, int32_t shell, int64_t *out) {
  if (n < 4 || site < 0 || site >= 4ll * n * n * n || shell < 1 || shell > 3 || !out)
    return Fail(LMC_ERR_ARG, "lmc_fcc_neighbors: bad argument");
  const lmc::I3 *tab = shell == 1 ? lmc::kNN1 : shell == 2 ? lmc::kNN2 : lmc::kNN3;
  const int cnt = shell == 1 ? 12 : shell == 2 ? 6 : 24;
  for (int i = 0; i < cnt; ++i) out[i] = lmc::SiteAdd(n, site, tab[i]);
  std::sort(out, out + cnt);
  return cnt;
}

static int FindDirection(int32_t n, int64_t v, int64_t nb) {
  for (int k = 0; k < 12; ++k)
    if (lmc::SiteAdd(n, v, lmc::kNN1[k]) == nb) return k;
  return -1;
}

int32_t lmc_jump_state_clusters(int32_t n, int64_t v, int64_t nb, int64_t *out, int32_t cap) {
  if (n < 4) return Fail(LMC_ERR_ARG, "n must be >= 4");
  const int k = FindDirection(n, v, nb);
  if (k < 0) return Fail(LMC_ERR_ARG, "sites are not first neighbours");
  const auto &jt = lmc::GetJumpTables();
  int32_t cnt = 0;
  for (const auto &c : jt.state) {
    if (cnt >= cap) return Fail(LMC_ERR_ARG, "capacity too small");
    int64_t *r = out + 4 * cnt;
    r[0] = c.label;
    for (int i = 0; i < 3; ++i)
      r[1 + i] = i < c.size ? lmc::SiteAdd(n, v, jt.off[k][c.slots[i]]) : -1;
    ++cnt;
  }
  return cnt;
}

int32_t lmc_jump_environment(int32_t n, int64_t v, int64_t nb, int64_t *out60) {
  if (n < 4) return Fail(LMC_ERR_ARG, "n must be >= 4");
  if (!out60) return Fail(LMC_ERR_ARG, "null output");
  const int k = FindDirection(n, v, nb);
  if (k < 0) return Fail(LMC_ERR_ARG, "sites are not first neighbours");
  const auto &jt = lmc::GetJumpTables();
  for (int j = 0; j < lmc::kEnvSlots; ++j) out60[j] = lmc::SiteAdd(n, v, jt.off[k][j]);
  return lmc::kEnvSlots;
}

int lmc_jump_symmetries(int32_t *perm) {
  return Guard([&] {
    if (!perm) throw std::invalid_argument("null output");
    const auto &jt = lmc::GetJumpTables();
    const int sy[4] = {1, 1, -1, -1}, sz[4] = {1, -1, 1, -1};
    for (int g = 0; g < 4; ++g)
      for (int j = 0; j < lmc::kEnvSlots; ++j) {
        const lmc::I3 want{jt.key[j].x, sy[g] * jt.key[j].y, sz[g] * jt.key[j].z};
        int found = -1;
        for (int u = 0; u < lmc::kEnvSlots; ++u)
          if (jt.key[u].x == want.x && jt.key[u].y == want.y && jt.key[u].z == want.z) found = u;
        if (found < 0) throw std::logic_error("environment is not symmetric");
        perm[g * lmc::kEnvSlots + j] = found;
      }
    return LMC_OK;
  });
}

int32_t lmc_jump_orbit_clusters(int32_t n, int64_t v, int64_t nb, int32_t which, int64_t *out,
                                int32_t cap) {
  if (n < 4) return Fail(LMC_ERR_ARG, "n must be >= 4");
  const int k = FindDirection(n, v, nb);
  if (k < 0) return Fail(LMC_ERR_ARG, "sites are not first neighbours");
  const auto &jt = lmc::GetJumpTables();
  const auto &ot = which == 0 ? jt.mmm : jt.mm2;
  int32_t cnt = 0;
  for (const auto &c : ot.clusters) {
    if (cnt >= cap) return Fail(LMC_ERR_ARG, "capacity too small");
    int64_t *r = out + 4 * cnt;
    r[0] = c.orbit;
    r[1] = c.sym ? 1 : 0;
    r[2] = lmc::SiteAdd(n, v, jt.off[k][c.s0]);
    r[3] = c.s1 < 0 ? -1 : lmc::SiteAdd(n, v, jt.off[k][c.s1]);
    ++cnt;
  }
  return cnt;
}

int lmc_eval_jumps(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                   const int32_t *d_ncell, const int64_t *d_vac, int32_t n_replicas,
                   int64_t *d_nb, double *d_ea, double *d_de, void *stream) {
  return Guard([&] {
    RequireModel(m);
    if (m->m->kind != LMC_MODEL_E0 && m->m->kind != LMC_MODEL_QUARTIC)
      throw std::invalid_argument("lmc_eval_jumps needs an E0 or Quartic model");
    RequireBuffers("lmc_eval_jumps", n_replicas,
                   {d_occ, d_offset, d_ncell, d_vac, d_nb, d_ea, d_de});
    if (n_replicas == 0) return LMC_OK;
    const auto &t = lmc::Upload(*m->m);
    Check(lmc::LaunchEvalJumps(t, d_occ, d_offset, d_ncell, d_vac, n_replicas, d_nb, d_ea, d_de,
                               m->m->dev_err, S(stream), m->m->formulation == 1),
          "eval jumps kernel");
    return LMC_OK;
  });
}

int lmc_vacancy_scan(const lmc_model *m, const uint8_t *d_occ, int32_t ncell, int64_t vacancy,
                     int32_t solvent, const int64_t *d_sites, int32_t n_sites, int64_t *d_nb,
                     double *d_ea, double *d_de, void *stream) {
  return Guard([&] {
    RequireModel(m);
    if (m->m->kind != LMC_MODEL_E0 && m->m->kind != LMC_MODEL_QUARTIC)
      throw std::invalid_argument("lmc_vacancy_scan needs an E0 or Quartic model");
    if (n_sites < 0) throw std::invalid_argument("n_sites < 0");
    if (ncell < 4) throw std::invalid_argument("ncell must be >= 4");
    const int64_t N = 4ll * ncell * ncell * ncell;
    if (vacancy < -1 || vacancy >= N) throw std::invalid_argument("vacancy site out of range");
    if (solvent < 0 || solvent >= m->m->S || solvent == m->m->x_code)
      throw std::invalid_argument("solvent must be an element code");
    if (n_sites > 0 && (!d_occ || !d_sites || !d_nb || !d_ea || !d_de))
      throw std::invalid_argument("null buffer");
    const auto &t = lmc::Upload(*m->m);
    Check(lmc::LaunchVacancyScan(t, d_occ, ncell, vacancy, solvent, d_sites, n_sites, d_nb, d_ea,
                                 d_de, m->m->dev_err, S(stream)),
          "vacancy scan kernel");
    return LMC_OK;
  });
}

int lmc_check_errors(const lmc_model *m) {
  return Guard([&] {
    RequireModel(m);
    if (!lmc::IsUploaded(*m->m)) return LMC_OK;
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    int32_t e = 0;
    Check(hipMemcpy(&e, m->m->dev_err, sizeof(e), hipMemcpyDeviceToHost), "hipMemcpy");
    if (e) {
      const int32_t z = 0;
      Check(hipMemcpy(m->m->dev_err, &z, sizeof(z), hipMemcpyHostToDevice), "hipMemcpy");
      throw ModelError("Cluster not found in ClusterIndexer or element without barrier parameters");
    }
    return LMC_OK;
  });
}

int lmc_eval_pair(const lmc_model *m, const uint8_t *d_occ, int32_t ncell, int64_t v, int64_t nb,
                  double *ea, double *de) {
  return Guard([&] {
    RequireModel(m);
    RequireBuffers("lmc_eval_pair", 1, {d_occ, ea, de});
    if (ncell < 4) throw std::invalid_argument("ncell must be >= 4");
    if (v < 0 || v >= 4ll * ncell * ncell * ncell) throw std::out_of_range("vacancy site out of range");
    const int k = FindDirection(ncell, v, nb);
    if (k < 0) throw std::out_of_range("Neighbor not found for lattice pair in GetPairFlatIndex");
    int64_t *buf = DevAlloc<int64_t>(2 + 12);
    double *fb = DevAlloc<double>(24);
    int32_t *nc = DevAlloc<int32_t>(1);
    const int64_t h[2] = {0, v};
    int rc = LMC_OK;
    try {
      Check(hipMemcpy(buf, h, sizeof(h), hipMemcpyHostToDevice), "hipMemcpy");
      Check(hipMemcpy(nc, &ncell, sizeof(ncell), hipMemcpyHostToDevice), "hipMemcpy");
      const auto &t = lmc::Upload(*m->m);
      Check(lmc::LaunchEvalJumps(t, d_occ, buf, nc, buf + 1, 1, buf + 2, fb, fb + 12,
                                 m->m->dev_err, nullptr, m->m->formulation == 1),
            "eval jumps kernel");
      int64_t nbs[12];
      double r[24];
      Check(hipMemcpy(nbs, buf + 2, sizeof(nbs), hipMemcpyDeviceToHost), "hipMemcpy");
      Check(hipMemcpy(r, fb, sizeof(r), hipMemcpyDeviceToHost), "hipMemcpy");
      for (int i = 0; i < 12; ++i)
        if (nbs[i] == nb) {
          *ea = r[i];
          *de = r[12 + i];
        }
    } catch (...) {
      (void)hipFree(buf);
      (void)hipFree(fb);
      (void)hipFree(nc);
      throw;
    }
    (void)hipFree(buf);
    (void)hipFree(fb);
    (void)hipFree(nc);
    rc = lmc_check_errors(m);
    return rc;
  });
}

int lmc_swap_de(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                const int32_t *d_ncell, const int64_t *d_pairs, int32_t n_replicas, double *d_de,
                void *stream) {
  return Guard([&] {
    RequireModel(m);
    RequireBuffers("lmc_swap_de", n_replicas, {d_occ, d_offset, d_ncell, d_pairs, d_de});
    if (n_replicas == 0) return LMC_OK;
    const auto &t = lmc::Upload(*m->m);
    Check(lmc::LaunchSwapDe(t, d_occ, d_offset, d_ncell, d_pairs, n_replicas, d_de, S(stream)),
          "swap dE kernel");
    return LMC_OK;
  });
}

int lmc_site_de(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                const int32_t *d_ncell, const int64_t *d_sites, const uint8_t *d_codes,
                int32_t n_replicas, double *d_de, void *stream) {
  return Guard([&] {
    RequireModel(m);
    RequireBuffers("lmc_site_de", n_replicas, {d_occ, d_offset, d_ncell, d_sites, d_codes, d_de});
    if (n_replicas == 0) return LMC_OK;
    const auto &t = lmc::Upload(*m->m);
    Check(lmc::LaunchSiteDe(t, d_occ, d_offset, d_ncell, d_sites, nullptr, d_codes, n_replicas,
                            d_de, S(stream)),
          "site dE kernel");
    return LMC_OK;
  });
}

int lmc_site_change_de(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                       const int32_t *d_ncell, const int64_t *d_sites, const uint8_t *d_from,
                       const uint8_t *d_codes, int32_t n_replicas, double *d_de, void *stream) {
  return Guard([&] {
    RequireModel(m);
    RequireBuffers("lmc_site_change_de", n_replicas,
                   {d_occ, d_offset, d_ncell, d_sites, d_from, d_codes, d_de});
    if (n_replicas == 0) return LMC_OK;
    const auto &t = lmc::Upload(*m->m);
    Check(lmc::LaunchSiteDe(t, d_occ, d_offset, d_ncell, d_sites, d_from, d_codes, n_replicas,
                            d_de, S(stream)),
          "site dE kernel");
    return LMC_OK;
  });
}

int lmc_total_energy(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                     const int32_t *d_ncell, int32_t n_replicas, double *d_energy, void *stream) {
  return Guard([&] {
    RequireModel(m);
    RequireBuffers("lmc_total_energy", n_replicas, {d_occ, d_offset, d_ncell, d_energy});
    if (n_replicas == 0) return LMC_OK;
    ComputeTotalEnergy(m, d_occ, d_offset, d_ncell, n_replicas, d_energy, S(stream));
    return LMC_OK;
  });
}

int lmc_total_energy_noalloc(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                             const int32_t *d_ncell, int32_t n_replicas, double *d_energy,
                             void *stream) {
  return Guard([&] {
    RequireModel(m);
    if (n_replicas < 0 || (n_replicas && (!d_occ || !d_offset || !d_ncell || !d_energy)))
      throw std::invalid_argument("lmc_total_energy_noalloc: bad arguments");
    // the model's tables must already be on this device: uploading them would allocate
    const lmc::KTables *tp = lmc::UploadedOn(*m->m, CurrentDevice());
    if (!tp)
      throw std::invalid_argument(
          "lmc_total_energy_noalloc: the model's tables are not on this device; run any compute "
          "entry point with this model on it first");
    const lmc::KTables &t = *tp;
    const size_t need = static_cast<size_t>(n_replicas) * static_cast<size_t>(t.n_eci);
    const size_t cap = lmc::TotalEnergyScratchCapacity();
    if (need > cap)
      throw std::invalid_argument("lmc_total_energy_noalloc: " + std::to_string(n_replicas) +
                                  " replicas need " + std::to_string(need) +
                                  " cluster counts, more than the scratch's " +
                                  std::to_string(cap) + " (at most " +
                                  std::to_string(cap / static_cast<size_t>(t.n_eci)) +
                                  " replicas for this model)");
    if (n_replicas == 0) return LMC_OK;
    unsigned long long *counts = nullptr;
    Check(lmc::TotalEnergyScratch(&counts), "scratch address");
    Check(lmc::LaunchTotalEnergy(t, d_occ, d_offset, d_ncell, n_replicas, counts, d_energy,
                                 S(stream)),
          "total energy kernel");
    return LMC_OK;
  });
}

int lmc_total_encode(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                     const int32_t *d_ncell, int32_t n_replicas, double *d_encode, void *stream) {
  return Guard([&] {
    RequireModel(m);
    if (n_replicas <= 0) throw std::invalid_argument("lmc_total_encode: n_replicas must be > 0");
    RequireBuffers("lmc_total_encode", n_replicas, {d_occ, d_offset, d_ncell, d_encode});
    const auto &t = lmc::Upload(*m->m);
    unsigned long long *counts =
        DevAlloc<unsigned long long>(static_cast<size_t>(n_replicas) * t.n_eci);
    const hipError_t e = lmc::LaunchTotalEncode(t, d_occ, d_offset, d_ncell, n_replicas, counts,
                                                d_encode, nullptr, S(stream));
    (void)hipStreamSynchronize(S(stream));
    (void)hipFree(counts);
    Check(e, "encode kernel");
    return LMC_OK;
  });
}

int lmc_cluster_energy(const lmc_model *m, const uint8_t *d_occ, int32_t n, const int64_t *sites,
                       int32_t k, double *encode, double *energy) {
  return Guard([&] {
    RequireModel(m);
    if (!d_occ || !sites || k <= 0 || n < 4 || n > 1000) throw std::invalid_argument("bad argument");
    const int64_t N = 4ll * n * n * n;
    std::vector<int64_t> set;
    for (int32_t i = 0; i < k; ++i) {
      if (sites[i] < 0 || sites[i] >= N) throw std::invalid_argument("site out of range");
      set.push_back(sites[i]);
      for (int shell = 1; shell <= 3; ++shell) {
        int64_t nb[24];
        const int c = lmc_fcc_neighbors(n, sites[i], shell, nb);
        set.insert(set.end(), nb, nb + c);
      }
    }
    std::sort(set.begin(), set.end());
    set.erase(std::unique(set.begin(), set.end()), set.end());
    if (set.size() > static_cast<size_t>(std::numeric_limits<int32_t>::max()))
      throw std::invalid_argument("too many sites");
    const auto &t = lmc::Upload(*m->m);
    int64_t *d_set = DevCopy(set);
    unsigned long long *counts = DevAlloc<unsigned long long>(static_cast<size_t>(t.n_eci));
    double *d_enc = DevAlloc<double>(static_cast<size_t>(t.n_eci));
    double *d_e = DevAlloc<double>(1);
    hipError_t e = lmc::LaunchClusterEncode(t, d_occ, n, d_set, static_cast<int32_t>(set.size()),
                                            counts, d_enc, d_e, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess && encode)
      e = hipMemcpy(encode, d_enc, sizeof(double) * t.n_eci, hipMemcpyDeviceToHost);
    if (e == hipSuccess && energy) e = hipMemcpy(energy, d_e, sizeof(double), hipMemcpyDeviceToHost);
    DevFree(d_set);
    DevFree(counts);
    DevFree(d_enc);
    DevFree(d_e);
    Check(e, "cluster encode kernel");
    return LMC_OK;
  });
}

int lmc_chemical_potential(const lmc_model *m, const char *solvent, double *mu) {
  return Guard([&] {
    RequireModel(m);
    if (!solvent || !mu) throw std::invalid_argument("null argument");
    const Model &md = *m->m;
    int sv = -1;
    for (int c = 0; c < md.S; ++c)
      if (md.eci.species[c] == lmc::CanonicalElement(solvent)) sv = c;
    if (sv < 0) throw std::invalid_argument(std::string("solvent ") + solvent + " is not in the model");
    // GenerateFCC({15, 15, 15}, solvent), and per other species the same box with atom 0 (lattice
    // site 0) replaced (EnergyPredictor.cpp:194-212): one batch of S replicas on the GPU.
    const int n = 15, S_ = md.S;
    const int64_t N = 4ll * n * n * n;
    std::vector<uint8_t> occ(static_cast<size_t>(N * S_), static_cast<uint8_t>(sv));
    std::vector<int64_t> off(S_);
    for (int c = 0; c < S_; ++c) {
      off[c] = c * N;
      occ[static_cast<size_t>(c * N)] = static_cast<uint8_t>(c);
    }
    uint8_t *d_occ = DevCopy(occ);
    int64_t *d_off = DevCopy(off);
    int32_t *d_nc = DevCopy(std::vector<int32_t>(S_, n));
    double *d_e = DevAlloc<double>(S_);
    std::vector<double> e(S_);
    try {
      ComputeTotalEnergy(m, d_occ, d_off, d_nc, S_, d_e, nullptr);
      ToHost(e.data(), d_e, e.size());
    } catch (...) {
      DevFree(d_occ);
      DevFree(d_off);
      DevFree(d_nc);
      DevFree(d_e);
      throw;
    }
    DevFree(d_occ);
    DevFree(d_off);
    DevFree(d_nc);
    DevFree(d_e);
    for (int c = 0; c < S_; ++c) mu[c] = c == sv ? 0.0 : e[c] - e[sv];
    return LMC_OK;
  });
}

int lmc_kmc_create(const lmc_model *barrier, const lmc_model *energy, int32_t R,
                   const int32_t *ncell, const uint8_t *occ, const lmc_kmc_replica *params,
                   lmc_kmc **out) {
  return Guard([&] {
    RequireModel(barrier);
    if (barrier->m->kind != LMC_MODEL_E0 && barrier->m->kind != LMC_MODEL_QUARTIC)
      throw std::invalid_argument("KMC needs an E0 or Quartic barrier model");
    if (!ncell || !occ || !params || !out) throw std::invalid_argument("null argument");
    const Batch b = Layout(R, ncell);
    const Model &m = *barrier->m;
    auto k = std::make_unique<lmc_kmc>();
    k->bm = barrier;
    k->d.R = R;
    std::vector<int64_t> vac(R);
    std::vector<double> cv(R), cs(R), smass(R), temp(R), time(R), en(R);
    std::vector<int32_t> flags(R), solvent(R);
    std::vector<uint64_t> steps(R), maxs(R), seed(R);
    std::vector<uint32_t> rid(R);
    int al = -1;
    for (int c = 0; c < m.S; ++c)
      if (m.eci.species[c] == "Al") al = c;
    for (int r = 0; r < R; ++r) {
      const uint8_t *o = occ + b.off[r];
      const int64_t N = 4ll * ncell[r] * ncell[r] * ncell[r];
      std::vector<int64_t> cnt(m.S, 0);
      int64_t nv = 0;
      for (int64_t i = 0; i < N; ++i) {
        if (o[i] >= m.S) throw std::invalid_argument("occupancy holds a species code >= S");
        cnt[o[i]]++;
        if (o[i] == m.x_code) {
          vac[r] = i;
          ++nv;
        }
      }
      if (nv != 1) throw std::invalid_argument("each replica needs exactly one vacancy");
      const lmc_kmc_replica &p = params[r];
      if (!(p.temperature > 0)) throw std::invalid_argument("temperature must be positive");
      cv[r] = static_cast<double>(nv) / static_cast<double>(N);
      int64_t ns = 0;
      for (int c = 0; c < m.S; ++c)
        if (c != al && c != m.x_code) ns += cnt[c];
      cs[r] = static_cast<double>(ns) / static_cast<double>(N);
      int64_t mx = 0;  // Config::GetSolventElement: first species with the largest count
      for (int c = 0; c < m.S; ++c)
        if (cnt[c] > mx) {
          mx = cnt[c];
          solvent[r] = c;
        }
      double sm = 0;
      for (int c = 0; c < m.S; ++c)
        if (c != solvent[r]) sm += ElementMass(m.eci.species[c]) * static_cast<double>(cnt[c]);
      smass[r] = sm;
      temp[r] = p.temperature;
      time[r] = p.restart_time;
      en[r] = p.restart_energy;
      steps[r] = p.restart_steps;
      maxs[r] = p.maximum_steps;
      seed[r] = p.seed;
      rid[r] = p.replica_id;
      flags[r] = (p.rate_corrector ? 1 : 0) | (p.early_stop ? 2 : 0) | (p.solute_disp ? 4 : 0);
    }
    for (int c = 0; c < m.S && c < lmc::kMaxSpecies; ++c)
      k->d.mass[c] = ElementMass(m.eci.species[c]);
    k->start_steps = steps;
    k->solute_disp.resize(R);
    for (int r = 0; r < R; ++r) k->solute_disp[r] = params[r].solute_disp ? 1 : 0;
    auto &d = k->d;
    d.occ = DevCopy(std::vector<uint8_t>(occ, occ + b.total));
    d.off = DevCopy(b.off);
    d.ncell = DevCopy(std::vector<int32_t>(ncell, ncell + R));
    d.vac = DevCopy(vac);
    d.steps = DevCopy(steps);
    d.max_steps = DevCopy(maxs);
    d.seed = DevCopy(seed);
    d.rid = DevCopy(rid);
    d.time = DevCopy(time);
    d.energy = DevCopy(en);
    d.abs_energy = DevCopy(std::vector<double>(R, 0.0));
    d.temp = DevCopy(temp);
    d.cv = DevCopy(cv);
    d.cs = DevCopy(cs);
    d.solute_mass = DevCopy(smass);
    d.flags = DevCopy(flags);
    d.solvent = DevCopy(solvent);
    d.escaped = DevCopy(std::vector<int32_t>(R, 0));
    d.esc_time = DevCopy(std::vector<double>(R, 0.0));
    d.esc_step = DevCopy(std::vector<uint64_t>(R, 0));
    d.esc_site = DevCopy(std::vector<int64_t>(R, -1));
    d.disp = DevCopy(std::vector<int64_t>(3 * R, 0));
    d.com = DevCopy(std::vector<double>(3 * R, 0.0));
    d.error = DevCopy(std::vector<int32_t>(1, 0));
    {  // chain steps start with the vacancy's first neighbour in id order as the previous site
      std::vector<int64_t> prev(R);
      for (int r = 0; r < R; ++r) {
        int64_t nb[12];
        (void)lmc_fcc_neighbors(ncell[r], vac[r], 1, nb);
        prev[r] = nb[0];
      }
      d.prev_j = DevCopy(prev);
    }
    lmc::Upload(*barrier->m);
    if (energy) {
      RequireModel(energy);
      ComputeTotalEnergy(energy, d.occ, d.off, d.ncell, R, d.abs_energy, nullptr);
    }
    k->formulation = barrier->m->formulation;
    k->total_sites = b.total;
    *out = k.release();
    return LMC_OK;
  });
}

void lmc_kmc_free(lmc_kmc *k) { delete k; }

int lmc_kmc_set_formulation(lmc_kmc *k, int32_t formulation) {
  if (!k || formulation < 0 || formulation > 3) return Fail(LMC_ERR_ARG, "bad formulation");
  if (formulation == 3 && k->sur) return Fail(LMC_ERR_ARG, "chain steps take no surrogate");
  k->formulation = formulation;
  return LMC_OK;
}

static int DeviceCUs() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  return cus;
}

int lmc_kmc_autotune(lmc_kmc *k, int32_t steps, int32_t *chosen, double *ms_dense,
                     double *ms_solvent) {
  double ms[3];
  const int rc = lmc_kmc_autotune_shapes(k, steps, chosen, ms);
  if (rc == LMC_OK) {
    if (ms_dense) *ms_dense = ms[0];
    if (ms_solvent) *ms_solvent = ms[1];
  }
  return rc;
}

int lmc_kmc_autotune_shapes(lmc_kmc *k, int32_t steps, int32_t *chosen, double *ms_per_step) {
  return Guard([&] {
    if (!k || steps <= 0) throw std::invalid_argument("bad argument");
    if (k->formulation == 3) {  // chain steps have one kernel shape
      if (chosen) *chosen = 3;
      if (ms_per_step)
        for (int f = 0; f < 3; ++f) ms_per_step[f] = std::numeric_limits<double>::quiet_NaN();
      return LMC_OK;
    }
    const auto &t = lmc::Upload(*k->bm->m);
    const size_t R = static_cast<size_t>(k->d.R);
    // Scratch copy of every array the KMC kernel writes; read-only arrays are shared.
    lmc::KmcDev s = k->d;
    s.trace_cap = 0;
    s.occ = nullptr;
    s.log_every = 0;  // no log records from the timing runs
    s.log = nullptr;
    s.log_n = nullptr;
    std::vector<void *> owned;
    auto clone = [&](auto *&dst, auto *src, size_t n) {
      using T = std::remove_pointer_t<std::remove_reference_t<decltype(dst)>>;
      dst = DevAlloc<T>(n);
      owned.push_back(dst);
      Check(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice), "hipMemcpy D2D");
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms[3] = {0, 0, 0};
    // the wide shape only when every replica has a CU of its own
    const int cus = DeviceCUs();
    const int forms = (cus > 0 && k->d.R <= cus) ? 3 : 2;
    try {
      Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
      Check(hipEventCreate(&e0), "hipEventCreate");
      Check(hipEventCreate(&e1), "hipEventCreate");
      for (int f = 0; f < forms; ++f) {
        for (void *p : owned) (void)hipFree(p);
        owned.clear();
        clone(s.occ, k->d.occ, static_cast<size_t>(k->total_sites));
        clone(s.vac, k->d.vac, R);
        clone(s.steps, k->d.steps, R);
        clone(s.time, k->d.time, R);
        clone(s.energy, k->d.energy, R);
        clone(s.abs_energy, k->d.abs_energy, R);
        clone(s.temp, k->d.temp, R);
        clone(s.escaped, k->d.escaped, R);
        clone(s.disp, k->d.disp, 3 * R);
        clone(s.com, k->d.com, 3 * R);
        clone(s.error, k->d.error, 1);
        // the timing runs ignore maximum_steps and the escape stop, so every replica runs all
        // `steps` steps whatever the driver's own limits are
        s.max_steps = DevCopy(std::vector<uint64_t>(R, std::numeric_limits<uint64_t>::max()));
        owned.push_back(s.max_steps);
        {
          std::vector<int32_t> fl(R);
          ToHost(fl.data(), k->d.flags, R);
          for (auto &f : fl) f &= ~2;
          s.flags = DevCopy(fl);
          owned.push_back(s.flags);
        }
        if (k->d.atom) {  // tracked atoms are written by every jump
          clone(s.atom, k->d.atom, static_cast<size_t>(k->total_sites));
          clone(s.shift, k->d.shift, 3 * static_cast<size_t>(k->total_sites));
        }
        Check(hipEventRecord(e0, nullptr), "hipEventRecord");
        Check(lmc::LaunchKmc(t, s, static_cast<uint64_t>(steps), 0, nullptr, f),
              "autotune KMC launch");
        Check(hipEventRecord(e1, nullptr), "hipEventRecord");
        Check(hipEventSynchronize(e1), "hipEventSynchronize");
        Check(hipEventElapsedTime(&ms[f], e0, e1), "hipEventElapsedTime");
        // divide by the steps every replica actually ran
        std::vector<uint64_t> a(R), b(R);
        ToHost(a.data(), s.steps, R);
        ToHost(b.data(), k->d.steps, R);
        uint64_t ran = static_cast<uint64_t>(steps);
        for (size_t r = 0; r < R; ++r) ran = std::min<uint64_t>(ran, a[r] - b[r]);
        if (ran == 0) throw std::runtime_error("autotune: a timing run advanced no step");
        ms[f] /= static_cast<float>(ran);
      }
    } catch (...) {
      for (void *p : owned) (void)hipFree(p);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      throw;
    }
    for (void *p : owned) (void)hipFree(p);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int best = 0;
    for (int f = 1; f < forms; ++f)
      if (ms[f] < ms[best]) best = f;
    k->formulation = best;
    if (chosen) *chosen = k->formulation;
    if (ms_per_step)
      for (int f = 0; f < 3; ++f)
        ms_per_step[f] = f < forms ? ms[f] : std::numeric_limits<double>::quiet_NaN();
    return LMC_OK;
  });
}

int lmc_kmc_set_schedule(lmc_kmc *k, int32_t n_points, const double *time, const double *temp) {
  return Guard([&] {
    if (!k || n_points < 0) throw std::invalid_argument("bad argument");
    std::vector<std::pair<double, double>> pts;
    for (int i = 0; i < n_points; ++i) pts.emplace_back(time[i], temp[i]);
    std::sort(pts.begin(), pts.end());  // TimeTemperatureInterpolator::SortPoints (:32-43)
    for (size_t i = 1; i < pts.size(); ++i)
      if (std::abs(pts[i].first - pts[i - 1].first) < 1e-12)
        throw std::invalid_argument("2 adjacent x values are equal.");
    std::vector<double> tv, Tv;
    for (auto &p : pts) {
      tv.push_back(p.first);
      Tv.push_back(p.second);
    }
    // a captured graph holds the old schedule's pointers and point count
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    k->DropGraph();
    DevFree(k->d.tt_time);
    DevFree(k->d.tt_temp);
    k->d.n_tt = n_points;
    if (n_points) {
      k->d.tt_time = DevCopy(tv);
      k->d.tt_temp = DevCopy(Tv);
    }
    return LMC_OK;
  });
}

int lmc_kmc_set_trace(lmc_kmc *k, uint64_t capacity) {
  return Guard([&] {
    if (!k) throw std::invalid_argument("null driver");
    DevFree(k->d.tr_sel);
    DevFree(k->d.tr_ea);
    DevFree(k->d.tr_de);
    DevFree(k->d.tr_time);
    k->d.trace_cap = capacity;
    k->trace_pos = 0;
    if (capacity) {
      const size_t n = static_cast<size_t>(capacity) * k->d.R;
      k->d.tr_sel = DevAlloc<int64_t>(n);
      k->d.tr_ea = DevAlloc<double>(n);
      k->d.tr_de = DevAlloc<double>(n);
      k->d.tr_time = DevAlloc<double>(n);
      Check(hipMemset(k->d.tr_sel, 0xff, n * sizeof(int64_t)), "hipMemset");
    }
    return LMC_OK;
  });
}

static hipError_t RunSurrogateKmc(lmc_kmc *k, uint64_t steps, hipStream_t s);
static hipError_t PrepareSurrogate(lmc_kmc *k);

static void BuildGraph(lmc_kmc *k, uint64_t steps_per_launch, int32_t launches);

int lmc_kmc_build_graph(lmc_kmc *k, uint64_t steps_per_launch, int32_t launches) {
  return Guard([&] {
    if (!k) throw std::invalid_argument("null driver");
    BuildGraph(k, steps_per_launch, launches);
    return LMC_OK;
  });
}

int lmc_kmc_run_graph(lmc_kmc *k, uint64_t steps_per_launch, int32_t launches, void *stream) {
  return Guard([&] {
    if (!k) throw std::invalid_argument("null driver");
    BuildGraph(k, steps_per_launch, launches);
    Check(hipGraphLaunch(k->graph, S(stream)), "hipGraphLaunch");
    return LMC_OK;
  });
}

static void BuildGraph(lmc_kmc *k, uint64_t steps_per_launch, int32_t launches) {
  {
    if (steps_per_launch == 0 || launches <= 0) throw std::invalid_argument("nothing to run");
    if (k->d.trace_cap) throw std::invalid_argument("graph replay needs tracing off");
    const auto &t = lmc::Upload(*k->bm->m);
    if (!k->graph || k->graph_steps != steps_per_launch || k->graph_launches != launches ||
        k->graph_form != k->formulation || k->graph_sur != static_cast<const void *>(k->sur)) {
      k->DropGraph();
      if (!k->capture) Check(hipStreamCreateWithFlags(&k->capture, hipStreamNonBlocking), "hipStreamCreate");
      if (k->sur)
        Check(PrepareSurrogate(k), "surrogate KMC set-up");
      else
        Check(lmc::PrepareKmc(t, k->d.R, k->formulation), "KMC set-up");
      Check(hipStreamBeginCapture(k->capture, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      hipError_t e = hipSuccess;
      for (int32_t i = 0; i < launches && e == hipSuccess; ++i)
        e = k->sur ? RunSurrogateKmc(k, steps_per_launch, k->capture)
                   : lmc::LaunchKmc(t, k->d, steps_per_launch, 0, k->capture, k->formulation);
      hipGraph_t g = nullptr;
      const hipError_t e2 = hipStreamEndCapture(k->capture, &g);
      Check(e, "KMC launch under capture");
      Check(e2, "hipStreamEndCapture");
      const hipError_t e3 = hipGraphInstantiate(&k->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      Check(e3, "hipGraphInstantiate");
      k->graph_steps = steps_per_launch;
      k->graph_launches = launches;
      k->graph_form = k->formulation;
      k->graph_sur = k->sur;
    }
  }
}

int lmc_kmc_run(lmc_kmc *k, uint64_t steps, void *stream) {
  return Guard([&] {
    if (!k) throw std::invalid_argument("null driver");
    const auto &t = lmc::Upload(*k->bm->m);
    if (k->sur)
      Check(RunSurrogateKmc(k, steps, S(stream)), "surrogate KMC kernel");
    else
      Check(lmc::LaunchKmc(t, k->d, steps, k->trace_pos, S(stream), k->formulation),
            "KMC kernel");
    k->trace_pos += steps;
    return LMC_OK;
  });
}

static int KmcErrors(lmc_kmc *k) {
  int32_t e = 0;
  Check(hipMemcpy(&e, k->d.error, sizeof(e), hipMemcpyDeviceToHost), "hipMemcpy");
  if (e) throw ModelError("Cluster not found in ClusterIndexer or element without barrier parameters");
  return LMC_OK;
}

int lmc_kmc_state(lmc_kmc *k, uint64_t *steps, double *time, double *energy, double *abs_energy,
                  double *temperature, int64_t *vacancy, int64_t *disp, double *com,
                  int32_t *escaped) {
  return Guard([&] {
    if (!k) throw std::invalid_argument("null driver");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const size_t R = static_cast<size_t>(k->d.R);
    ToHost(steps, k->d.steps, R);
    ToHost(time, k->d.time, R);
    ToHost(energy, k->d.energy, R);
    ToHost(abs_energy, k->d.abs_energy, R);
    ToHost(temperature, k->d.temp, R);
    ToHost(vacancy, k->d.vac, R);
    ToHost(disp, k->d.disp, 3 * R);
    ToHost(com, k->d.com, 3 * R);
    ToHost(escaped, k->d.escaped, R);
    return KmcErrors(k);
  });
}

int lmc_kmc_escape_info(lmc_kmc *k, double *t_exit, uint64_t *step, int64_t *site_before) {
  return Guard([&] {
    if (!k) throw std::invalid_argument("null driver");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const size_t R = static_cast<size_t>(k->d.R);
    ToHost(t_exit, k->d.esc_time, R);
    ToHost(step, k->d.esc_step, R);
    ToHost(site_before, k->d.esc_site, R);
    return LMC_OK;
  });
}

int lmc_kmc_occupancy(lmc_kmc *k, uint8_t *occ) {
  return Guard([&] {
    if (!k || !occ) throw std::invalid_argument("null argument");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    std::vector<int64_t> off(k->d.R);
    std::vector<int32_t> nc(k->d.R);
    ToHost(off.data(), k->d.off, off.size());
    ToHost(nc.data(), k->d.ncell, nc.size());
    const int64_t total = off.back() + 4ll * nc.back() * nc.back() * nc.back();
    ToHost(occ, k->d.occ, static_cast<size_t>(total));
    return LMC_OK;
  });
}

int lmc_kmc_trace(lmc_kmc *k, int64_t *sel, double *ea, double *de, double *time) {
  return Guard([&] {
    if (!k || !k->d.trace_cap) throw std::invalid_argument("trace is off");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const size_t n = static_cast<size_t>(k->d.trace_cap) * k->d.R;
    ToHost(sel, k->d.tr_sel, n);
    ToHost(ea, k->d.tr_ea, n);
    ToHost(de, k->d.tr_de, n);
    ToHost(time, k->d.tr_time, n);
    return LMC_OK;
  });
}

int lmc_kmc_set_atoms(lmc_kmc *k, const int64_t *lattice_to_atom, const int32_t *map_shift) {
  return Guard([&] {
    if (!k) throw std::invalid_argument("null driver");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    UploadAtoms(k->d.R, k->d.off, k->d.ncell, lattice_to_atom, map_shift, k->d.atom, k->d.shift);
    k->DropGraph();
    return LMC_OK;
  });
}

int lmc_kmc_atoms(lmc_kmc *k, int64_t *lattice_to_atom, int32_t *map_shift) {
  return Guard([&] {
    if (!k) throw std::invalid_argument("null driver");
    DownloadAtoms(k->d.R, k->d.off, k->d.ncell, k->d.atom, k->d.shift, lattice_to_atom, map_shift);
    return LMC_OK;
  });
}

}  // extern "C"

namespace {
constexpr uint64_t kMaxLogEvery = 100000000000000000ull;  // 1e17: 10 L stays in 64 bits
constexpr uint64_t kMaxAveragingSteps = 1ull << 31;
constexpr uint64_t kAveragingChunk = 1ull << 16;  // trials per launch with the averaging ring on
}  // namespace

extern "C" {

int lmc_kmc_set_log(lmc_kmc *k, uint64_t log_dump_steps, uint64_t capacity,
                    const double *box_edge) {
  return Guard([&] {
    if (!k) throw std::invalid_argument("null driver");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const int R = k->d.R;
    DevFree(k->d.log_n);
    DevFree(k->d.log);
    DevFree(k->d.log_edge);
    k->d.log_every = 0;
    k->d.log_cap = 0;
    k->DropGraph();
    if (!log_dump_steps) return LMC_OK;
    if (!capacity) throw std::invalid_argument("log capacity must be positive");
    if (log_dump_steps > kMaxLogEvery) throw std::invalid_argument("log_dump_steps is too large");
    std::vector<int32_t> nc(R);
    ToHost(nc.data(), k->d.ncell, nc.size());
    std::vector<double> edge(R);
    for (int r = 0; r < R; ++r) {
      edge[r] = box_edge ? box_edge[r] : 4.046 * nc[r];  // GenerateFCC's basis (Constants.hpp:6)
      if (!(edge[r] > 0)) throw std::invalid_argument("box edges must be positive");
    }
    k->d.log_n = DevCopy(std::vector<uint64_t>(R, 0));
    k->d.log = DevAlloc<lmc::KmcLogRecord>(static_cast<size_t>(capacity) * R);
    k->d.log_edge = DevCopy(edge);
    k->d.log_every = log_dump_steps;
    k->d.log_cap = capacity;
    k->log_begun.assign(R, 0);
    k->log_fixed.assign(R, 0);
    // a restarted run skips its first dump (McAbstract's is_restarted_, KineticMcAbstract.cpp:60-63)
    k->log_skip_first.resize(R);
    for (int r = 0; r < R; ++r) k->log_skip_first[r] = k->start_steps[r] > 0 ? 1 : 0;
    return LMC_OK;
  });
}

namespace {
// A double as McAbstract's log stream prints it: precision(16), the default float format until
// the first vector has set std::fixed (VectorMatrix.hpp:41-51).
void PutDouble(std::string &o, double v, bool fixed) {
  char buf[64];
  const int n = std::snprintf(buf, sizeof(buf), fixed ? "%.16f" : "%.16g", v);
  o.append(buf, static_cast<size_t>(n));
}
}  // namespace

int lmc_kmc_write_log(lmc_kmc *k, const char *const *filenames) {
  return Guard([&] {
    if (!k || !filenames) throw std::invalid_argument("null argument");
    if (!k->d.log_every) throw std::invalid_argument("logging is off");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const int R = k->d.R;
    std::vector<uint64_t> cnt(R);
    ToHost(cnt.data(), k->d.log_n, cnt.size());
    std::vector<lmc::KmcLogRecord> rec(static_cast<size_t>(k->d.log_cap) * R);
    ToHost(rec.data(), k->d.log, rec.size());
    Check(hipMemset(k->d.log_n, 0, sizeof(uint64_t) * R), "hipMemset");
    uint64_t lost = 0;
    for (int r = 0; r < R; ++r) {
      if (!filenames[r]) continue;
      std::string o;
      const bool fresh = !k->log_begun[r];
      if (fresh && k->start_steps[r] == 0) {  // header of a run from step 0 (Dump, :64-70)
        o += "steps\ttime\ttemperature\tenergy\tEa\tdE\tselected\tvac1\tvac2\tvac3";
        if (k->solute_disp[r]) o += "\tsolute_com1\tsolute_com2\tsolute_com3";
        o += '\n';
      }
      const uint64_t n = std::min(cnt[r], k->d.log_cap);
      lost += cnt[r] - n;
      for (uint64_t i = 0; i < n; ++i) {
        const lmc::KmcLogRecord &L = rec[static_cast<size_t>(r) * k->d.log_cap + i];
        if (k->log_skip_first[r]) {
          k->log_skip_first[r] = 0;
          if (L.steps == k->start_steps[r]) continue;
        }
        const bool fx = k->log_fixed[r] != 0;
        o += std::to_string(L.steps);
        for (double v : {L.time, L.temperature, L.energy, L.ea, L.de}) {
          o += '\t';
          PutDouble(o, v, fx);
        }
        o += '\t';
        o += std::to_string(L.atom);
        o += '\t';
        for (int d = 0; d < 3; ++d) {
          if (d) o += ' ';
          PutDouble(o, L.vac[d], true);
        }
        k->log_fixed[r] = 1;
        if (k->solute_disp[r]) {
          o += '\t';
          for (int d = 0; d < 3; ++d) {
            if (d) o += ' ';
            PutDouble(o, L.com[d], true);
          }
        }
        o += '\n';
      }
      // a run from step 0 truncates (std::ofstream::out); a restarted one appends
      std::FILE *f = std::fopen(filenames[r], fresh && k->start_steps[r] == 0 ? "wb" : "ab");
      if (!f) throw std::runtime_error(std::string("Cannot open ") + filenames[r]);
      const bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
      if (std::fclose(f) != 0 || !ok) throw std::runtime_error(std::string("Cannot write ") + filenames[r]);
      k->log_begun[r] = 1;
    }
    if (lost)
      throw std::invalid_argument(std::to_string(lost) +
                                  " log records exceeded the capacity and were dropped");
    return LMC_OK;
  });
}

int lmc_kmc_device_views(lmc_kmc *k, uint8_t **d_occ, const int64_t **d_offset,
                         const int32_t **d_ncell, const int64_t **d_vac) {
  if (!k) return Fail(LMC_ERR_ARG, "null driver");
  if (d_occ) *d_occ = k->d.occ;
  if (d_offset) *d_offset = k->d.off;
  if (d_ncell) *d_ncell = k->d.ncell;
  if (d_vac) *d_vac = k->d.vac;
  return LMC_OK;
}

int lmc_cmc_create(const lmc_model *pairsite, const lmc_model *energy, int32_t R,
                   const int32_t *ncell, const uint8_t *occ, const lmc_cmc_replica *params,
                   lmc_cmc **out) {
  return Guard([&] {
    RequireModel(pairsite);
    if (!ncell || !occ || !params || !out) throw std::invalid_argument("null argument");
    const Batch b = Layout(R, ncell);
    const Model &m = *pairsite->m;
    auto c = std::make_unique<lmc_cmc>();
    c->pm = pairsite;
    c->d.R = R;
    for (int r = 0; r < R; ++r)
      c->d.max_sites = std::max<int64_t>(c->d.max_sites, 4ll * ncell[r] * ncell[r] * ncell[r]);
    std::vector<uint64_t> steps(R), maxs(R), seed(R);
    std::vector<uint32_t> rid(R);
    std::vector<double> t0(R), t1(R), en(R);
    for (int r = 0; r < R; ++r) {
      const int64_t N = 4ll * ncell[r] * ncell[r] * ncell[r];
      const uint8_t *o = occ + b.off[r];
      bool two = false;
      for (int64_t i = 0; i < N; ++i) {
        if (o[i] >= m.S) throw std::invalid_argument("occupancy holds a species code >= S");
        two |= o[i] != o[0];
      }
      if (!two) throw std::invalid_argument("canonical MC needs at least two species");
      const lmc_cmc_replica &p = params[r];
      if (!(p.t_start > 0) || !(p.t_end > 0) || p.maximum_steps == 0)
        throw std::invalid_argument("temperatures and maximum_steps must be positive");
      steps[r] = p.restart_steps;
      maxs[r] = p.maximum_steps;
      seed[r] = p.seed;
      rid[r] = p.replica_id;
      t0[r] = p.t_start;
      t1[r] = p.t_end;
      en[r] = p.restart_energy;
    }
    auto &d = c->d;
    d.occ = DevCopy(std::vector<uint8_t>(occ, occ + b.total));
    d.off = DevCopy(b.off);
    d.ncell = DevCopy(std::vector<int32_t>(ncell, ncell + R));
    d.steps = DevCopy(steps);
    d.max_steps = DevCopy(maxs);
    d.seed = DevCopy(seed);
    d.accepted = DevCopy(std::vector<uint64_t>(R, 0));
    d.rid = DevCopy(rid);
    d.t0 = DevCopy(t0);
    d.t1 = DevCopy(t1);
    d.energy = DevCopy(en);
    d.abs_energy = DevCopy(std::vector<double>(R, 0.0));
    d.error = DevCopy(std::vector<int32_t>(1, 0));
    c->start_steps = steps;
    lmc::Upload(*pairsite->m);
    if (energy) {
      RequireModel(energy);
      ComputeTotalEnergy(energy, d.occ, d.off, d.ncell, R, d.abs_energy, nullptr);
    }
    *out = c.release();
    return LMC_OK;
  });
}

void lmc_cmc_free(lmc_cmc *c) { delete c; }

int lmc_cmc_set_trace(lmc_cmc *c, uint64_t capacity) {
  return Guard([&] {
    if (!c) throw std::invalid_argument("null driver");
    DevFree(c->d.tr_p);
    DevFree(c->d.tr_q);
    DevFree(c->d.tr_de);
    DevFree(c->d.tr_acc);
    c->d.trace_cap = capacity;
    c->trace_pos = 0;
    if (capacity) {
      const size_t n = static_cast<size_t>(capacity) * c->d.R;
      c->d.tr_p = DevAlloc<int64_t>(n);
      c->d.tr_q = DevAlloc<int64_t>(n);
      c->d.tr_de = DevAlloc<double>(n);
      c->d.tr_acc = DevAlloc<int8_t>(n);
      Check(hipMemset(c->d.tr_p, 0xff, n * sizeof(int64_t)), "hipMemset");
    }
    return LMC_OK;
  });
}

int lmc_cmc_set_atoms(lmc_cmc *c, const int64_t *lattice_to_atom, const int32_t *map_shift) {
  return Guard([&] {
    if (!c) throw std::invalid_argument("null driver");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    UploadAtoms(c->d.R, c->d.off, c->d.ncell, lattice_to_atom, map_shift, c->d.atom, c->d.shift);
    return LMC_OK;
  });
}

int lmc_cmc_atoms(lmc_cmc *c, int64_t *lattice_to_atom, int32_t *map_shift) {
  return Guard([&] {
    if (!c) throw std::invalid_argument("null driver");
    DownloadAtoms(c->d.R, c->d.off, c->d.ncell, c->d.atom, c->d.shift, lattice_to_atom, map_shift);
    return LMC_OK;
  });
}

int lmc_cmc_set_lds_occupancy(lmc_cmc *c, int32_t enable) {
  return Guard([&] {
    if (!c) throw std::invalid_argument("null driver");
    c->d.lds_occupancy = enable ? 1 : 0;
    return LMC_OK;
  });
}

int lmc_cmc_set_parallel_trials(lmc_cmc *c, int32_t trials_per_round) {
  return Guard([&] {
    if (!c) throw std::invalid_argument("null driver");
    if (trials_per_round < 0 || trials_per_round > 64)
      throw std::invalid_argument("trials per round must be in [0, 64]");
    if (trials_per_round && c->d.anneal)
      throw std::invalid_argument("CanonicalMcOmp rounds take no annealing schedule");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    DevFree(c->d.omp_p);
    DevFree(c->d.omp_q);
    DevFree(c->d.omp_de);
    DevFree(c->d.omp_len);
    DevFree(c->d.omp_pos);
    c->d.omp_k = trials_per_round;
    if (!trials_per_round) return LMC_OK;
    const size_t R = static_cast<size_t>(c->d.R), n = R * static_cast<size_t>(trials_per_round);
    c->d.omp_p = DevAlloc<int64_t>(n);
    c->d.omp_q = DevAlloc<int64_t>(n);
    c->d.omp_de = DevAlloc<double>(n);
    c->d.omp_len = DevCopy(std::vector<int32_t>(R, 0));
    c->d.omp_pos = DevCopy(std::vector<int32_t>(R, 0));
    return LMC_OK;
  });
}

int lmc_cmc_round_left(lmc_cmc *c, int32_t *left) {
  return Guard([&] {
    if (!c || !left) throw std::invalid_argument("null argument");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const size_t R = static_cast<size_t>(c->d.R);
    if (!c->d.omp_k) {
      std::fill(left, left + R, 0);
      return LMC_OK;
    }
    std::vector<int32_t> len(R), pos(R);
    ToHost(len.data(), c->d.omp_len, R);
    ToHost(pos.data(), c->d.omp_pos, R);
    for (size_t r = 0; r < R; ++r) left[r] = len[r] - pos[r];
    return LMC_OK;
  });
}

int lmc_cmc_set_annealing(lmc_cmc *c, int32_t enable) {
  return Guard([&] {
    if (!c) throw std::invalid_argument("null driver");
    if (enable && c->d.omp_k)
      throw std::invalid_argument("CanonicalMcOmp rounds take no annealing schedule");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    DevFree(c->d.anneal);
    c->d.anneal = nullptr;
    if (!enable) return LMC_OK;
    const size_t R = static_cast<size_t>(c->d.R);
    std::vector<double> t0(R), en(R);
    std::vector<uint64_t> st(R);
    Check(hipMemcpy(t0.data(), c->d.t0, R * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    Check(hipMemcpy(en.data(), c->d.energy, R * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    Check(hipMemcpy(st.data(), c->d.steps, R * sizeof(uint64_t), hipMemcpyDeviceToHost), "hipMemcpy");
    std::vector<lmc::CmcAnneal> a(R);
    for (size_t r = 0; r < R; ++r) {  // SimulatedAnnealing constructor (SimulatedAnnealing.cpp:66-73)
      if (!(t0[r] > 0)) throw std::invalid_argument("annealing needs a positive t_start");
      a[r] = lmc::CmcAnneal{t0[r], en[r], en[r], st[r], 0, 0, 0, 0, 0};
    }
    c->d.anneal = DevCopy(a);
    return LMC_OK;
  });
}

int lmc_cmc_annealing_state(lmc_cmc *c, double *temperature, double *lowest, uint32_t *reheats) {
  return Guard([&] {
    if (!c) throw std::invalid_argument("null driver");
    if (!c->d.anneal) throw std::invalid_argument("annealing is off");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const size_t R = static_cast<size_t>(c->d.R);
    std::vector<lmc::CmcAnneal> a(R);
    Check(hipMemcpy(a.data(), c->d.anneal, R * sizeof(lmc::CmcAnneal), hipMemcpyDeviceToHost),
          "hipMemcpy");
    for (size_t r = 0; r < R; ++r) {
      if (temperature) temperature[r] = a[r].T;
      if (lowest) lowest[r] = a[r].lowest;
      if (reheats) reheats[r] = a[r].reheats;
    }
    return LMC_OK;
  });
}

int lmc_cmc_run(lmc_cmc *c, uint64_t trials, void *stream) {
  return Guard([&] {
    if (!c) throw std::invalid_argument("null driver");
    const auto &t = lmc::Upload(*c->pm->m);
    // with the averaging ring on, launches of at most ring_cap - avg_m trials, each followed by
    // the averages of its records (the ring still holds every record's window then)
    const uint64_t chunk = c->d.avg_m ? c->d.ring_cap - c->d.avg_m : trials;
    for (uint64_t done = 0; done < trials;) {
      const uint64_t k = std::min(chunk, trials - done);
      Check(lmc::LaunchCmc(t, c->d, k, c->trace_pos, S(stream)), "CMC kernel");
      Check(lmc::LaunchCmcAverage(c->d, S(stream)), "CMC averaging kernel");
      c->trace_pos += k;
      done += k;
    }
    return LMC_OK;
  });
}

int lmc_cmc_set_log(lmc_cmc *c, uint64_t log_dump_steps, uint64_t capacity,
                    uint64_t thermodynamic_averaging_steps) {
  return Guard([&] {
    if (!c) throw std::invalid_argument("null driver");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    c->FreeLog();
    if (!log_dump_steps) return LMC_OK;
    if (!capacity) throw std::invalid_argument("log capacity must be positive");
    if (log_dump_steps > kMaxLogEvery) throw std::invalid_argument("log_dump_steps is too large");
    const int R = c->d.R;
    auto &d = c->d;
    d.log_n = DevCopy(std::vector<uint64_t>(R, 0));
    d.log_lo = DevCopy(std::vector<uint64_t>(R, 0));
    d.log = DevAlloc<lmc::CmcLogRecord>(static_cast<size_t>(capacity) * R);
    if (thermodynamic_averaging_steps) {
      if (thermodynamic_averaging_steps > kMaxAveragingSteps)
        throw std::invalid_argument("thermodynamic_averaging_steps is too large");
      d.ring_cap = thermodynamic_averaging_steps + kAveragingChunk;
      d.ring = DevAlloc<double>(static_cast<size_t>(d.ring_cap) * R);
      d.ring_n = DevCopy(std::vector<uint64_t>(R, 0));
      d.avg_m = thermodynamic_averaging_steps;
    }
    d.log_every = log_dump_steps;
    d.log_cap = capacity;
    c->log_begun.assign(R, 0);
    // a restarted canonical run skips its first dump (CanonicalMcAbstract.cpp:54-57)
    c->log_skip_first.resize(R);
    for (int r = 0; r < R; ++r) c->log_skip_first[r] = c->start_steps[r] > 0 ? 1 : 0;
    return LMC_OK;
  });
}

int lmc_cmc_write_log(lmc_cmc *c, const char *const *filenames) {
  return Guard([&] {
    if (!c || !filenames) throw std::invalid_argument("null argument");
    if (!c->d.log_every) throw std::invalid_argument("logging is off");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const int R = c->d.R;
    const bool anneal = c->d.anneal != nullptr;
    std::vector<uint64_t> cnt(R);
    ToHost(cnt.data(), c->d.log_n, cnt.size());
    std::vector<lmc::CmcLogRecord> rec(static_cast<size_t>(c->d.log_cap) * R);
    ToHost(rec.data(), c->d.log, rec.size());
    Check(hipMemset(c->d.log_n, 0, sizeof(uint64_t) * R), "hipMemset");
    uint64_t lost = 0;
    for (int r = 0; r < R; ++r) {
      if (!filenames[r]) continue;
      std::string o;
      const bool fresh = !c->log_begun[r];
      if (fresh && c->start_steps[r] == 0)  // header of a run from step 0
        o += anneal ? "steps\ttemperature\tenergy\tlowest_energy\tabsolute_energy\n"
                    : "steps\ttemperature\tenergy\taverage_energy\tabsolute_energy\n";
      const uint64_t n = std::min(cnt[r], c->d.log_cap);
      lost += cnt[r] - n;
      for (uint64_t i = 0; i < n; ++i) {
        const lmc::CmcLogRecord &L = rec[static_cast<size_t>(r) * c->d.log_cap + i];
        if (c->log_skip_first[r]) {
          c->log_skip_first[r] = 0;
          if (!anneal && L.steps == c->start_steps[r]) continue;
        }
        o += std::to_string(L.steps);
        for (double v : {L.temperature, L.energy, L.second, L.abs_energy}) {
          o += '\t';
          PutDouble(o, v, false);  // the stream never switches to std::fixed here
        }
        o += '\n';
      }
      // a run from step 0 truncates (std::ofstream::out); a restarted one appends
      std::FILE *f = std::fopen(filenames[r], fresh && c->start_steps[r] == 0 ? "wb" : "ab");
      if (!f) throw std::runtime_error(std::string("Cannot open ") + filenames[r]);
      const bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
      if (std::fclose(f) != 0 || !ok) throw std::runtime_error(std::string("Cannot write ") + filenames[r]);
      c->log_begun[r] = 1;
    }
    if (lost)
      throw std::invalid_argument(std::to_string(lost) +
                                  " log records past the capacity were lost; write the log more "
                                  "often or raise its capacity");
    return LMC_OK;
  });
}

int lmc_cmc_state(lmc_cmc *c, uint64_t *steps, double *energy, double *abs_energy,
                  uint64_t *accepted) {
  return Guard([&] {
    if (!c) throw std::invalid_argument("null driver");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const size_t R = static_cast<size_t>(c->d.R);
    ToHost(steps, c->d.steps, R);
    ToHost(energy, c->d.energy, R);
    ToHost(abs_energy, c->d.abs_energy, R);
    ToHost(accepted, c->d.accepted, R);
    int32_t e = 0;
    Check(hipMemcpy(&e, c->d.error, sizeof(e), hipMemcpyDeviceToHost), "hipMemcpy");
    if (e) throw ModelError("no pair of distinct species found");
    return LMC_OK;
  });
}

int lmc_cmc_occupancy(lmc_cmc *c, uint8_t *occ) {
  return Guard([&] {
    if (!c || !occ) throw std::invalid_argument("null argument");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    std::vector<int64_t> off(c->d.R);
    std::vector<int32_t> nc(c->d.R);
    ToHost(off.data(), c->d.off, off.size());
    ToHost(nc.data(), c->d.ncell, nc.size());
    const int64_t total = off.back() + 4ll * nc.back() * nc.back() * nc.back();
    ToHost(occ, c->d.occ, static_cast<size_t>(total));
    return LMC_OK;
  });
}

int lmc_cmc_trace(lmc_cmc *c, int64_t *p, int64_t *q, double *de, int8_t *accepted) {
  return Guard([&] {
    if (!c || !c->d.trace_cap) throw std::invalid_argument("trace is off");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const size_t n = static_cast<size_t>(c->d.trace_cap) * c->d.R;
    ToHost(p, c->d.tr_p, n);
    ToHost(q, c->d.tr_q, n);
    ToHost(de, c->d.tr_de, n);
    ToHost(accepted, c->d.tr_acc, n);
    return LMC_OK;
  });
}

}  // extern "C"

// ---- Learned surrogate --------------------------------------------------------------------

struct lmc_surrogate {
  int M = 0, C = 0, Kp = 0, S = 0, x_code = 0, base_code = 0;
  int log_ekra = 0;  // E0 models: output 0 is ln E_kra (the model's own logit scale)
  int v_slot = 0;
  int max_solutes = lmc::kEnvUsed;  // most non-base, non-vacancy slots seen in training
  lmc::SurDev Dev(double threshold, double window, uint64_t *n_sur, uint64_t *n_fb) const {
    lmc::SurDev d{};
    d.w = View();
    for (int i = 0; i < 2; ++i) {
      d.mu[i] = mu[i];
      d.sigma[i] = sigma[i];
    }
    d.log_ekra = log_ekra;
    d.base_code = base_code;
    d.max_solutes = max_solutes;
    d.threshold = threshold;
    d.window = window;
    d.n_sur = n_sur;
    d.n_fb = n_fb;
    return d;
  }
  int chan[4] = {-1, -1, -1, -1};
  lmc::MlpLayout lay{};
  double mu[2] = {0.0, 0.0}, sigma[2] = {0.0, 0.0};
  float *master = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  uint16_t *w1t = nullptr, *w2t = nullptr, *w2k = nullptr, *w3 = nullptr, *w1f = nullptr,
           *w2f = nullptr;
  uint64_t adam_t = 0;
  ~lmc_surrogate() {
    DevFree(master);
    DevFree(adam_m);
    DevFree(adam_v);
    DevFree(w1t);
    DevFree(w2t);
    DevFree(w2k);
    DevFree(w3);
    DevFree(w1f);
    DevFree(w2f);
  }
  lmc::MlpDev View() const {
    lmc::MlpDev v{};
    v.M = M;
    v.C = C;
    v.Kp = Kp;
    v.lay = lay;
    for (int i = 0; i < 4; ++i) v.chan[i] = chan[i];
    v.master = master;
    v.w1t = w1t;
    v.w2t = w2t;
    v.w2k = w2k;
    v.w3 = w3;
    v.w1f = w1f;
    v.w2f = w2f;
    return v;
  }
  lmc::MlpMut Mut() {
    lmc::MlpMut v{};
    v.M = M;
    v.C = C;
    v.Kp = Kp;
    v.lay = lay;
    v.master = master;
    v.adam_m = adam_m;
    v.adam_v = adam_v;
    v.w1t = w1t;
    v.w2t = w2t;
    v.w2k = w2k;
    v.w3 = w3;
    v.w1f = w1f;
    v.w2f = w2f;
    return v;
  }
  void RefreshCopies() {  // bf16 copies from the master parameters
    const lmc::AdamStep a{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    Check(lmc::LaunchMlpReduceAdam(Mut(), nullptr, 0, nullptr, 0, a, nullptr), "surrogate copy");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  }
};

namespace {

void RequireSurrogate(const lmc_surrogate *s) {
  if (!s || !s->master) throw std::invalid_argument("null surrogate");
}

// Uniform (-a, a) from Philox (seed, counter i, member, tag).
float InitUniform(uint64_t seed, uint64_t i, uint32_t member, float a) {
  const lmc::U4 u = lmc::Philox4x32_10(seed, i, member, 0x5EEDu);
  const double x = (u.x >> 8) * (1.0 / 16777216.0);  // [0, 1)
  return static_cast<float>((2.0 * x - 1.0) * a);
}

}  // namespace

int lmc_surrogate_create(const lmc_model *barrier, int32_t n_members, uint64_t seed,
                         lmc_surrogate **out) {
  return Guard([&] {
    RequireModel(barrier);
    if (!out) throw std::invalid_argument("null output");
    if (barrier->m->kind != LMC_MODEL_E0 && barrier->m->kind != LMC_MODEL_QUARTIC)
      throw std::invalid_argument("the surrogate stands in for an E0 or Quartic model");
    if (n_members < 1 || n_members > 8) throw std::invalid_argument("n_members must be in [1, 8]");
    const Model &m = *barrier->m;
    std::unique_ptr<lmc_surrogate> s(new lmc_surrogate);
    s->M = n_members;
    s->S = m.S;
    s->log_ekra = m.kind == LMC_MODEL_E0 ? 1 : 0;
    s->v_slot = lmc::GetJumpTables().v_slot;
    s->x_code = m.x_code;
    s->base_code = lmc::Upload(*barrier->m).solvent_default;
    for (int c = 0; c < m.S; ++c)
      if (c != s->x_code && c != s->base_code) {
        if (s->C == 3) throw std::invalid_argument("the surrogate supports at most 4 elements");
        s->chan[s->C++] = c;
      }
    if (s->C < 2) throw std::invalid_argument("the surrogate needs at least 3 elements");
    s->Kp = (lmc::kEnvUsed * s->C + 31) / 32 * 32;
    s->lay = lmc::MakeMlpLayout(s->Kp);
    const lmc::MlpLayout &L = s->lay;
    const size_t total = static_cast<size_t>(s->M) * L.P;
    std::vector<float> p(total, 0.f);
    for (int mm = 0; mm < s->M; ++mm) {
      float *q = p.data() + static_cast<size_t>(mm) * L.P;
      for (int k = 0; k < lmc::kEnvUsed * s->C; ++k)  // rows past 60 C stay zero
        for (int n = 0; n < lmc::kHidden; ++n)
          q[L.oW1 + k * lmc::kHidden + n] = InitUniform(seed, L.oW1 + k * lmc::kHidden + n, mm, 0.6f);
      for (int i = 0; i < lmc::kHidden * lmc::kHidden; ++i)
        q[L.oW2 + i] = InitUniform(seed, L.oW2 + i, mm, 0.153f);
      for (int i = 0; i < lmc::kHidden * lmc::kMlpOut; ++i)
        q[L.oW3 + i] = InitUniform(seed, L.oW3 + i, mm, 0.21f);
    }
    s->master = DevCopy(p);
    s->adam_m = DevCopy(std::vector<float>(total, 0.f));
    s->adam_v = DevCopy(std::vector<float>(total, 0.f));
    s->w1t = DevAlloc<uint16_t>(static_cast<size_t>(s->M) * lmc::kHidden * s->Kp);
    s->w2t = DevAlloc<uint16_t>(static_cast<size_t>(s->M) * lmc::kHidden * lmc::kHidden);
    s->w2k = DevAlloc<uint16_t>(static_cast<size_t>(s->M) * lmc::kHidden * lmc::kHidden);
    s->w3 = DevAlloc<uint16_t>(static_cast<size_t>(s->M) * lmc::kHidden * lmc::kMlpOut);
    s->w1f = DevAlloc<uint16_t>(static_cast<size_t>(s->M) * lmc::kHidden * s->Kp);
    s->w2f = DevAlloc<uint16_t>(static_cast<size_t>(s->M) * lmc::kHidden * lmc::kHidden);
    s->RefreshCopies();
    *out = s.release();
    return LMC_OK;
  });
}

void lmc_surrogate_free(lmc_surrogate *s) { delete s; }

int lmc_surrogate_layout(const lmc_surrogate *s, int32_t *n_members, int32_t *n_inputs,
                         int32_t *n_params, int32_t *n_channels, int32_t *channel_codes4,
                         int32_t *log_ekra) {
  return Guard([&] {
    RequireSurrogate(s);
    if (log_ekra) *log_ekra = s->log_ekra;
    if (n_members) *n_members = s->M;
    if (n_inputs) *n_inputs = s->Kp;
    if (n_params) *n_params = s->lay.P;
    if (n_channels) *n_channels = s->C;
    if (channel_codes4)
      for (int i = 0; i < 4; ++i) channel_codes4[i] = s->chan[i];
    return LMC_OK;
  });
}

int lmc_surrogate_get_params(const lmc_surrogate *s, float *params) {
  return Guard([&] {
    RequireSurrogate(s);
    if (!params) throw std::invalid_argument("null output");
    ToHost(params, s->master, static_cast<size_t>(s->M) * s->lay.P);
    return LMC_OK;
  });
}

int lmc_surrogate_set_params(lmc_surrogate *s, const float *params) {
  return Guard([&] {
    RequireSurrogate(s);
    if (!params) throw std::invalid_argument("null input");
    Check(hipMemcpy(s->master, params, sizeof(float) * s->M * s->lay.P, hipMemcpyHostToDevice),
          "hipMemcpy");
    s->RefreshCopies();
    return LMC_OK;
  });
}

int lmc_surrogate_set_normalization(lmc_surrogate *s, const double *mu2, const double *sigma2) {
  return Guard([&] {
    RequireSurrogate(s);
    if (!mu2 || !sigma2 || !(sigma2[0] > 0) || !(sigma2[1] > 0))
      throw std::invalid_argument("normalisation needs positive sigmas");
    for (int i = 0; i < 2; ++i) {
      s->mu[i] = mu2[i];
      s->sigma[i] = sigma2[i];
    }
    return LMC_OK;
  });
}

int lmc_surrogate_get_normalization(const lmc_surrogate *s, double *mu2, double *sigma2) {
  return Guard([&] {
    RequireSurrogate(s);
    for (int i = 0; i < 2; ++i) {
      if (mu2) mu2[i] = s->mu[i];
      if (sigma2) sigma2[i] = s->sigma[i];
    }
    return LMC_OK;
  });
}

int lmc_jump_samples(const lmc_model *m, const uint8_t *d_occ, const int64_t *d_offset,
                     const int32_t *d_ncell, const int64_t *d_vac, int32_t n_replicas,
                     uint8_t *d_env, double *d_ekra, double *d_de, void *stream) {
  return Guard([&] {
    RequireModel(m);
    if (m->m->kind != LMC_MODEL_E0 && m->m->kind != LMC_MODEL_QUARTIC)
      throw std::invalid_argument("lmc_jump_samples needs an E0 or Quartic model");
    if (!d_env || !d_ekra || !d_de) throw std::invalid_argument("lmc_jump_samples: null buffer");
    RequireBuffers("lmc_jump_samples", n_replicas, {d_occ, d_offset, d_ncell, d_vac});
    if (n_replicas == 0) return LMC_OK;
    const auto &t = lmc::Upload(*m->m);
    int64_t *nb = DevAlloc<int64_t>(12 * static_cast<size_t>(n_replicas));
    double *ea = DevAlloc<double>(12 * static_cast<size_t>(n_replicas));
    const hipError_t e = lmc::LaunchEvalJumps(t, d_occ, d_offset, d_ncell, d_vac, n_replicas, nb,
                                              ea, d_de, m->m->dev_err, S(stream),
                                              m->m->formulation == 1, d_env, d_ekra);
    (void)hipStreamSynchronize(S(stream));
    DevFree(nb);
    DevFree(ea);
    Check(e, "jump samples kernel");
    return LMC_OK;
  });
}

int lmc_surrogate_forward(const lmc_surrogate *s, const uint8_t *d_env, int32_t n_rows,
                          float *d_out, void *stream) {
  return Guard([&] {
    RequireSurrogate(s);
    if (n_rows < 0 || (n_rows && (!d_env || !d_out))) throw std::invalid_argument("bad arguments");
    if (n_rows) Check(lmc::LaunchMlpForward(s->View(), d_env, n_rows, d_out, S(stream)), "forward");
    return LMC_OK;
  });
}

int lmc_surrogate_gradients(lmc_surrogate *s, const uint8_t *d_env, const float *d_target,
                            int32_t n_rows, float *d_grad, double *loss) {
  return Guard([&] {
    RequireSurrogate(s);
    if (n_rows <= 0 || !d_env || !d_target || !d_grad) throw std::invalid_argument("bad arguments");
    const int tiles = (n_rows + lmc::MlpTileRows() - 1) / lmc::MlpTileRows();
    float *part = DevAlloc<float>(static_cast<size_t>(tiles) * s->M * s->lay.P);
    float *lp = DevAlloc<float>(static_cast<size_t>(tiles) * s->M);
    hipError_t e = lmc::LaunchMlpTrain(s->View(), d_env, nullptr, 0, n_rows, d_target, part, lp, nullptr);
    const lmc::AdamStep a{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (e == hipSuccess) e = lmc::LaunchMlpReduceAdam(s->Mut(), part, tiles, d_grad, 0, a, nullptr);
    std::vector<float> hl(static_cast<size_t>(tiles) * s->M);
    if (e == hipSuccess) e = hipMemcpy(hl.data(), lp, hl.size() * sizeof(float), hipMemcpyDeviceToHost);
    DevFree(part);
    DevFree(lp);
    Check(e, "surrogate gradients");
    if (loss)
      for (int mm = 0; mm < s->M; ++mm) {
        double acc = 0.0;
        for (int b = 0; b < tiles; ++b) acc += hl[static_cast<size_t>(b) * s->M + mm];
        loss[mm] = acc;
      }
    return LMC_OK;
  });
}

int lmc_surrogate_train(lmc_surrogate *s, const uint8_t *d_env, const double *d_ekra,
                        const double *d_de, int32_t n_rows, int32_t epochs, int32_t batch,
                        double lr, uint64_t shuffle_seed, double *mse) {
  return Guard([&] {
    RequireSurrogate(s);
    if (n_rows <= 0 || epochs <= 0 || batch <= 0 || !d_env || !d_ekra || !d_de || !(lr > 0))
      throw std::invalid_argument("bad arguments");
    const size_t n = static_cast<size_t>(n_rows);
    std::vector<double> ek(n), de(n);
    ToHost(ek.data(), d_ekra, n);
    ToHost(de.data(), d_de, n);
    if (s->log_ekra)
      for (double &x : ek) {
        if (!(x > 0)) throw std::invalid_argument("E_kra samples of an E0 model must be positive");
        x = std::log(x);
      }
    if (!(s->sigma[0] > 0) || !(s->sigma[1] > 0)) {  // normalisation from the samples
      const std::vector<double> *v[2] = {&ek, &de};
      for (int o = 0; o < 2; ++o) {
        long double sum = 0, sq = 0;
        for (double x : *v[o]) sum += x;
        const double mean = static_cast<double>(sum / n);
        for (double x : *v[o]) sq += (x - mean) * (x - mean);
        s->mu[o] = mean;
        s->sigma[o] = std::max(1e-6, static_cast<double>(std::sqrt(sq / n)));
      }
    }
    {  // out-of-distribution bound: the most solute slots of any training row
      std::vector<uint8_t> rows(n * lmc::kEnvRow);
      ToHost(rows.data(), d_env, rows.size());
      int most = 0;
      for (size_t i = 0; i < n; ++i) {
        int c = 0;
        for (int j = 0; j < lmc::kEnvUsed; ++j)
          c += (j != s->v_slot && rows[i * lmc::kEnvRow + j] != s->base_code) ? 1 : 0;
        most = std::max(most, c);
      }
      s->max_solutes = most;
    }
    std::vector<float> tg(2 * n);
    for (size_t i = 0; i < n; ++i) {
      tg[2 * i] = static_cast<float>((ek[i] - s->mu[0]) / s->sigma[0]);
      tg[2 * i + 1] = static_cast<float>((de[i] - s->mu[1]) / s->sigma[1]);
    }
    const int rows = lmc::MlpTileRows();
    const int bsz = std::min<int>(batch, n_rows);
    const int tiles_max = (bsz + rows - 1) / rows;
    const int nbatch = (n_rows + bsz - 1) / bsz;
    float *dt = DevCopy(tg);
    int32_t *order = DevAlloc<int32_t>(n);
    float *part = DevAlloc<float>(static_cast<size_t>(tiles_max) * s->M * s->lay.P);
    float *lp = DevAlloc<float>(static_cast<size_t>(nbatch) * tiles_max * s->M);
    std::vector<int32_t> perm(n);
    std::vector<float> hl(static_cast<size_t>(nbatch) * tiles_max * s->M);
    const double total_steps = static_cast<double>(epochs) * nbatch;
    hipError_t e = hipSuccess;
    double step = 0;
    for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) {
      for (size_t i = 0; i < n; ++i) perm[i] = static_cast<int32_t>(i);
      for (size_t i = n - 1; i > 0; --i) {  // Fisher-Yates on Philox (shuffle_seed, epoch)
        const lmc::U4 u = lmc::Philox4x32_10(shuffle_seed, i, static_cast<uint32_t>(ep), 0x5A5Au);
        const uint64_t bits = (static_cast<uint64_t>(u.x) << 32) | u.y;
        const size_t j = static_cast<size_t>((static_cast<unsigned __int128>(bits) * (i + 1)) >> 64);
        std::swap(perm[i], perm[j]);
      }
      e = hipMemcpy(order, perm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMemset(lp, 0, hl.size() * sizeof(float));
      for (int b = 0; b < nbatch && e == hipSuccess; ++b, step += 1) {
        const int row0 = b * bsz, nb = std::min(bsz, n_rows - row0);
        e = lmc::LaunchMlpTrain(s->View(), d_env, order, row0, nb, dt, part,
                                lp + static_cast<size_t>(b) * tiles_max * s->M, nullptr);
        if (e != hipSuccess) break;
        ++s->adam_t;
        const double frac = step / std::max(1.0, total_steps - 1);
        const double lr_t = lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(M_PI * frac)));
        const float b1 = 0.9f, b2 = 0.999f;
        const lmc::AdamStep a{static_cast<float>(lr_t), b1, b2, 1e-8f,
                              static_cast<float>(1.0 / (1.0 - std::pow(b1, s->adam_t))),
                              static_cast<float>(1.0 / (1.0 - std::pow(b2, s->adam_t)))};
        e = lmc::LaunchMlpReduceAdam(s->Mut(), part, (nb + rows - 1) / rows, nullptr, 1, a, nullptr);
      }
      if (e == hipSuccess && ep == epochs - 1)
        e = hipMemcpy(hl.data(), lp, hl.size() * sizeof(float), hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    DevFree(dt);
    DevFree(order);
    DevFree(part);
    DevFree(lp);
    Check(e, "surrogate training");
    if (mse)
      for (int mm = 0; mm < s->M; ++mm) {
        double acc = 0.0;
        for (int b = 0; b < nbatch * tiles_max; ++b) acc += hl[static_cast<size_t>(b) * s->M + mm];
        mse[mm] = acc / (2.0 * n_rows);
      }
    return LMC_OK;
  });
}

static hipError_t PrepareSurrogate(lmc_kmc *k) {
  return lmc::PrepareSurrogateKmc(lmc::Upload(*k->bm->m), k->sur->C);
}

static hipError_t RunSurrogateKmc(lmc_kmc *k, uint64_t steps, hipStream_t s) {
  const auto &t = lmc::Upload(*k->bm->m);
  const lmc::SurDev sd = k->sur->Dev(k->sur_threshold, k->sur_window, k->n_sur, k->n_fb);
  const hipError_t e = lmc::LaunchSurrogateKmc(t, k->d, sd, steps, k->trace_pos, s);
  return e;
}

int lmc_kmc_set_surrogate(lmc_kmc *k, const lmc_surrogate *s, double threshold, double window) {
  return Guard([&] {
    if (!k) throw std::invalid_argument("null driver");
    if (s && s->M != 4) throw std::invalid_argument("the KMC surrogate path needs 4 members");
    if (s && s->S != k->bm->m->S) throw std::invalid_argument("surrogate built for another model");
    if (s && k->formulation == 3) throw std::invalid_argument("chain steps take no surrogate");
    // a captured graph holds the surrogate's parameters by value
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    k->DropGraph();
    k->sur = s;
    k->sur_threshold = threshold;
    k->sur_window = window;
    if (s && !k->n_sur) {
      k->n_sur = DevCopy(std::vector<uint64_t>(k->d.R, 0));
      k->n_fb = DevCopy(std::vector<uint64_t>(k->d.R, 0));
    }
    return LMC_OK;
  });
}

int lmc_kmc_surrogate_stats(lmc_kmc *k, uint64_t *events, uint64_t *fallbacks) {
  return Guard([&] {
    if (!k) throw std::invalid_argument("null driver");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const size_t R = static_cast<size_t>(k->d.R);
    std::vector<uint64_t> a(R, 0), b(R, 0);
    if (k->n_sur) {
      ToHost(a.data(), k->n_sur, R);
      ToHost(b.data(), k->n_fb, R);
    }
    uint64_t sa = 0, sb = 0;
    for (size_t i = 0; i < R; ++i) {
      sa += a[i];
      sb += b[i];
    }
    if (events) *events = sa + sb;
    if (fallbacks) *fallbacks = sb;
    return LMC_OK;
  });
}

int lmc_surrogate_eval_jumps(const lmc_surrogate *s, const lmc_model *m, const uint8_t *d_occ,
                             const int64_t *d_offset, const int32_t *d_ncell,
                             const int64_t *d_vac, const double *d_temp, int32_t n_replicas,
                             double threshold, double window, int64_t *d_nb, double *d_ea,
                             double *d_de, int8_t *d_fallback, void *stream) {
  return Guard([&] {
    RequireSurrogate(s);
    RequireModel(m);
    if (s->M != 4) throw std::invalid_argument("the surrogate evaluation needs 4 members");
    if (m->m->kind != LMC_MODEL_E0 && m->m->kind != LMC_MODEL_QUARTIC)
      throw std::invalid_argument("the surrogate stands in for an E0 or Quartic model");
    if (m->m->S != s->S) throw std::invalid_argument("surrogate built for another model");
    if (n_replicas < 0 || (n_replicas && (!d_occ || !d_offset || !d_ncell || !d_vac || !d_temp ||
                                           !d_nb || !d_ea || !d_de || !d_fallback)))
      throw std::invalid_argument("bad arguments");
    if (n_replicas == 0) return LMC_OK;
    const auto &t = lmc::Upload(*m->m);
    lmc::KmcDev kd{};
    kd.R = n_replicas;
    kd.occ = const_cast<uint8_t *>(d_occ);
    kd.off = const_cast<int64_t *>(d_offset);
    kd.ncell = const_cast<int32_t *>(d_ncell);
    kd.vac = const_cast<int64_t *>(d_vac);
    kd.temp = const_cast<double *>(d_temp);
    kd.error = m->m->dev_err;
    const lmc::SurOut out{d_nb, d_ea, d_de, d_fallback};
    Check(lmc::LaunchSurrogateEval(t, kd, s->Dev(threshold, window, nullptr, nullptr), out,
                                   S(stream)),
          "surrogate evaluation kernel");
    return LMC_OK;
  });
}
