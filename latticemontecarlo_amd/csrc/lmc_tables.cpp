// Exact-integer construction of the jump/site cluster tables and the ECI key space.
//
// The reference builds the same tables per lattice site from floating-point fractional
// positions rotated into each pair's frame and sorted with a 1e-4 tolerance
// (lmc/pred/src/EnergyUtility.cpp:45-104, :169-287, lmc/cfg/include/LatticeCluster.hpp:62-234,
// lmc/cfg/src/Config.cpp:228-264, :1047-1058). Here every key is an integer:
//   R  = 2*(site - vacancy) - d        (site relative to the pair centre, in a/4 units)
//   Xk = R.d, Yk = R.p, Zk = R.q       (rotated coordinates scaled by 2*sqrt2, 2*sqrt2, 2)
//   N8 = Xk^2 + Yk^2 + 2 Zk^2          (8 x squared distance from the centre)
// so the reference's orderings become exact lexicographic orders on these keys.
#include "lmc_tables.h"

#include <algorithm>
#include <map>
#include <set>
#include <stdexcept>
#include <tuple>

namespace lmc {

const I3 kNN1[12] = {{1, 1, 0},  {1, -1, 0},  {-1, 1, 0}, {-1, -1, 0}, {1, 0, 1},  {1, 0, -1},
                     {-1, 0, 1}, {-1, 0, -1}, {0, 1, 1},  {0, 1, -1},  {0, -1, 1}, {0, -1, -1}};
const I3 kNN2[6] = {{2, 0, 0}, {-2, 0, 0}, {0, 2, 0}, {0, -2, 0}, {0, 0, 2}, {0, 0, -2}};
const I3 kNN3[24] = {{2, 1, 1},  {2, 1, -1},  {2, -1, 1},  {2, -1, -1},  {-2, 1, 1},  {-2, 1, -1},
                     {-2, -1, 1}, {-2, -1, -1}, {1, 2, 1},  {1, 2, -1},   {-1, 2, 1},  {-1, 2, -1},
                     {1, -2, 1},  {1, -2, -1},  {-1, -2, 1}, {-1, -2, -1}, {1, 1, 2},  {1, -1, 2},
                     {-1, 1, 2},  {-1, -1, 2},  {1, 1, -2},  {1, -1, -2},  {-1, 1, -2}, {-1, -1, -2}};

static inline I3 Add(const I3 &a, const I3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline I3 Sub(const I3 &a, const I3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline I3 Scale(const I3 &a, int s) { return {a.x * s, a.y * s, a.z * s}; }
static inline int Dot(const I3 &a, const I3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline I3 Cross(const I3 &a, const I3 &b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline bool Eq(const I3 &a, const I3 &b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

int ShellOf(const I3 &d) {
  const int ax = std::abs(d.x), ay = std::abs(d.y), az = std::abs(d.z);
  const int r2 = ax * ax + ay * ay + az * az;
  // FCC shells in a/2 units: <110> r2=2, <200> r2=4, <211> r2=6 (Constants.hpp:8-10 cutoffs
  // 3.5/4.8/5.3 Angstrom separate exactly these at a=4.046).
  if ((ax + ay + az) & 1) return 0;
  if (r2 == 2) return 1;
  if (r2 == 4) return 2;
  if (r2 == 6) return 3;
  return 0;
}

int64_t SiteAdd(int n, int64_t id, const I3 &d) {
  const I3 p = SiteXYZ(n, id);
  const int m = 2 * n;
  auto w = [m](int v) { return ((v % m) + m) % m; };
  return SiteId(n, w(p.x + d.x), w(p.y + d.y), w(p.z + d.z));
}

// Frame of a jump direction d (a <110> vector): p is the perpendicular <110> vector in the
// same {100} plane, q = (d x p)/2 completes a right-handed frame. Any two such frames differ
// by a lattice rotation, and the reference's choice of p (first perpendicular 1NN of the
// vacancy in id order, Config.cpp:253-260) only differs from this one by a half turn about d,
// under which every table built here is invariant.
static void Frame(const I3 &d, I3 &p, I3 &q) {
  int k = d.x == 0 ? 0 : (d.y == 0 ? 1 : 2);
  int c[3] = {d.x, d.y, d.z};
  int pc[3] = {0, 0, 0};
  int i = (k == 0) ? 1 : 0;
  int j = (k == 2) ? 1 : 2;
  pc[i] = c[i];
  pc[j] = -c[j];
  p = {pc[0], pc[1], pc[2]};
  const I3 x = Cross(d, p);
  q = {x.x / 2, x.y / 2, x.z / 2};
}

static int TripletType(int a, int b, int c) {  // GetLabel (EnergyUtility.cpp:353-376)
  int s[3] = {a, b, c};
  std::sort(s, s + 3);
  if (s[0] != 1) return -1;
  if (s[1] == 1 && s[2] == 1) return 4;
  if (s[1] == 1 && s[2] == 2) return 5;
  if (s[1] == 1 && s[2] == 3) return 6;
  if (s[1] == 2 && s[2] == 3) return 7;
  return -1;  // types 8..10 are dropped by the reference (EnergyUtility.cpp:458-469)
}

// Clusters (singlets, pairs, triplets of types 4..7) among `off` containing one of `centre`.
static std::vector<StateCluster> EnumerateClusters(const std::vector<I3> &off,
                                                   const std::vector<int> &centre) {
  auto is_c = [&](int s) { return std::find(centre.begin(), centre.end(), s) != centre.end(); };
  std::vector<StateCluster> out;
  const int n = static_cast<int>(off.size());
  for (int a = 0; a < n; ++a)
    if (is_c(a)) out.push_back({0, {a, -1, -1}, 1});
  for (int a = 0; a < n; ++a)
    for (int b = a + 1; b < n; ++b) {
      if (!is_c(a) && !is_c(b)) continue;
      const int sh = ShellOf(Sub(off[a], off[b]));
      if (sh) out.push_back({sh, {a, b, -1}, 2});
    }
  for (int a = 0; a < n; ++a)
    for (int b = a + 1; b < n; ++b)
      for (int c = b + 1; c < n; ++c) {
        if (!is_c(a) && !is_c(b) && !is_c(c)) continue;
        const int t = TripletType(ShellOf(Sub(off[a], off[b])), ShellOf(Sub(off[b], off[c])),
                                  ShellOf(Sub(off[c], off[a])));
        if (t > 0) out.push_back({t, {a, b, c}, 3});
      }
  std::stable_sort(out.begin(), out.end(),
                   [](const StateCluster &x, const StateCluster &y) { return x.label < y.label; });
  return out;
}

// Orbit tables over the non-endpoint slots (GetAverageClusterParametersMappingMMM/MM2 and
// GetAverageParametersMappingFromLatticeClusterVectorHelperMMM/MM2, EnergyUtility.cpp:106-259).
// Group key: MMM (N8, |Xk|) (GroupCompareMMM), MM2 (N8, Xk) (GroupCompareMM2).
// Within-cluster order: group key, then Xk, Yk, Zk (PositionCompareMMM/MM2).
static OrbitTable BuildOrbits(const JumpTables &t, const std::vector<I3> &off, bool mm2) {
  using GK = std::pair<int, int>;
  auto gkey = [&](int s) {
    const I3 &k = t.key[s];
    const int n8 = k.x * k.x + k.y * k.y + 2 * k.z * k.z;
    return GK{n8, mm2 ? k.x : std::abs(k.x)};
  };
  auto pos_less = [&](int a, int b) {
    const GK ga = gkey(a), gb = gkey(b);
    if (ga != gb) return ga < gb;
    const I3 &ka = t.key[a], &kb = t.key[b];
    return std::make_tuple(ka.x, ka.y, ka.z) < std::make_tuple(kb.x, kb.y, kb.z);
  };
  std::vector<int> slots;
  for (int s = 0; s < kEnvSlots; ++s)
    if (s != t.v_slot && s != t.n_slot) slots.push_back(s);
  struct C {
    int s0, s1;
    bool sym;
    std::vector<GK> keys;
  };
  std::vector<C> cat[4];  // singlets, 1NN, 2NN, 3NN pairs
  for (size_t i = 0; i < slots.size(); ++i) {
    const int a = slots[i];
    cat[0].push_back({a, -1, false, {gkey(a)}});
    for (size_t j = 0; j < i; ++j) {
      int b = slots[j];
      const int sh = ShellOf(Sub(off[a], off[b]));
      if (!sh) continue;
      int x = a, y = b;
      if (pos_less(y, x)) std::swap(x, y);
      cat[sh].push_back({x, y, gkey(x) == gkey(y), {gkey(x), gkey(y)}});
    }
  }
  OrbitTable ot;
  for (int c = 0; c < 4; ++c) {
    auto &v = cat[c];
    std::stable_sort(v.begin(), v.end(), [](const C &a, const C &b) { return a.keys < b.keys; });
    size_t lo = 0;
    while (lo < v.size()) {
      size_t hi = lo;
      while (hi < v.size() && v[hi].keys == v[lo].keys) ++hi;
      const int o = static_cast<int>(ot.orbit_size.size());
      for (size_t i = lo; i < hi; ++i) ot.clusters.push_back({o, v[i].s0, v[i].s1, v[i].sym});
      ot.orbit_size.push_back(static_cast<int>(hi - lo));
      ot.orbit_kind.push_back(c == 0 ? 0 : (v[lo].sym ? 2 : 1));
      ot.orbit_shell.push_back(c);
      lo = hi;
    }
  }
  return ot;
}

void OrbitTable::Layout(int ne, std::vector<int> &base, std::vector<int> &len, int &total) const {
  base.clear();
  len.clear();
  total = 0;
  for (size_t o = 0; o < orbit_size.size(); ++o) {
    const int l = orbit_kind[o] == 0 ? ne : (orbit_kind[o] == 1 ? ne * ne : ne * (ne + 1) / 2);
    base.push_back(total);
    len.push_back(l);
    total += l;
  }
}

static JumpTables BuildJumpTables() {
  JumpTables t;
  const I3 d0 = kNN1[0];
  I3 p0, q0;
  Frame(d0, p0, q0);
  // Environment of the canonical jump 0 -> d0: sites within 3NN of either endpoint
  // (Config::GetNeighborsLatticeIdSetOfPair, Config.cpp:337-352).
  std::vector<I3> env;
  for (int x = -4; x <= 4; ++x)
    for (int y = -4; y <= 4; ++y)
      for (int z = -4; z <= 4; ++z) {
        if ((x + y + z) & 1) continue;
        const I3 o{x, y, z};
        const bool near_v = (x == 0 && y == 0 && z == 0) || ShellOf(o);
        const I3 r = Sub(o, d0);
        const bool near_n = (r.x == 0 && r.y == 0 && r.z == 0) || ShellOf(r);
        if (near_v || near_n) env.push_back(o);
      }
  if (env.size() != kEnvSlots) throw std::logic_error("jump environment is not 60 sites");
  auto keyof = [&](const I3 &o) {
    const I3 R = Sub(Scale(o, 2), d0);
    return I3{Dot(R, d0), Dot(R, p0), Dot(R, q0)};
  };
  // State order (PositionCompareState, LatticeCluster.hpp:62-74): x, y, z of the rotated frame.
  std::sort(env.begin(), env.end(), [&](const I3 &a, const I3 &b) {
    const I3 ka = keyof(a), kb = keyof(b);
    return std::make_tuple(ka.x, ka.y, ka.z) < std::make_tuple(kb.x, kb.y, kb.z);
  });
  for (int s = 0; s < kEnvSlots; ++s) {
    t.key[s] = keyof(env[s]);
    if (Eq(env[s], I3{0, 0, 0})) t.v_slot = s;
    if (Eq(env[s], d0)) t.n_slot = s;
  }
  // Offsets for every direction from the slot keys: R = Xk d/2 + Yk p/2 + Zk q, o = (R + d)/2.
  for (int k = 0; k < 12; ++k) {
    const I3 d = kNN1[k];
    I3 p, q;
    Frame(d, p, q);
    for (int s = 0; s < kEnvSlots; ++s) {
      const I3 &kk = t.key[s];
      const I3 R = Add(Add(Scale(d, kk.x / 2), Scale(p, kk.y / 2)), Scale(q, kk.z));
      const I3 o2 = Add(R, d);
      t.off[k][s] = {o2.x / 2, o2.y / 2, o2.z / 2};
    }
  }
  std::vector<I3> off0(t.off[0].begin(), t.off[0].end());
  t.state = EnumerateClusters(off0, {t.v_slot, t.n_slot});
  for (const auto &c : t.state) {
    bool hv = false, hn = false;
    for (int i = 0; i < c.size; ++i) {
      hv |= c.slots[i] == t.v_slot;
      hn |= c.slots[i] == t.n_slot;
    }
    if (hv && !hn) t.state_v_only.push_back(c);
    if (hn && !hv) t.state_n_only.push_back(c);
  }
  t.mmm = BuildOrbits(t, off0, false);
  t.mm2 = BuildOrbits(t, off0, true);
  for (int s = 0; s < kEnvSlots; ++s) {
    const I3 m{-t.key[s].x, -t.key[s].y, t.key[s].z};
    t.mirror[s] = -1;
    for (int u = 0; u < kEnvSlots; ++u)
      if (Eq(t.key[u], m)) t.mirror[s] = u;
    if (t.mirror[s] < 0) throw std::logic_error("environment is not symmetric");
  }
  for (int s = 0; s < kEnvSlots; ++s)
    if (ShellOf(off0[s])) t.vac_shell_slots.push_back(s);
  return t;
}

static SiteTables BuildSiteTables() {
  SiteTables t;
  std::vector<I3> off{{0, 0, 0}};
  for (const auto &d : kNN1) off.push_back(d);
  for (const auto &d : kNN2) off.push_back(d);
  for (const auto &d : kNN3) off.push_back(d);
  for (int s = 0; s < kSiteSlots; ++s) t.off[s] = off[s];
  t.clusters = EnumerateClusters(off, {0});
  return t;
}

const JumpTables &GetJumpTables() {
  static const JumpTables t = BuildJumpTables();
  return t;
}
const SiteTables &GetSiteTables() {
  static const SiteTables t = BuildSiteTables();
  return t;
}

int EciSpace::Find(int label, std::vector<int> codes) const {
  std::sort(codes.begin(), codes.end());
  for (size_t i = 0; i < keys.size(); ++i)
    if (keys[i].label == label && keys[i].codes == codes) return static_cast<int>(i);
  return -1;
}

EciSpace BuildEciSpace(const std::vector<std::string> &elements_without_x) {
  EciSpace e;
  std::set<std::string> names(elements_without_x.begin(), elements_without_x.end());
  names.insert("X");
  e.species.assign(names.begin(), names.end());  // std::string order == Element::operator<
  for (int i = 0; i < e.S(); ++i)
    if (e.species[i] == "X") e.x_code = i;
  // InitializeClusterHashMap (EnergyUtility.cpp:314-343): X allowed only as the first element.
  std::set<std::pair<std::pair<int, int>, std::vector<int>>> ks;
  const int S = e.S();
  for (int a = 0; a < S; ++a) {
    ks.insert({{1, 0}, {a}});
    for (int b = 0; b < S; ++b) {
      if (b == e.x_code) continue;
      for (int l = 1; l <= 3; ++l) {
        std::vector<int> v{a, b};
        std::sort(v.begin(), v.end());
        ks.insert({{2, l}, v});
      }
      for (int c = 0; c < S; ++c) {
        if (c == e.x_code) continue;
        for (int l = 4; l < 8; ++l) {
          std::vector<int> v{a, b, c};
          std::sort(v.begin(), v.end());
          ks.insert({{3, l}, v});
        }
      }
    }
  }
  // std::set order = (size, label, codes): ElementCluster::operator< (ElementCluster.hpp:40-50).
  for (const auto &k : ks) e.keys.push_back({k.first.second, k.second});
  e.t1.assign(S, -1);
  e.t2.assign(3 * S * S, -1);
  e.t3.assign(4 * S * S * S, -1);
  for (size_t i = 0; i < e.keys.size(); ++i) {
    const auto &k = e.keys[i];
    const int idx = static_cast<int>(i);
    if (k.codes.size() == 1) e.t1[k.codes[0]] = idx;
    if (k.codes.size() == 2) {
      const int l = k.label - 1;
      e.t2[(l * S + k.codes[0]) * S + k.codes[1]] = idx;
      e.t2[(l * S + k.codes[1]) * S + k.codes[0]] = idx;
    }
    if (k.codes.size() == 3) {
      const int l = k.label - 4;
      int c[3] = {k.codes[0], k.codes[1], k.codes[2]};
      std::sort(c, c + 3);
      do {
        e.t3[((l * S + c[0]) * S + c[1]) * S + c[2]] = idx;
      } while (std::next_permutation(c, c + 3));
    }
  }
  return e;
}

}  // namespace lmc
