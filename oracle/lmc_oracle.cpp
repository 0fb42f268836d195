// CPU oracle for the LatticeMC event-evaluation path -- TEST INFRASTRUCTURE ONLY.
//
// Restates, in FP64 C++, the reference algorithms of zhucongx/LatticeMonteCarlo (paths relative
// to the reference repository root). Geometry is restated the way the reference does it:
// floating-point fractional positions, rotation matrices and kEpsilon = 1e-4 tolerance sorts,
// so that it is an independent check of the GPU engine's exact-integer tables.
//
// Deviations from the reference, each forced by reproducibility:
//  * random numbers come from Philox4x32-10 keyed by (seed, replica) with the step counter as
//    counter, instead of a clock-seeded mt19937_64 (lmc/mc/src/McAbstract.cpp:32);
//  * the total rate is summed in event-index order instead of an OpenMP reduction
//    (lmc/mc/src/KineticMcFirstOmp.cpp:55-68), and every dot product is summed left to right
//    instead of Eigen's vectorised order.
//
// What pins each part (tests/test_oracle_pins.py; the reference ships no golden vectors):
//  * Philox4x32-10: the published known-answer vectors;
//  * lattice ids and neighbour shells: closed forms and brute-force distance enumeration;
//  * cluster tables: the counts the reference states (EnergyUtility.cpp:751) and brute force;
//  * GetDe / swap and site dE: equal to differences of the brute-force total CE energy;
//  * E0 and Quartic barriers: the KRA identity Ea(i->j) - Ea(j->i) = dE, special cases with
//    closed-form orbit averages, and the quartic closed form;
//  * KMC/CMC drivers: invariants and the event-choice statistics.
// Parity unpinned: the ORDER of the MMM/MM2 orbit features (GetAverageClusterParametersMapping*,
// EnergyUtility.cpp:169-259). Their grouping and lengths are pinned; the order follows a reading
// of IsClusterSmallerSymmetrically* (LatticeCluster.hpp:211-234) that no reference output
// checks, so coefficient files trained by the reference may index the features differently.
// The LRU key hash (boost::hash_combine) is not restated; it does not change results.
#include "lmc_oracle.h"

#include <algorithm>
#include <atomic>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ------------------------------------------------------------------------------------------
// Constants (lmc/cfg/include/Constants.hpp:6-32, lmc/cfg/include/VectorMatrix.hpp:65)
constexpr double kLatticeConstant = 4.046;
constexpr double kCut1 = 3.5, kCut2 = 4.8, kCut3 = 5.3;
constexpr double kEpsilon = 1e-4;
constexpr double kBoltzmann = 8.617333262145e-5;
constexpr double kPrefactor = 1e13;
constexpr size_t kSizeMax = SIZE_MAX;

// ------------------------------------------------------------------------------------------
// Elements (lmc/cfg/include/Element.hpp:4-68): unknown strings become X; order is by string.
enum class EN : int { X = 0, Al, Mg, Zn, Cu, Sn };
static const char *kElemStr[] = {"X", "Al", "Mg", "Zn", "Cu", "Sn"};
static const double kElemMass[] = {0.0, 26.98, 24.31, 65.38, 63.55, 118.71};
inline std::string Str(EN e) { return kElemStr[static_cast<int>(e)]; }
inline EN FromStr(const std::string &s) {
  for (int i = 0; i < 6; ++i)
    if (s == kElemStr[i]) return static_cast<EN>(i);
  return EN::X;
}
struct ElemLess {
  bool operator()(EN a, EN b) const { return Str(a) < Str(b); }
};

// ------------------------------------------------------------------------------------------
// Vectors (lmc/cfg/include/VectorMatrix.hpp)
using V3 = std::array<double, 3>;
using M3 = std::array<V3, 3>;
inline V3 operator-(const V3 &a, const V3 &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 operator+(const V3 &a, const V3 &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 operator*(const V3 &a, double f) { return {a[0] * f, a[1] * f, a[2] * f}; }
inline V3 operator-(const V3 &a, double f) { return {a[0] - f, a[1] - f, a[2] - f}; }
inline double Inner(const V3 &a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }
inline double Dot(const V3 &a, const V3 &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 Cross(const V3 &a, const V3 &b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline V3 Normalize(const V3 &a) { return a * (1.0 / std::sqrt(Inner(a))); }
inline V3 Floor(const V3 &a) { return {std::floor(a[0]), std::floor(a[1]), std::floor(a[2])}; }
// row vector times matrix (VectorMatrix.hpp, operator*(Vector_d, Matrix_d))
inline V3 MulVM(const V3 &v, const M3 &m) {
  return {v[0] * m[0][0] + v[1] * m[1][0] + v[2] * m[2][0],
          v[0] * m[0][1] + v[1] * m[1][1] + v[2] * m[2][1],
          v[0] * m[0][2] + v[1] * m[1][2] + v[2] * m[2][2]};
}
inline M3 Transpose(const M3 &m) {
  return {{{m[0][0], m[1][0], m[2][0]}, {m[0][1], m[1][1], m[2][1]}, {m[0][2], m[1][2], m[2][2]}}};
}
// Vector_d operator< with tolerance (VectorMatrix.hpp:78-88)
inline bool VecLess(const V3 &l, const V3 &r) {
  const double dx = l[0] - r[0];
  if (dx < -kEpsilon) return true;
  if (dx > kEpsilon) return false;
  const double dy = l[1] - r[1];
  if (dy < -kEpsilon) return true;
  if (dy > kEpsilon) return false;
  return l[2] < r[2] - kEpsilon;
}
// GetRelativeDistanceVectorLattice (lmc/cfg/include/Lattice.hpp:62-74)
inline V3 RelDist(const V3 &first, const V3 &second) {
  V3 d = second - first;
  for (int k = 0; k < 3; ++k) {
    while (d[k] >= 0.5) d[k] -= 1;
    while (d[k] < -0.5) d[k] += 1;
  }
  return d;
}

struct Site {
  size_t id;
  V3 rel;
};

// Tolerance comparators (lmc/cfg/include/LatticeCluster.hpp:62-154)
inline bool PositionCompareState(const Site &l, const Site &r) {
  const double dx = l.rel[0] - r.rel[0], dy = l.rel[1] - r.rel[1], dz = l.rel[2] - r.rel[2];
  if (dx < -kEpsilon) return true;
  if (dx > kEpsilon) return false;
  if (dy < -kEpsilon) return true;
  if (dy > kEpsilon) return false;
  return dz < -kEpsilon;
}
inline bool GroupCompareMMM(const Site &l, const Site &r) {
  const double dn = Inner(l.rel - 0.5) - Inner(r.rel - 0.5);
  if (dn < -kEpsilon) return true;
  if (dn > kEpsilon) return false;
  const double dxs = std::abs(l.rel[0] - 0.5) - std::abs(r.rel[0] - 0.5);
  return dxs < -kEpsilon;
}
inline bool GroupCompareMM2(const Site &l, const Site &r) {
  const double dn = Inner(l.rel - 0.5) - Inner(r.rel - 0.5);
  if (dn < -kEpsilon) return true;
  if (dn > kEpsilon) return false;
  return l.rel[0] - r.rel[0] < -kEpsilon;
}
inline bool PositionCompareMMM(const Site &l, const Site &r) {
  const double dn = Inner(l.rel - 0.5) - Inner(r.rel - 0.5);
  if (dn < -kEpsilon) return true;
  if (dn > kEpsilon) return false;
  const double dxs = std::abs(l.rel[0] - 0.5) - std::abs(r.rel[0] - 0.5);
  if (dxs < -kEpsilon) return true;
  if (dxs > kEpsilon) return false;
  for (int k = 0; k < 3; ++k) {
    const double d = l.rel[k] - r.rel[k];
    if (d < -kEpsilon) return true;
    if (d > kEpsilon) return false;
  }
  return false;
}
inline bool PositionCompareMM2(const Site &l, const Site &r) {
  const double dn = Inner(l.rel - 0.5) - Inner(r.rel - 0.5);
  if (dn < -kEpsilon) return true;
  if (dn > kEpsilon) return false;
  for (int k = 0; k < 3; ++k) {
    const double d = l.rel[k] - r.rel[k];
    if (d < -kEpsilon) return true;
    if (d > kEpsilon) return false;
  }
  return false;
}

// ------------------------------------------------------------------------------------------
// Lattice: GenerateFCC + UpdateNeighbors + ReassignLatticeVector (lmc/cfg/src/Config.cpp)
struct Lattice {
  int n = 0;
  size_t N = 0;
  double L = 0;  // box edge (Angstrom); basis = diag(L)
  std::vector<V3> rel;
  std::vector<std::vector<size_t>> nn[3];
};

static Lattice BuildLattice(int n) {
  Lattice lat;
  lat.n = n;
  lat.L = kLatticeConstant * n;
  const size_t N = 4ull * n * n * n;
  lat.N = N;
  // GenerateFCC (Config.cpp:1060-1095): id = running count over k, j, i, 4 basis sites.
  std::vector<V3> rel;
  rel.reserve(N);
  const double len = static_cast<double>(n);
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double x = i, y = j, z = k;
        rel.push_back({x / len, y / len, z / len});
        rel.push_back({(x + 0.5) / len, (y + 0.5) / len, z / len});
        rel.push_back({(x + 0.5) / len, y / len, (z + 0.5) / len});
        rel.push_back({x / len, (y + 0.5) / len, (z + 0.5) / len});
      }
  // UpdateNeighbors (Config.cpp:955-1045): cell list of edge >= kCut3, periodic images.
  std::vector<std::vector<size_t>> nn[3];
  for (auto &v : nn) v.assign(N, {});
  const size_t nc = static_cast<size_t>(std::floor(lat.L / kCut3));
  std::vector<std::vector<size_t>> cells(nc * nc * nc);
  for (size_t id = 0; id < N; ++id) {
    size_t c[3];
    for (int d = 0; d < 3; ++d) c[d] = static_cast<size_t>(std::floor(nc * rel[id][d]));
    cells.at((c[0] * nc + c[1]) * nc + c[2]).push_back(id);
  }
  const double c1 = kCut1 * kCut1, c2 = kCut2 * kCut2, c3 = kCut3 * kCut3;
  for (size_t ci = 0; ci < cells.size(); ++ci) {
    const size_t i = ci / (nc * nc), j = (ci % (nc * nc)) / nc, k = ci % nc;
    for (int di = -1; di <= 1; ++di)
      for (int dj = -1; dj <= 1; ++dj)
        for (int dk = -1; dk <= 1; ++dk) {
          const size_t ni = (nc + i + di) % nc, nj = (nc + j + dj) % nc, nk = (nc + k + dk) % nc;
          const auto &ncell = cells[(ni * nc + nj) * nc + nk];
          for (size_t id1 : cells[ci])
            for (size_t id2 : ncell) {
              if (id2 >= id1) continue;
              const V3 d = RelDist(rel[id1], rel[id2]) * lat.L;
              if (std::abs(d[0]) > kCut3 || std::abs(d[1]) > kCut3 || std::abs(d[2]) > kCut3)
                continue;
              const double r2 = Inner(d);
              int shell = -1;
              if (r2 < c1) shell = 0;
              else if (r2 < c2) shell = 1;
              else if (r2 < c3) shell = 2;
              if (shell < 0) continue;
              nn[shell][id1].push_back(id2);
              nn[shell][id2].push_back(id1);
            }
        }
  }
  // ReassignLatticeVector (Config.cpp:466-552): ids := rank of positions under VecLess.
  std::vector<size_t> order(N);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(),
            [&](size_t a, size_t b) { return VecLess(rel[a], rel[b]); });
  std::vector<size_t> old_to_new(N);
  for (size_t newid = 0; newid < N; ++newid) old_to_new[order[newid]] = newid;
  lat.rel.resize(N);
  for (size_t newid = 0; newid < N; ++newid) lat.rel[newid] = rel[order[newid]];
  for (int s = 0; s < 3; ++s) {
    lat.nn[s].assign(N, {});
    for (size_t old = 0; old < N; ++old)
      for (size_t o2 : nn[s][old]) lat.nn[s][old_to_new[old]].push_back(old_to_new[o2]);
    for (auto &v : lat.nn[s]) std::sort(v.begin(), v.end());
  }
  return lat;
}

// FindDistanceLabelBetweenLattice (Config.cpp:354-371)
static int DistanceLabel(const Lattice &lat, size_t a, size_t b) {
  for (int s = 0; s < 3; ++s) {
    const auto &v = lat.nn[s][a];
    if (std::find(v.begin(), v.end(), b) != v.end()) return s + 1;
  }
  return -1;
}
// GetLabel (lmc/pred/src/EnergyUtility.cpp:345-379)
static int GetLabel(const Lattice &lat, const std::vector<size_t> &ids) {
  if (ids.size() == 1) return 0;
  if (ids.size() == 2) return DistanceLabel(lat, ids[0], ids[1]);
  if (ids.size() == 3) {
    int b[3] = {DistanceLabel(lat, ids[0], ids[1]), DistanceLabel(lat, ids[1], ids[2]),
                DistanceLabel(lat, ids[2], ids[0])};
    std::sort(b, b + 3);
    if (b[0] == 1 && b[1] == 1 && b[2] == 1) return 4;
    if (b[0] == 1 && b[1] == 1 && b[2] == 2) return 5;
    if (b[0] == 1 && b[1] == 1 && b[2] == 3) return 6;
    if (b[0] == 1 && b[1] == 2 && b[2] == 3) return 7;
    if (b[0] == 1 && b[1] == 3 && b[2] == 3) return 8;
    if (b[0] == 2 && b[1] == 3 && b[2] == 3) return 9;
    if (b[0] == 3 && b[1] == 3 && b[2] == 3) return 10;
  }
  return -1;
}
// GetNeighborsLatticeIdSetOfPair / OfSite (Config.cpp:312-352)
static std::vector<size_t> NeighborSetOfSite(const Lattice &lat, size_t s) {
  std::set<size_t> h{s};
  for (int k = 0; k < 3; ++k) h.insert(lat.nn[k][s].begin(), lat.nn[k][s].end());
  return {h.begin(), h.end()};
}
static std::vector<size_t> NeighborSetOfPair(const Lattice &lat, size_t a, size_t b) {
  std::set<size_t> h;
  for (size_t s : {a, b}) {
    h.insert(s);
    for (int k = 0; k < 3; ++k) h.insert(lat.nn[k][s].begin(), lat.nn[k][s].end());
  }
  return {h.begin(), h.end()};
}
// GetLatticePairCenter (Config.cpp:228-245)
static V3 PairCenter(const Lattice &lat, size_t first, size_t second) {
  V3 c{};
  for (int k = 0; k < 3; ++k) {
    double f = lat.rel[first][k];
    const double s = lat.rel[second][k];
    double dist = f - s;
    int period = static_cast<int>(dist / 0.5);
    while (period != 0) {
      f -= static_cast<double>(period);
      dist = f - s;
      period = static_cast<int>(dist / 0.5);
    }
    c[k] = 0.5 * (f + s);
  }
  return c;
}
// GetLatticePairRotationMatrix (Config.cpp:247-264); cubic basis diag(L).
static M3 PairRotation(const Lattice &lat, size_t first, size_t second) {
  const V3 pd = Normalize(RelDist(lat.rel[first], lat.rel[second]) * lat.L);
  V3 vert{};
  for (size_t idx : lat.nn[0][first]) {
    const V3 jv = Normalize(RelDist(lat.rel[first], lat.rel[idx]) * lat.L);
    if (std::abs(Dot(pd, jv)) < kEpsilon) {
      vert = jv;
      break;
    }
  }
  return Transpose(M3{pd, vert, Cross(pd, vert)});
}
// RotateLatticeVector (Config.cpp:1047-1058)
static void Rotate(std::vector<Site> &v, const M3 &m) {
  const V3 half{0.5, 0.5, 0.5};
  const V3 move = half - MulVM(half, m);
  for (auto &s : v) {
    V3 p = MulVM(s.rel, m);
    p = p + move;
    p = p - Floor(p);
    s.rel = p;
  }
}
static std::vector<Site> MovedToCenter(const Lattice &lat, const std::vector<size_t> &ids,
                                       const V3 &center, size_t skip_a, size_t skip_b) {
  const V3 move = V3{0.5, 0.5, 0.5} - center;
  std::vector<Site> out;
  for (size_t id : ids) {
    V3 p = lat.rel[id] + move;
    p = p - Floor(p);
    if (id == skip_a || id == skip_b) continue;
    out.push_back({id, p});
  }
  return out;
}
// GetSortedLatticeVectorStateOfPair (EnergyUtility.cpp:261-287)
static std::vector<Site> SortedStateOfPair(const Lattice &lat, size_t a, size_t b) {
  auto v = MovedToCenter(lat, NeighborSetOfPair(lat, a, b), PairCenter(lat, a, b), kSizeMax,
                         kSizeMax);
  Rotate(v, PairRotation(lat, a, b));
  std::sort(v.begin(), v.end(), PositionCompareState);
  return v;
}
// GetSymmetricallySortedLatticeVectorMMM/MM2 (EnergyUtility.cpp:45-104)
static std::vector<Site> SortedSymOfPair(const Lattice &lat, size_t a, size_t b, bool mm2) {
  auto v = MovedToCenter(lat, NeighborSetOfPair(lat, a, b), PairCenter(lat, a, b), a, b);
  Rotate(v, PairRotation(lat, a, b));
  std::sort(v.begin(), v.end(), mm2 ? PositionCompareMM2 : PositionCompareMMM);
  return v;
}
// GetSortedLatticeVectorStateOfSite (EnergyUtility.cpp:288-313): no rotation.
static std::vector<Site> SortedStateOfSite(const Lattice &lat, size_t s) {
  auto v = MovedToCenter(lat, NeighborSetOfSite(lat, s), lat.rel[s], kSizeMax, kSizeMax);
  std::sort(v.begin(), v.end(), PositionCompareState);
  return v;
}

using Mapping = std::vector<std::vector<std::vector<size_t>>>;

// GetClusterParametersMappingStatePair / StateSite (EnergyUtility.cpp:393-581): clusters (as
// indices into the sorted list) that contain one of the centre sites, grouped by label 0..7.
static Mapping StateMapping(const Lattice &lat, const std::vector<Site> &sorted,
                            const std::vector<size_t> &centre) {
  auto is_c = [&](size_t id) {
    return std::find(centre.begin(), centre.end(), id) != centre.end();
  };
  Mapping m(8);
  for (size_t i1 = 0; i1 < sorted.size(); ++i1) {
    const size_t id1 = sorted[i1].id;
    if (is_c(id1)) m[0].push_back({i1});
    for (size_t i2 = 0; i2 < i1; ++i2) {
      const size_t id2 = sorted[i2].id;
      if (is_c(id1) || is_c(id2)) {
        const int lab = GetLabel(lat, {id1, id2});
        if (lab < 1 || lab > 3) continue;  // `default: continue` skips the triplet loop
        m[lab].push_back({i1, i2});
      }
      for (size_t i3 = 0; i3 < i2; ++i3) {
        const size_t id3 = sorted[i3].id;
        if (is_c(id1) || is_c(id2) || is_c(id3)) {
          const int lab = GetLabel(lat, {id1, id2, id3});
          if (lab >= 4 && lab <= 7) m[lab].push_back({i1, i2, i3});
        }
      }
    }
  }
  return m;
}

// LatticeClusterMMM/MM2 (LatticeCluster.hpp:155-209) and the orbit grouping of
// GetAverageParametersMappingFromLatticeClusterVectorHelperMMM/MM2 (EnergyUtility.cpp:106-167).
struct SymCluster {
  std::vector<Site> s;
  bool sym = false;
};
static SymCluster MakeSymCluster(std::vector<Site> s, bool mm2) {
  std::sort(s.begin(), s.end(), mm2 ? PositionCompareMM2 : PositionCompareMMM);
  SymCluster c{s, false};
  auto gc = mm2 ? GroupCompareMM2 : GroupCompareMMM;
  for (size_t a = 0; a < s.size(); ++a)
    for (size_t b = 0; b < a; ++b)
      if (!gc(s[a], s[b]) && !gc(s[b], s[a])) c.sym = true;
  return c;
}
static bool SymLess(const SymCluster &l, const SymCluster &r, bool mm2) {
  auto gc = mm2 ? GroupCompareMM2 : GroupCompareMMM;
  for (size_t i = 0; i < l.s.size(); ++i) {
    if (gc(l.s[i], r.s[i])) return true;
    if (gc(r.s[i], l.s[i])) return false;
  }
  return false;
}
static void GroupOrbits(std::vector<SymCluster> v, bool mm2, Mapping &out) {
  std::sort(v.begin(), v.end(),
            [mm2](const SymCluster &a, const SymCluster &b) { return SymLess(a, b, mm2); });
  size_t lo = 0;
  while (lo < v.size()) {
    size_t hi = lo + 1;
    while (hi < v.size() && !SymLess(v[lo], v[hi], mm2)) ++hi;
    std::vector<std::vector<size_t>> orbit;
    for (size_t i = lo; i < hi; ++i) {
      std::vector<size_t> idx;
      if (v[i].sym) idx.push_back(kSizeMax);
      for (auto &s : v[i].s) idx.push_back(s.id);
      orbit.push_back(idx);
    }
    out.push_back(orbit);
    lo = hi;
  }
}
// GetAverageClusterParametersMappingMMM/MM2 (EnergyUtility.cpp:169-259)
static Mapping AverageMapping(const Lattice &lat, const std::vector<Site> &sorted, bool mm2) {
  std::vector<SymCluster> singles, pairs[3];
  for (size_t i1 = 0; i1 < sorted.size(); ++i1) {
    Site l1{i1, sorted[i1].rel};
    singles.push_back(MakeSymCluster({l1}, mm2));
    for (size_t i2 = 0; i2 < i1; ++i2) {
      Site l2{i2, sorted[i2].rel};
      const int lab = DistanceLabel(lat, sorted[i1].id, sorted[i2].id);
      if (lab >= 1 && lab <= 3) pairs[lab - 1].push_back(MakeSymCluster({l1, l2}, mm2));
    }
  }
  Mapping m;
  GroupOrbits(singles, mm2, m);
  for (auto &p : pairs) GroupOrbits(p, mm2, m);
  return m;
}

// ------------------------------------------------------------------------------------------
// Element clusters and the ECI key space (lmc/cfg/include/ElementCluster.hpp:40-50,
// EnergyUtility.cpp:314-343, :798-819)
struct EC {
  int label;
  std::vector<EN> e;  // sorted by string
  EC(int l, std::vector<EN> v) : label(l), e(std::move(v)) {
    std::sort(e.begin(), e.end(), ElemLess());
  }
};
struct ECLess {
  bool operator()(const EC &l, const EC &r) const {
    if (l.e.size() != r.e.size()) return l.e.size() < r.e.size();
    if (l.label != r.label) return l.label < r.label;
    for (size_t i = 0; i < l.e.size(); ++i) {
      if (ElemLess()(l.e[i], r.e[i])) return true;
      if (ElemLess()(r.e[i], l.e[i])) return false;
    }
    return false;
  }
};
static std::map<EC, int, ECLess> EciKeySpace(const std::vector<EN> &with_x) {
  std::map<EC, int, ECLess> m;
  for (EN e1 : with_x) {
    m[EC(0, {e1})] = 0;
    for (EN e2 : with_x) {
      if (e2 == EN::X) continue;
      for (int l = 1; l <= 3; ++l) m[EC(l, {e1, e2})] = 0;
      for (EN e3 : with_x) {
        if (e3 == EN::X) continue;
        for (int l = 4; l < 8; ++l) m[EC(l, {e1, e2, e3})] = 0;
      }
    }
  }
  int i = 0;
  for (auto &kv : m) kv.second = i++;
  return m;
}

// ------------------------------------------------------------------------------------------
// Minimal JSON reader (the reference parses with nlohmann::json, EnergyUtility.h:15-33).
struct Json {
  enum T { Null, Num, Str, Arr, Obj } t = Null;
  double num = 0;
  std::string str;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;
  const Json &at(const std::string &k) const {
    for (auto &kv : obj)
      if (kv.first == k) return kv.second;
    throw std::runtime_error("missing key " + k);
  }
  std::vector<double> vec() const {
    std::vector<double> v;
    for (auto &a : arr) v.push_back(a.num);
    return v;
  }
};
struct JsonParser {
  const char *p;
  void ws() {
    while (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r') ++p;
  }
  Json parse() {
    ws();
    Json j;
    if (*p == '{') {
      j.t = Json::Obj;
      ++p;
      ws();
      if (*p == '}') { ++p; return j; }
      while (true) {
        ws();
        std::string k = parse().str;
        ws();
        if (*p++ != ':') throw std::runtime_error("json: expected ':'");
        j.obj.emplace_back(k, parse());
        ws();
        if (*p == ',') { ++p; continue; }
        if (*p == '}') { ++p; break; }
        throw std::runtime_error("json: bad object");
      }
    } else if (*p == '[') {
      j.t = Json::Arr;
      ++p;
      ws();
      if (*p == ']') { ++p; return j; }
      while (true) {
        j.arr.push_back(parse());
        ws();
        if (*p == ',') { ++p; continue; }
        if (*p == ']') { ++p; break; }
        throw std::runtime_error("json: bad array");
      }
    } else if (*p == '"') {
      j.t = Json::Str;
      ++p;
      while (*p && *p != '"') j.str += *p++;
      if (*p++ != '"') throw std::runtime_error("json: bad string");
    } else if (!std::strncmp(p, "null", 4)) {
      p += 4;
    } else {
      char *end = nullptr;
      j.t = Json::Num;
      j.num = std::strtod(p, &end);
      if (end == p) throw std::runtime_error("json: bad value");
      p = end;
    }
    return j;
  }
};

// ------------------------------------------------------------------------------------------
// Model: the predictor classes of lmc/pred/.
struct ParamsBarrier {  // ParametersE0 / ParametersQuartic (EnergyUtility.h:78-101)
  std::vector<double> mu_mmm, sigma_mmm, mu_mm2, sigma_mm2;
  std::vector<std::vector<double>> U_mmm, U_mm2;
  std::vector<double> theta_a, theta_b;  // E0: theta_e0 | Quartic: theta_D, theta_Ks
  double mu_a = 0, sigma_a = 0, mu_b = 0, sigma_b = 0;
};

struct Model {
  int kind = 0;
  const Lattice *lat = nullptr;
  std::vector<EN> elements;  // element_set_ (no X), string order
  std::vector<EN> species;   // element_set_ + X, string order: occupancy codes
  std::vector<double> theta;
  std::map<EC, int, ECLess> eci;
  std::vector<double> total_bonds;  // per ECI index
  std::map<EN, ParamsBarrier> params;
  std::unordered_map<std::string, std::vector<double>> one_hot;
  Mapping map_state, map_mmm, map_mm2, map_site;
  // Per directed 1NN pair tables (site*12 + slot), built lazily: 60 state ids, 58 MMM, 58 MM2.
  mutable std::vector<std::vector<size_t>> t_state, t_mmm, t_mm2;
  mutable std::vector<std::vector<size_t>> t_site;
  mutable std::mutex mu;

  int EciIndex(int label, const std::vector<EN> &v) const {
    auto it = eci.find(EC(label, v));
    if (it == eci.end()) throw std::out_of_range("Cluster not found in ClusterIndexer");
    return it->second;
  }
  size_t Flat(size_t a, size_t b) const {  // GetPairFlatIndex (...Quartic.cpp:278-285)
    const auto &v = lat->nn[0][a];
    auto it = std::find(v.begin(), v.end(), b);
    if (it == v.end()) throw std::out_of_range("Neighbor not found for lattice pair");
    return a * 12 + static_cast<size_t>(it - v.begin());
  }
  void EnsurePair(size_t a, size_t b) const {
    const size_t f = Flat(a, b);
    std::lock_guard<std::mutex> g(mu);
    if (!t_state[f].empty()) return;
    for (auto &s : SortedStateOfPair(*lat, a, b)) t_state[f].push_back(s.id);
    for (auto &s : SortedSymOfPair(*lat, a, b, false)) t_mmm[f].push_back(s.id);
    for (auto &s : SortedSymOfPair(*lat, a, b, true)) t_mm2[f].push_back(s.id);
  }
  const std::vector<size_t> &Site43(size_t s) const {
    if (sites_ready) return t_site[s];  // every table built: read without the lock
    std::lock_guard<std::mutex> g(mu);
    if (t_site[s].empty())
      for (auto &x : SortedStateOfSite(*lat, s)) t_site[s].push_back(x.id);
    return t_site[s];
  }
  // Builds every site table up front (before concurrent runs).
  void PrepareSites() const {
    for (size_t s = 0; s < lat->N; ++s) (void)Site43(s);
    sites_ready = true;
  }
  mutable std::atomic<bool> sites_ready{false};
};

static double DotSeq(const std::vector<double> &a, const std::vector<double> &b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// GetOneHotEncodeHashmap (EnergyUtility.cpp:7-43)
static std::unordered_map<std::string, std::vector<double>> OneHotMap(const std::vector<EN> &es) {
  std::unordered_map<std::string, std::vector<double>> m;
  const size_t n = es.size();
  for (size_t i = 0; i < n; ++i) {
    std::vector<double> v(n, 0);
    v[i] = 1;
    m[Str(es[i])] = v;
  }
  size_t c = 0;
  for (EN a : es)
    for (EN b : es) {
      std::vector<double> v(n * n, 0);
      v[c++] = 1;
      m[Str(a) + Str(b)] = v;
    }
  c = 0;
  for (size_t i = 0; i < n; ++i)
    for (size_t j = i; j < n; ++j) {
      std::vector<double> v(n * (n + 1) / 2, 0);
      v[c++] = 1;
      m[Str(es[i]) + '-' + Str(es[j])] = v;
    }
  return m;
}

// GetOneHotParametersFromMap (EnergyUtility.cpp:743-796)
static std::vector<double> OneHotEncode(const Model &m, const std::vector<EN> &enc,
                                        const Mapping &mapping) {
  std::vector<double> out;
  const size_t ne = m.elements.size();
  for (const auto &orbit : mapping) {
    size_t len = orbit[0][0] == kSizeMax ? ne * (ne + 1) / 2
                                        : static_cast<size_t>(std::pow(ne, orbit[0].size()));
    std::vector<double> sum(len, 0.0);
    for (const auto &cl : orbit) {
      std::string key;
      if (cl[0] == kSizeMax) {
        std::vector<std::string> sv;
        for (size_t i = 1; i < cl.size(); ++i) sv.push_back(Str(enc[cl[i]]));
        std::sort(sv.begin(), sv.end());
        key = sv[0];
        for (size_t i = 1; i < sv.size(); ++i) key += '-' + sv[i];
      } else {
        for (size_t idx : cl) key += Str(enc[idx]);
      }
      const auto &oh = m.one_hot.at(key);
      for (size_t i = 0; i < len; ++i) sum[i] += oh[i];
    }
    const double cnt = static_cast<double>(orbit.size());
    for (auto &x : sum) x /= cnt;
    out.insert(out.end(), sum.begin(), sum.end());
  }
  return out;
}

static EN ElemAt(const Model &m, const uint8_t *occ, size_t site) { return m.species[occ[site]]; }

// GetDe (VacancyMigrationPredictorQuartic.cpp:112-166 == ...E0.cpp:77-126)
static double GetDe(const Model &m, const uint8_t *occ, size_t a, size_t b) {
  m.EnsurePair(a, b);
  const EN mig = ElemAt(m, occ, b);
  const auto &ids = m.t_state[m.Flat(a, b)];
  std::vector<int> start(m.eci.size(), 0), end(m.eci.size(), 0);
  for (size_t label = 0; label < m.map_state.size(); ++label)
    for (const auto &cl : m.map_state[label]) {
      std::vector<EN> vs, ve;
      for (size_t idx : cl) {
        const size_t id = ids[idx];
        const EN e = ElemAt(m, occ, id);
        vs.push_back(e);
        if (id == a) ve.push_back(mig);
        else if (id == b) ve.push_back(EN::X);
        else ve.push_back(e);
      }
      start[m.EciIndex(static_cast<int>(label), vs)]++;
      end[m.EciIndex(static_cast<int>(label), ve)]++;
    }
  std::vector<double> enc(m.eci.size());
  for (size_t i = 0; i < enc.size(); ++i)
    enc[i] = (static_cast<double>(end[i]) - static_cast<double>(start[i])) / m.total_bonds[i];
  return DotSeq(m.theta, enc);
}

static std::vector<EN> Gather(const Model &m, const uint8_t *occ, const std::vector<size_t> &ids) {
  std::vector<EN> v;
  for (size_t id : ids) v.push_back(ElemAt(m, occ, id));
  return v;
}

// (U * ((x - mu) / sigma)) . theta, then * sigma_y + mu_y, exp (...E0.cpp:142-150,
// ...Quartic.cpp:204-214, :237-245)
static double LogLinearExp(std::vector<double> x, const std::vector<double> &mu,
                           const std::vector<double> &sigma,
                           const std::vector<std::vector<double>> &U,
                           const std::vector<double> &theta, double sig_y, double mu_y) {
  for (size_t i = 0; i < x.size(); ++i) x[i] = (x[i] - mu.at(i)) / sigma.at(i);
  std::vector<double> ux(U.size());
  for (size_t r = 0; r < U.size(); ++r) ux[r] = DotSeq(U[r], x);
  double l = DotSeq(ux, theta);
  l *= sig_y;
  l += mu_y;
  return std::exp(l);
}

static void BarrierDe(const Model &m, const uint8_t *occ, size_t a, size_t b, double &ea,
                      double &de) {
  m.EnsurePair(a, b);
  const EN mig = ElemAt(m, occ, b);
  de = GetDe(m, occ, a, b);
  const auto &p = m.params.at(mig);
  const size_t f = m.Flat(a, b);
  const auto xm = OneHotEncode(m, Gather(m, occ, m.t_mmm[f]), m.map_mmm);
  if (m.kind == ORC_E0) {
    // VacancyMigrationPredictorE0::GetE0 / GetBarrierAndDiffFromLatticeIdPair (:128-159)
    const double e0 = LogLinearExp(xm, p.mu_mmm, p.sigma_mmm, p.U_mmm, p.theta_a, p.sigma_a,
                                   p.mu_a);
    ea = std::max(0.0, e0 + de / 2);
    return;
  }
  // VacancyMigrationPredictorQuartic::GetD / GetKs / GetBarrier... (:167-276)
  const double D = LogLinearExp(xm, p.mu_mmm, p.sigma_mmm, p.U_mmm, p.theta_a, p.sigma_a, p.mu_a);
  m.EnsurePair(b, a);
  auto xf = OneHotEncode(m, Gather(m, occ, m.t_mm2[f]), m.map_mm2);
  const auto xb = OneHotEncode(m, Gather(m, occ, m.t_mm2[m.Flat(b, a)]), m.map_mm2);
  for (size_t i = 0; i < xf.size(); ++i) xf[i] += xb[i];
  const double Ks = LogLinearExp(xf, p.mu_mm2, p.sigma_mm2, p.U_mm2, p.theta_b, p.sigma_b, p.mu_b);
  const double bq = 4 * de / (D * D * D);
  const double aq = Ks / (4 * D * D);
  const double c = (9 * bq * bq - 16 * aq * aq * D * D) / (32 * aq);
  const double delta = std::sqrt(std::abs(9 * bq * bq - 32 * aq * c));
  ea = (3 * bq + delta) * (3 * bq + delta) * (3 * bq * bq - 16 * aq * c + bq * delta) /
       std::pow(aq, 3) / 2048;
}

// EnergyChangePredictorPairSite (lmc/pred/src/EnergyChangePredictorPairSite.cpp)
static double DeHelper(const Model &m, const std::vector<int> &s, const std::vector<int> &e) {
  std::vector<double> enc(m.eci.size());
  for (size_t i = 0; i < enc.size(); ++i)
    enc[i] = (static_cast<double>(e[i]) - static_cast<double>(s[i])) / m.total_bonds[i];
  return DotSeq(m.theta, enc);
}
static double SiteDe(const Model &m, const uint8_t *occ, size_t site, EN ne) {
  const EN old = ElemAt(m, occ, site);
  if (old == ne) return 0.0;
  const auto &ids = m.Site43(site);
  std::vector<int> s(m.eci.size(), 0), e(m.eci.size(), 0);
  for (size_t label = 0; label < m.map_site.size(); ++label)
    for (const auto &cl : m.map_site[label]) {
      std::vector<EN> vs, ve;
      for (size_t idx : cl) {
        const size_t id = ids[idx];
        vs.push_back(ElemAt(m, occ, id));
        ve.push_back(id == site ? ne : ElemAt(m, occ, id));
      }
      s[m.EciIndex(static_cast<int>(label), vs)]++;
      e[m.EciIndex(static_cast<int>(label), ve)]++;
    }
  return DeHelper(m, s, e);
}
static double SwapDe(const Model &m, const uint8_t *occ, size_t p, size_t q) {
  const EN ep = ElemAt(m, occ, p), eq = ElemAt(m, occ, q);
  if (ep == eq) return 0.0;
  const auto near = NeighborSetOfSite(*m.lat, p);
  if (std::find(near.begin(), near.end(), q) == near.end())
    return SiteDe(m, occ, p, eq) + SiteDe(m, occ, q, ep);  // ...WithoutCoupling (:136-146)
  // ...WithCoupling (:96-134) via GetClusterParametersMappingStatePairOf (EnergyUtility.cpp:582)
  const auto ids = NeighborSetOfPair(*m.lat, p, q);
  std::vector<Site> sites;
  for (size_t id : ids) sites.push_back({id, m.lat->rel[id]});
  const Mapping map = StateMapping(*m.lat, sites, {p, q});
  std::vector<int> s(m.eci.size(), 0), e(m.eci.size(), 0);
  for (size_t label = 0; label < map.size(); ++label)
    for (const auto &cl : map[label]) {
      std::vector<EN> vs, ve;
      for (size_t idx : cl) {
        const size_t id = sites[idx].id;
        vs.push_back(ElemAt(m, occ, id));
        ve.push_back(id == p ? eq : id == q ? ep : ElemAt(m, occ, id));
      }
      s[m.EciIndex(static_cast<int>(label), vs)]++;
      e[m.EciIndex(static_cast<int>(label), ve)]++;
    }
  return DeHelper(m, s, e);
}

// EnergyPredictor::GetEncode (lmc/pred/src/EnergyPredictor.cpp:40-96), normalisers :8
static std::vector<double> TotalEncode(const Model &m, const uint8_t *occ) {
  static const double kCounter[] = {256, 3072, 1536, 6144, 12288, 6144, 12288, 6144};
  const Lattice &L = *m.lat;
  std::vector<long long> cnt(m.eci.size(), 0);
  auto has = [&](int shell, size_t a, size_t b) {
    const auto &v = L.nn[shell][a];
    return std::find(v.begin(), v.end(), b) != v.end();
  };
  for (size_t s1 = 0; s1 < L.N; ++s1) {
    const EN e1 = ElemAt(m, occ, s1);
    cnt[m.EciIndex(0, {e1})]++;
    for (size_t s2 : L.nn[0][s1]) {
      const EN e2 = ElemAt(m, occ, s2);
      cnt[m.EciIndex(1, {e1, e2})]++;
      for (size_t s3 : L.nn[0][s2]) {
        const EN e3 = ElemAt(m, occ, s3);
        if (has(0, s1, s3)) cnt[m.EciIndex(4, {e1, e2, e3})]++;
        else if (has(1, s1, s3)) cnt[m.EciIndex(5, {e1, e2, e3})]++;
        else if (has(2, s1, s3)) cnt[m.EciIndex(6, {e1, e2, e3})]++;
      }
      for (size_t s3 : L.nn[1][s2]) {
        const EN e3 = ElemAt(m, occ, s3);
        if (has(2, s1, s3)) cnt[m.EciIndex(7, {e1, e2, e3})]++;
      }
    }
    for (size_t s2 : L.nn[1][s1]) cnt[m.EciIndex(2, {e1, ElemAt(m, occ, s2)})]++;
    for (size_t s2 : L.nn[2][s1]) cnt[m.EciIndex(3, {e1, ElemAt(m, occ, s2)})]++;
  }
  std::vector<double> enc(m.eci.size());
  size_t i = 0;
  for (auto &kv : m.eci) {
    enc[i] = static_cast<double>(cnt[i]) / kCounter[kv.first.label];
    ++i;
  }
  return enc;
}

static Model *LoadModel(const Lattice *lat, const std::string &path, const std::string &csv,
                        int kind) {
  auto m = std::make_unique<Model>();
  m->kind = kind;
  m->lat = lat;
  std::set<EN, ElemLess> es;
  std::stringstream ss(csv);
  std::string tok;
  while (std::getline(ss, tok, ','))
    if (!tok.empty()) es.insert(FromStr(tok));
  m->elements.assign(es.begin(), es.end());
  es.insert(EN::X);
  m->species.assign(es.begin(), es.end());
  m->eci = EciKeySpace(m->species);
  // ΔE normalisers (...Quartic.cpp:13-14)
  static const double kDe[] = {256, 1536, 768, 3072, 2048, 3072, 6144, 6144};
  for (auto &kv : m->eci) m->total_bonds.push_back(kDe[kv.first.label]);
  m->one_hot = OneHotMap(m->elements);

  std::ifstream ifs(path);
  if (!ifs) throw std::runtime_error("Cannot open " + path);
  std::stringstream buf;
  buf << ifs.rdbuf();
  const std::string text = buf.str();
  JsonParser jp{text.c_str()};
  const Json root = jp.parse();
  for (auto &kv : root.obj) {
    if (kv.first == "Base") {
      m->theta = kv.second.at("theta").vec();
      continue;
    }
    if (kind != ORC_E0 && kind != ORC_QUARTIC) continue;
    const Json &j = kv.second;
    ParamsBarrier p;
    auto mat = [](const Json &a) {
      std::vector<std::vector<double>> r;
      for (auto &row : a.arr) r.push_back(row.vec());
      return r;
    };
    p.mu_mmm = j.at("mu_x_mmm").vec();
    p.sigma_mmm = j.at("sigma_x_mmm").vec();
    p.U_mmm = mat(j.at("U_mmm"));
    if (kind == ORC_E0) {
      p.theta_a = j.at("theta_e0").vec();
      p.mu_a = j.at("mu_e0").num;
      p.sigma_a = j.at("sigma_e0").num;
    } else {
      p.mu_mm2 = j.at("mu_x_mm2").vec();
      p.sigma_mm2 = j.at("sigma_x_mm2").vec();
      p.U_mm2 = mat(j.at("U_mm2"));
      p.theta_a = j.at("theta_D").vec();
      p.theta_b = j.at("theta_Ks").vec();
      p.mu_a = j.at("mu_D").num;
      p.mu_b = j.at("mu_Ks").num;
      p.sigma_a = j.at("sigma_D").num;
      p.sigma_b = j.at("sigma_Ks").num;
    }
    m->params[FromStr(kv.first)] = p;
  }
  if (m->theta.size() != m->eci.size())
    throw std::runtime_error("Base.theta size does not match the cluster key space");
  // Mappings from the reference pair {0, nn1(0)[0]} and site 0 (EnergyUtility.cpp:169-259,
  // :393-581).
  const size_t a0 = 0, b0 = lat->nn[0][0][0];
  const auto st = SortedStateOfPair(*lat, a0, b0);
  m->map_state = StateMapping(*lat, st, {a0, b0});
  m->map_mmm = AverageMapping(*lat, SortedSymOfPair(*lat, a0, b0, false), false);
  m->map_mm2 = AverageMapping(*lat, SortedSymOfPair(*lat, a0, b0, true), true);
  const auto s43 = SortedStateOfSite(*lat, 0);
  m->map_site = StateMapping(*lat, s43, {0});
  m->t_state.assign(lat->N * 12, {});
  m->t_mmm.assign(lat->N * 12, {});
  m->t_mm2.assign(lat->N * 12, {});
  m->t_site.assign(lat->N, {});
  return m.release();
}

static size_t FeatureCount(const Model &m, const Mapping &map) {
  const size_t ne = m.elements.size();
  size_t n = 0;
  for (auto &o : map)
    n += o[0][0] == kSizeMax ? ne * (ne + 1) / 2 : static_cast<size_t>(std::pow(ne, o[0].size()));
  return n;
}

// ------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11): 10 rounds, key schedule Weyl increments.
static void Philox(uint32_t k0, uint32_t k1, const uint32_t in[4], uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
    const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
    const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}
static void Draw(uint64_t seed, uint64_t ctr, uint32_t replica, uint32_t tag, uint32_t out[4]) {
  const uint32_t in[4] = {static_cast<uint32_t>(ctr), static_cast<uint32_t>(ctr >> 32), replica,
                          tag};
  Philox(static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32), in, out);
}
static inline uint64_t Bits64(uint32_t lo, uint32_t hi) {
  return static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32);
}
static inline double UnitOpenLeft(uint64_t b) {  // (0, 1]
  return static_cast<double>((b >> 11) + 1) * 0x1.0p-53;
}
static inline double UnitOpenRight(uint64_t b) {  // [0, 1)
  return static_cast<double>(b >> 11) * 0x1.0p-53;
}
static inline uint64_t Below(uint64_t b, uint64_t n) {  // floor(b * n / 2^64)
  return static_cast<uint64_t>((static_cast<unsigned __int128>(b) * n) >> 64);
}

// ------------------------------------------------------------------------------------------
// LruCache (lmc/pred/include/LruCache.hpp:8-50) keyed by a hash of the environment species
// (VacancyMigrationPredictorQuarticLru.cpp:36-48, boost::hash_combine).
struct Lru {
  size_t cap;
  std::list<std::pair<size_t, std::pair<double, double>>> lst;
  std::unordered_map<size_t, decltype(lst.begin())> map;
  std::mutex mu;
  uint64_t hits = 0;
  explicit Lru(size_t c) : cap(c) {}
  bool Get(size_t k, std::pair<double, double> &v) {
    std::lock_guard<std::mutex> g(mu);
    auto it = map.find(k);
    if (it == map.end()) return false;
    lst.splice(lst.begin(), lst, it->second);
    v = it->second->second;
    ++hits;
    return true;
  }
  void Add(size_t k, const std::pair<double, double> &v) {
    std::lock_guard<std::mutex> g(mu);
    auto it = map.find(k);
    if (it != map.end()) lst.erase(it->second);
    lst.push_front({k, v});
    map[k] = lst.begin();
    if (map.size() > cap) {
      map.erase(lst.back().first);
      lst.pop_back();
    }
  }
};
static inline void HashCombine(size_t &seed, size_t v) {
  seed ^= v + 0x9e3779b9 + (seed << 6) + (seed >> 2);
}

// ------------------------------------------------------------------------------------------
// KineticMcFirst driver
struct Kmc {
  const Model *bm;
  const Lattice *lat;
  std::vector<uint8_t> occ;
  double T, beta, time, energy, abs_energy;
  uint64_t steps;
  uint64_t seed;
  uint32_t replica;
  size_t vac;
  int64_t disp[3] = {0, 0, 0};
  double com[3] = {0, 0, 0};
  bool escaped = false;
  bool rate_corrector, early_stop, solute_disp;
  double cv = 0, cs = 0;
  EN solvent = EN::Al;
  double solute_mass = 0;
  std::vector<std::pair<double, double>> tt;
  std::unique_ptr<Lru> lru;
  int threads = 1;
  // current events
  int32_t ev_nb[12];
  double ev_ea[12], ev_de[12], ev_rate[12], ev_cum[12];
  double total_rate = 0;
  // second-order (chain) steps: per first neighbour i of the vacancy k, the event k -> i as the
  // reverse of i -> k after the jump, and the total rate out of i
  bool chain = false;
  size_t prev_j = 0;  // the site the vacancy came from (KineticMcChainAbstract, :252)
  double ch_barrier[12], ch_de[12], ch_rate[12], ch_back[12], ch_total_i[12];

  // TimeTemperatureInterpolator::GetTemperature (TimeTemperatureInterpolator.cpp:44-65)
  double TAt(double t) const {
    auto it = std::lower_bound(tt.begin(), tt.end(), t,
                               [](const std::pair<double, double> &p, double x) { return p.first < x; });
    if (it == tt.end()) return (tt.end() - 1)->second;
    if (it == tt.begin() && t <= tt.begin()->first) return tt.begin()->second;
    const double ux = it->first, uy = it->second, lx = (it - 1)->first, ly = (it - 1)->second;
    return ly + ((t - lx) / (ux - lx)) * (uy - ly);
  }
  std::pair<double, double> Predict(size_t a, size_t b) {
    if (lru) {
      bm->EnsurePair(a, b);
      const size_t f = bm->Flat(a, b);
      size_t key = 0;
      for (size_t id : bm->t_state[f]) HashCombine(key, static_cast<size_t>(bm->species[occ[id]]));
      for (size_t id : bm->t_mmm[f]) HashCombine(key, static_cast<size_t>(bm->species[occ[id]]));
      for (size_t id : bm->t_mm2[f]) HashCombine(key, static_cast<size_t>(bm->species[occ[id]]));
      std::pair<double, double> v;
      if (lru->Get(key, v)) return v;
      BarrierDe(*bm, occ.data(), a, b, v.first, v.second);
      lru->Add(key, v);
      return v;
    }
    std::pair<double, double> v;
    BarrierDe(*bm, occ.data(), a, b, v.first, v.second);
    return v;
  }
  // BuildEventList (KineticMcFirstOmp.cpp:52-78) + JumpEvent (JumpEvent.cpp:6-13,51-57)
  void BuildEvents() {
    const auto &nb = lat->nn[0][vac];
#pragma omp parallel for num_threads(threads) schedule(static) if (threads > 1)
    for (int i = 0; i < 12; ++i) {
      auto r = Predict(vac, nb[i]);
      ev_nb[i] = static_cast<int32_t>(nb[i]);
      ev_ea[i] = r.first;
      ev_de[i] = r.second;
      ev_rate[i] = std::exp(-r.first * beta);
    }
    total_rate = 0.0;
    for (int i = 0; i < 12; ++i) total_rate += ev_rate[i];
    double cum = 0.0;
    for (int i = 0; i < 12; ++i) {
      cum += ev_rate[i] / total_rate;
      ev_cum[i] = cum;
    }
  }
  // KineticMcChainOmpi::BuildEventList (KineticMcChainOmpi.cpp:56-91), one pass per first
  // neighbour i of the vacancy k (the reference runs one MPI rank per i); the rank sums are
  // taken in neighbour order.
  void BuildChainEvents() {
    const size_t k = vac;
    const auto &nbk = lat->nn[0][k];
    for (int i = 0; i < 12; ++i) {
      const size_t is = nbk[i];
      std::swap(occ[k], occ[is]);  // LatticeJump({k, i})
      const auto &nbi = lat->nn[0][is];
      double tot = 0.0;
      for (int l = 0; l < 12; ++l) {
        const auto r = Predict(is, nbi[l]);
        if (nbi[l] == k) {  // GetReverseJumpEvent (JumpEvent.cpp:51-54)
          ch_barrier[i] = r.first - r.second;
          ch_de[i] = -r.second;
        }
        tot += std::exp(-r.first * beta);
      }
      std::swap(occ[k], occ[is]);  // LatticeJump({i, k})
      ch_total_i[i] = tot;
      ch_rate[i] = std::exp(-ch_barrier[i] * beta);
      ch_back[i] = std::exp((ch_de[i] - ch_barrier[i]) * beta);  // GetBackwardRate
    }
    total_rate = 0.0;
    for (int i = 0; i < 12; ++i) total_rate += ch_rate[i];  // MPI_Allreduce
  }
  // KineticMcChainOmpi::CalculateTime (KineticMcChainOmpi.cpp:93-151): second-order time and
  // event probabilities (cumulative in ev_cum).
  double ChainTime() {
    const auto &nbk = lat->nn[0][vac];
    const double t1 = 1 / total_rate / kPrefactor;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0}, b[12];
    bool prev[12];
    for (int i = 0; i < 12; ++i) {
      const double p_ki = ch_rate[i] / total_rate;
      const double p_ik = ch_back[i] / ch_total_i[i];
      const double bb = p_ki * p_ik;
      b[i] = p_ki * (1 - p_ik);
      prev[i] = nbk[i] == prev_j;
      const double ti = 1 / ch_total_i[i] / kPrefactor;
      s[0] += bb;
      s[1] += b[i];
      s[2] += prev[i] ? 0.0 : bb;
      s[3] += prev[i] ? 0.0 : b[i];
      s[4] += prev[i] ? b[i] : 0.0;
      s[5] += prev[i] ? p_ki : 0.0;
      s[6] += (t1 + ti) * bb;
      s[7] += prev[i] ? 0.0 : (t1 + ti) * bb;
    }
    const double ts = s[6] / s[0], ts_j = s[7] / s[2];
    const double inv = 1 / (1 - s[5]);
    const double t2 = inv * (s[3] * t1 + s[2] * (ts_j + t1 + s[0] / s[1] * ts));
    double cum = 0.0;
    for (int i = 0; i < 12; ++i) {
      const double p = prev[i] ? inv * (s[2] / s[1]) * s[4] : inv * (1 + s[2] / s[1]) * b[i];
      cum += p;
      ev_cum[i] = cum;
      ev_nb[i] = static_cast<int32_t>(nbk[i]);
      ev_ea[i] = ch_barrier[i];
      ev_de[i] = ch_de[i];
    }
    return t2;
  }
  bool IsSolventAll() const {  // IsEscaped (KineticMcAbstract.cpp:192-223)
    for (int s = 0; s < 3; ++s)
      for (size_t id : lat->nn[s][vac])
        if (bm->species[occ[id]] != solvent) return false;
    return true;
  }
  // OneStepSimulation (KineticMcAbstract.cpp:140-182)
  void Step(uint64_t maximum_steps, int32_t *sel_out, double *ea_out, double *de_out) {
    if (!tt.empty()) {
      T = TAt(time);
      beta = 1.0 / kBoltzmann / T;
    }
    uint32_t r[4];
    Draw(seed, steps, replica, 0u, r);
    const double u1 = UnitOpenLeft(Bits64(r[0], r[1]));
    const double u2 = UnitOpenRight(Bits64(r[2], r[3]));
    double dt;
    if (chain) {
      BuildChainEvents();
      dt = ChainTime();
    } else {
      BuildEvents();
      dt = -std::log(u1) / total_rate / kPrefactor;  // CalculateTime (:79-82)
    }
    if (rate_corrector) {  // RateCorrector::GetTimeCorrectionFactor (RateCorrector.hpp:17-24)
      dt *= cv / (1.64 * std::exp(-(0.66 / kBoltzmann / T - 0.7))) / (1 - 13 * cs);
    }
    int sel = 11;  // SelectEvent (KineticMcAbstract.cpp:106-123): lower_bound, else last
    for (int i = 0; i < 12; ++i)
      if (!(ev_cum[i] < u2)) {
        sel = i;
        break;
      }
    if (early_stop && IsSolventAll()) {
      escaped = true;
      steps = maximum_steps + 1;
    }
    const size_t jump = static_cast<size_t>(ev_nb[sel]);
    if (sel_out) *sel_out = ev_nb[sel];
    if (ea_out) *ea_out = ev_ea[sel];
    if (de_out) *de_out = ev_de[sel];
    time += dt;
    energy += ev_de[sel];
    abs_energy += ev_de[sel];
    const V3 rd = RelDist(lat->rel[vac], lat->rel[jump]);
    if (solute_disp) {
      const EN je = bm->species[occ[jump]];
      if (je != solvent) {
        const V3 vd = rd * lat->L;
        const double w = kElemMass[static_cast<int>(je)] / solute_mass;
        for (int k = 0; k < 3; ++k) com[k] -= vd[k] * w;
      }
    }
    for (int k = 0; k < 3; ++k) disp[k] += static_cast<int64_t>(std::llround(rd[k] * 2 * lat->n));
    std::swap(occ[vac], occ[jump]);  // Config::LatticeJump (Config.cpp:431-456)
    ++steps;
    prev_j = vac;  // KineticMcChainAbstract::OneStepSimulation (KineticMcAbstract.cpp:260-263)
    vac = jump;
  }
};

static Kmc *NewKmc(const Model *bm, const Model *em, const Lattice *lat, const uint8_t *occ,
                   const orc_kmc_params *p) {
  auto k = std::make_unique<Kmc>();
  k->bm = bm;
  k->lat = lat;
  k->occ.assign(occ, occ + lat->N);
  k->T = p->temperature;
  k->beta = 1.0 / kBoltzmann / k->T;
  k->time = p->restart_time;
  k->energy = p->restart_energy;
  k->steps = p->restart_steps;
  k->seed = p->seed;
  k->replica = p->replica;
  k->abs_energy = 0;
  if (em) k->abs_energy = DotSeq(em->theta, TotalEncode(*em, occ));  // McAbstract.cpp:26
  size_t nv = 0, vac = 0;
  std::map<EN, size_t, ElemLess> count;
  for (size_t i = 0; i < lat->N; ++i) {
    const EN e = bm->species[occ[i]];
    count[e]++;
    if (e == EN::X) {
      ++nv;
      vac = i;
    }
  }
  if (nv == 0) throw std::runtime_error("vacancy not found");
  k->vac = vac;
  k->rate_corrector = p->rate_corrector != 0;
  k->early_stop = p->early_stop != 0;
  k->solute_disp = p->solute_disp != 0;
  // RateCorrector(GetVacancyConcentration, GetSoluteConcentration("Al")) (KineticMcAbstract.cpp:35)
  k->cv = static_cast<double>(nv) / static_cast<double>(lat->N);
  size_t ns = 0;
  for (auto &kv : count)
    if (kv.first != EN::Al && kv.first != EN::X) ns += kv.second;
  k->cs = static_cast<double>(ns) / static_cast<double>(lat->N);
  size_t mx = 0;  // GetSolventElement (Config.cpp:404-415)
  for (auto &kv : count)
    if (kv.second > mx) {
      mx = kv.second;
      k->solvent = kv.first;
    }
  for (auto &kv : count)  // GetTotalSoluteMass (Config.cpp:153-165)
    if (kv.first != k->solvent) k->solute_mass += kElemMass[static_cast<int>(kv.first)] * kv.second;
  for (int i = 0; i < p->n_tt; ++i) k->tt.emplace_back(p->tt_time[i], p->tt_temp[i]);
  std::sort(k->tt.begin(), k->tt.end());
  if (p->lru_capacity > 0) k->lru = std::make_unique<Lru>(static_cast<size_t>(p->lru_capacity));
  k->threads = std::max(1, p->omp_threads);
  k->chain = p->chain != 0;
  k->prev_j = lat->nn[0][vac][0];  // first neighbour in id order (KineticMcAbstract.cpp:252)
  return k.release();
}

// ------------------------------------------------------------------------------------------
// CanonicalMcSerial driver
struct Cmc {
  const Model *pm;
  const Lattice *lat;
  std::vector<uint8_t> occ;
  double t0, t1, energy, abs_energy;
  uint64_t steps, seed, accepted = 0;
  uint32_t replica;
  // CanonicalMcOmp rounds (CanonicalMcOmp.cpp:40-92): the pairs and dE of the current round,
  // all evaluated on the occupancy at the round's start, and the next one to try
  std::vector<std::pair<size_t, size_t>> round;
  std::vector<double> round_de;
  size_t round_pos = 0;
  // SimulatedAnnealing schedule (lmc/mc/include/SimulatedAnnealing.h:31-71)
  bool anneal = false;
  double T = 0, lowest = 0, recent_best = 0;
  uint64_t last_improvement = 0, last_reheat = 0;
  uint32_t window_trials = 0, window_accepts = 0, reheats = 0;
  // SimulatedAnnealing::UpdateTemperature (SimulatedAnnealing.cpp:99-133), after the trial of
  // step `steps` with maximum_steps N
  void UpdateTemperature(bool accepted_trial, uint64_t N) {
    const uint64_t window = std::max<uint64_t>(1, static_cast<uint64_t>(N * 0.001));
    const uint64_t trigger = std::max<uint64_t>(1, static_cast<uint64_t>(N * 0.05));
    const uint64_t cooldown = std::max<uint64_t>(1, static_cast<uint64_t>(N * 0.10));
    ++window_trials;
    if (accepted_trial) {
      ++window_accepts;
      if (energy < recent_best - 1e-4) {
        recent_best = energy;
        last_improvement = steps;
      }
    }
    if (window_trials >= window) {
      const double acc = static_cast<double>(window_accepts) / static_cast<double>(window_trials);
      if (acc > 0.50) T *= 0.99;
      window_trials = 0;
      window_accepts = 0;
    }
    const double acc_est = window_trials > 0u
                               ? static_cast<double>(window_accepts) / static_cast<double>(window_trials)
                               : 1.0;
    if (reheats < 5 && steps - last_improvement >= trigger && steps - last_reheat >= cooldown &&
        acc_est < 0.05) {
      T *= 1.10;
      last_improvement = steps;
      last_reheat = steps;
      recent_best = energy;
      ++reheats;
    }
    T *= std::exp(-3.0 / static_cast<double>(std::max<uint64_t>(1, N)));
  }
};

}  // namespace orc

using namespace orc;

static void SetErr(char *err, int len, const char *msg) {
  if (err && len > 0) {
    std::strncpy(err, msg, static_cast<size_t>(len - 1));
    err[len - 1] = 0;
  }
}

extern "C" {

void *orc_lattice_new(int n) {
  if (n < 4) return nullptr;
  return new Lattice(BuildLattice(n));
}
void orc_lattice_free(void *lat) { delete static_cast<Lattice *>(lat); }
int orc_lattice_num_sites(const void *lat) {
  return static_cast<int>(static_cast<const Lattice *>(lat)->N);
}
int orc_lattice_neighbors(const void *lat, int site, int shell, int32_t *out) {
  const auto &v = static_cast<const Lattice *>(lat)->nn[shell - 1][static_cast<size_t>(site)];
  for (size_t i = 0; i < v.size(); ++i) out[i] = static_cast<int32_t>(v[i]);
  return static_cast<int>(v.size());
}
void orc_lattice_position(const void *lat, int site, double *rel3) {
  const auto &p = static_cast<const Lattice *>(lat)->rel[static_cast<size_t>(site)];
  rel3[0] = p[0];
  rel3[1] = p[1];
  rel3[2] = p[2];
}

void *orc_model_new(const void *lat, const char *json_path, const char *elements_csv, int kind,
                    char *err, int errlen) {
  try {
    return LoadModel(static_cast<const Lattice *>(lat), json_path, elements_csv, kind);
  } catch (const std::exception &e) {
    SetErr(err, errlen, e.what());
    return nullptr;
  }
}
void orc_model_free(void *m) { delete static_cast<Model *>(m); }
int orc_model_num_eci(const void *m) { return static_cast<int>(static_cast<const Model *>(m)->eci.size()); }
int orc_model_num_features(const void *mv, int which) {
  const Model &m = *static_cast<const Model *>(mv);
  return static_cast<int>(FeatureCount(m, which == 0 ? m.map_mmm : m.map_mm2));
}
int orc_model_num_species(const void *m) {
  return static_cast<int>(static_cast<const Model *>(m)->species.size());
}
int orc_model_code(const void *mv, const char *element) {
  const Model &m = *static_cast<const Model *>(mv);
  const std::string s(element);
  for (size_t i = 0; i < m.species.size(); ++i)
    if (Str(m.species[i]) == s) return static_cast<int>(i);
  return -1;
}

int orc_barrier_de(const void *mv, const void *, const uint8_t *occ, int v, int nb, double *ea,
                   double *de) {
  try {
    BarrierDe(*static_cast<const Model *>(mv), occ, static_cast<size_t>(v),
              static_cast<size_t>(nb), *ea, *de);
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}
int orc_encode(const void *mv, const void *, const uint8_t *occ, int v, int nb, int which,
               double *out) {
  try {
    const Model &m = *static_cast<const Model *>(mv);
    const size_t a = static_cast<size_t>(v), b = static_cast<size_t>(nb);
    m.EnsurePair(a, b);
    std::vector<double> x;
    if (which == 0) x = OneHotEncode(m, Gather(m, occ, m.t_mmm[m.Flat(a, b)]), m.map_mmm);
    else if (which == 1) x = OneHotEncode(m, Gather(m, occ, m.t_mm2[m.Flat(a, b)]), m.map_mm2);
    else {
      m.EnsurePair(b, a);
      x = OneHotEncode(m, Gather(m, occ, m.t_mm2[m.Flat(b, a)]), m.map_mm2);
    }
    std::copy(x.begin(), x.end(), out);
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}
int orc_swap_de(const void *mv, const void *, const uint8_t *occ, int p, int q, double *de) {
  try {
    *de = SwapDe(*static_cast<const Model *>(mv), occ, static_cast<size_t>(p), static_cast<size_t>(q));
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}
int orc_site_de(const void *mv, const void *, const uint8_t *occ, int site, int new_code,
                double *de) {
  try {
    const Model &m = *static_cast<const Model *>(mv);
    *de = SiteDe(m, occ, static_cast<size_t>(site), m.species.at(static_cast<size_t>(new_code)));
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}
int orc_total_encode(const void *mv, const void *, const uint8_t *occ, double *out) {
  try {
    const auto x = TotalEncode(*static_cast<const Model *>(mv), occ);
    std::copy(x.begin(), x.end(), out);
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}
int orc_total_energy(const void *mv, const void *, const uint8_t *occ, double *e) {
  try {
    const Model &m = *static_cast<const Model *>(mv);
    *e = DotSeq(m.theta, TotalEncode(m, occ));
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}

int orc_state_clusters(const void *mv, const void *, int v, int nb, int32_t *out, int cap) {
  const Model &m = *static_cast<const Model *>(mv);
  const size_t a = static_cast<size_t>(v), b = static_cast<size_t>(nb);
  m.EnsurePair(a, b);
  const auto &ids = m.t_state[m.Flat(a, b)];
  int n = 0;
  for (size_t label = 0; label < m.map_state.size(); ++label)
    for (const auto &cl : m.map_state[label]) {
      if (n >= cap) return -1;
      int32_t *r = out + 4 * n;
      r[0] = static_cast<int32_t>(label);
      for (int i = 0; i < 3; ++i)
        r[1 + i] = i < static_cast<int>(cl.size()) ? static_cast<int32_t>(ids[cl[i]]) : -1;
      ++n;
    }
  return n;
}
int orc_orbit_clusters(const void *mv, const void *, int v, int nb, int which, int32_t *out,
                       int cap) {
  const Model &m = *static_cast<const Model *>(mv);
  const size_t a = static_cast<size_t>(v), b = static_cast<size_t>(nb);
  m.EnsurePair(a, b);
  const auto &ids = which == 0 ? m.t_mmm[m.Flat(a, b)] : m.t_mm2[m.Flat(a, b)];
  const Mapping &map = which == 0 ? m.map_mmm : m.map_mm2;
  int n = 0;
  for (size_t o = 0; o < map.size(); ++o)
    for (const auto &cl : map[o]) {
      if (n >= cap) return -1;
      int32_t *r = out + 4 * n;
      const bool sym = cl[0] == kSizeMax;
      r[0] = static_cast<int32_t>(o);
      r[1] = sym ? 1 : 0;
      const size_t off = sym ? 1 : 0;
      r[2] = static_cast<int32_t>(ids[cl[off]]);
      r[3] = cl.size() > off + 1 ? static_cast<int32_t>(ids[cl[off + 1]]) : -1;
      ++n;
    }
  return n;
}

void orc_philox(uint32_t k0, uint32_t k1, const uint32_t ctr[4], uint32_t out[4]) {
  Philox(k0, k1, ctr, out);
}

void *orc_kmc_new(const void *bm, const void *em, const void *lat, const uint8_t *occ,
                  const orc_kmc_params *p, char *err, int errlen) {
  try {
    return NewKmc(static_cast<const Model *>(bm), static_cast<const Model *>(em),
                  static_cast<const Lattice *>(lat), occ, p);
  } catch (const std::exception &e) {
    SetErr(err, errlen, e.what());
    return nullptr;
  }
}
void orc_kmc_free(void *k) { delete static_cast<Kmc *>(k); }
int orc_kmc_run(void *kv, uint64_t steps, uint64_t maximum_steps, int32_t *tr_sel, double *tr_ea,
                double *tr_de, double *tr_time, double *tr_rate) {
  Kmc &k = *static_cast<Kmc *>(kv);
  try {
    for (uint64_t i = 0; i < steps && k.steps <= maximum_steps; ++i) {
      k.Step(maximum_steps, tr_sel ? tr_sel + i : nullptr, tr_ea ? tr_ea + i : nullptr,
             tr_de ? tr_de + i : nullptr);
      if (tr_time) tr_time[i] = k.time;
      if (tr_rate) tr_rate[i] = k.total_rate;
    }
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}
void orc_kmc_state(const void *kv, uint64_t *steps, double *time, double *energy,
                   double *abs_energy, double *temperature, int32_t *vac, int64_t *disp3,
                   double *com3, int *escaped) {
  const Kmc &k = *static_cast<const Kmc *>(kv);
  *steps = k.steps;
  *time = k.time;
  *energy = k.energy;
  *abs_energy = k.abs_energy;
  *temperature = k.T;
  *vac = static_cast<int32_t>(k.vac);
  for (int i = 0; i < 3; ++i) {
    disp3[i] = k.disp[i];
    com3[i] = k.com[i];
  }
  *escaped = k.escaped ? 1 : 0;
}
void orc_kmc_occupancy(const void *kv, uint8_t *occ) {
  const Kmc &k = *static_cast<const Kmc *>(kv);
  std::copy(k.occ.begin(), k.occ.end(), occ);
}
int orc_kmc_events(void *kv, int32_t *nb, double *ea, double *de, double *rate, double *cum) {
  Kmc &k = *static_cast<Kmc *>(kv);
  try {
    if (k.chain) {
      k.BuildChainEvents();
      (void)k.ChainTime();
      for (int i = 0; i < 12; ++i) k.ev_rate[i] = k.ch_rate[i];
    } else {
      k.BuildEvents();
    }
  } catch (const std::exception &) {
    return -1;
  }
  for (int i = 0; i < 12; ++i) {
    nb[i] = k.ev_nb[i];
    ea[i] = k.ev_ea[i];
    de[i] = k.ev_de[i];
    rate[i] = k.ev_rate[i];
    cum[i] = k.ev_cum[i];
  }
  return 0;
}
double orc_kmc_step_time(void *kv) {
  Kmc &k = *static_cast<Kmc *>(kv);
  try {
    if (k.chain) {
      k.BuildChainEvents();
      return k.ChainTime();
    }
    k.BuildEvents();
    return 1 / k.total_rate / kPrefactor;
  } catch (const std::exception &) {
    return -1.0;
  }
}
uint64_t orc_kmc_lru_hits(const void *kv) {
  const Kmc &k = *static_cast<const Kmc *>(kv);
  return k.lru ? k.lru->hits : 0;
}

void *orc_cmc_new(const void *pm, const void *em, const void *lat, const uint8_t *occ,
                  const orc_cmc_params *p, char *err, int errlen) {
  try {
    auto c = std::make_unique<Cmc>();
    c->pm = static_cast<const Model *>(pm);
    c->lat = static_cast<const Lattice *>(lat);
    c->occ.assign(occ, occ + c->lat->N);
    c->t0 = p->t_start;
    c->t1 = p->t_end;
    c->seed = p->seed;
    c->replica = p->replica;
    c->steps = p->restart_steps;
    c->energy = p->restart_energy;
    c->abs_energy = 0;
    c->anneal = p->anneal != 0;
    c->T = p->t_start;
    c->lowest = c->recent_best = c->energy;
    c->last_improvement = c->steps;
    if (em) {
      const Model &m = *static_cast<const Model *>(em);
      c->abs_energy = DotSeq(m.theta, TotalEncode(m, occ));
    }
    return c.release();
  } catch (const std::exception &e) {
    SetErr(err, errlen, e.what());
    return nullptr;
  }
}
void orc_cmc_free(void *c) { delete static_cast<Cmc *>(c); }
void orc_cmc_anneal_state(const void *cv, double *temperature, double *lowest, uint32_t *reheats) {
  const Cmc &c = *static_cast<const Cmc *>(cv);
  *temperature = c.T;
  *lowest = c.lowest;
  *reheats = c.reheats;
}
int orc_cmc_run(void *cv, uint64_t trials, uint64_t maximum_steps, int32_t *tr_p, int32_t *tr_q,
                double *tr_de, int8_t *tr_acc) {
  Cmc &c = *static_cast<Cmc *>(cv);
  const Model &m = *c.pm;
  const uint64_t N = c.lat->N;
  try {
    for (uint64_t i = 0; i < trials && c.steps <= maximum_steps; ++i) {
      // GenerateLatticeIdJumpPair (CanonicalMcAbstract.cpp:43-51)
      size_t p = 0, q = 0;
      for (uint32_t att = 0;; ++att) {
        if (att >= (1u << 20)) throw std::runtime_error("no distinct pair");
        uint32_t r[4];
        Draw(c.seed, c.steps, c.replica, 2u * att + 1u, r);
        p = Below(Bits64(r[0], r[1]), N);
        q = Below(Bits64(r[2], r[3]), N);
        if (c.occ[p] != c.occ[q]) break;
      }
      const double de = SwapDe(m, c.occ.data(), p, q);
      double beta;
      if (c.anneal) {
        if (c.energy < c.lowest - 1e-4) c.lowest = c.energy;  // Dump (SimulatedAnnealing.cpp:80-97)
        beta = 1.0 / kBoltzmann / std::max(c.T, 1e-12);
      } else {
        const double T = c.t0 + (c.t1 - c.t0) * (static_cast<double>(c.steps) /
                                                 static_cast<double>(maximum_steps));
        beta = 1.0 / kBoltzmann / T;
      }
      bool acc = false;  // SelectEvent (CanonicalMcAbstract.cpp:86-101)
      if (de < 0) {
        acc = true;
      } else {
        uint32_t r[4];
        Draw(c.seed, c.steps, c.replica, 0xFFFFFFFEu, r);
        acc = UnitOpenRight(Bits64(r[0], r[1])) < std::exp(-de * beta);
      }
      if (acc) {
        std::swap(c.occ[p], c.occ[q]);
        c.energy += de;
        c.abs_energy += de;
        ++c.accepted;
      }
      if (tr_p) tr_p[i] = static_cast<int32_t>(p);
      if (tr_q) tr_q[i] = static_cast<int32_t>(q);
      if (tr_de) tr_de[i] = de;
      if (tr_acc) tr_acc[i] = acc ? 1 : 0;
      if (c.anneal) c.UpdateTemperature(acc, maximum_steps);
      ++c.steps;
    }
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}
// CanonicalMcOmp::BuildEventVector (CanonicalMcOmp.cpp:40-78) for the round starting at the
// current step s0, with up to K events: the attempts are numbered through the round, attempt a
// giving the pair of Philox counter (s0, replica, 2a + 1), redrawn while the species agree
// (GenerateLatticeIdJumpPair, CanonicalMcAbstract.cpp:43-51). Each event takes the first drawn
// pair with neither site in the unavailable set; the 50th draw of an event ends the round
// (the `ct == 50` breaks). An accepted pair adds both sites and their 1st-3rd neighbour shells
// to the set. Every dE is taken on the occupancy at the round's start.
static void BuildCmcRound(Cmc &c, int K) {
  const Model &m = *c.pm;
  const uint64_t N = c.lat->N;
  std::vector<char> unavailable(N, 0);
  c.round.clear();
  c.round_de.clear();
  c.round_pos = 0;
  uint32_t att = 0;
  for (int i = 0; i < K; ++i) {
    int ct = 0;
    size_t p = 0, q = 0;
    do {
      for (;; ++att) {  // GenerateLatticeIdJumpPair
        if (att >= (1u << 20)) throw std::runtime_error("no distinct pair");
        uint32_t r[4];
        Draw(c.seed, c.steps, c.replica, 2u * att + 1u, r);
        p = Below(Bits64(r[0], r[1]), N);
        q = Below(Bits64(r[2], r[3]), N);
        if (c.occ[p] != c.occ[q]) {
          ++att;
          break;
        }
      }
      ++ct;
      if (ct == 50) break;
    } while (unavailable[p] || unavailable[q]);
    if (ct == 50) break;
    for (size_t site : {p, q}) {
      unavailable[site] = 1;
      for (int sh = 0; sh < 3; ++sh)
        for (size_t nb : c.lat->nn[sh][site]) unavailable[nb] = 1;
    }
    c.round.emplace_back(p, q);
  }
  for (const auto &pq : c.round) c.round_de.push_back(SwapDe(m, c.occ.data(), pq.first, pq.second));
}

// CanonicalMcOmp::Simulate (CanonicalMcOmp.cpp:79-92): a new round starts only while steps <=
// maximum_steps, and a started round runs to its end; each event is the Metropolis step of
// CanonicalMcAbstract::SelectEvent (:86-101) at its own step, uniform of Philox counter (step,
// replica, 0xFFFFFFFE), temperature of the linear ramp. Rounds carry over between calls.
int orc_cmc_omp_run(void *cv, uint64_t trials, uint64_t maximum_steps, int K, int32_t *tr_p,
                    int32_t *tr_q, double *tr_de, int8_t *tr_acc) {
  Cmc &c = *static_cast<Cmc *>(cv);
  if (K < 1 || c.anneal) return -1;
  try {
    for (uint64_t i = 0; i < trials; ++i) {
      if (c.round_pos == c.round.size()) {
        if (c.steps > maximum_steps) break;
        BuildCmcRound(c, K);
      }
      const auto [p, q] = c.round[c.round_pos];
      const double de = c.round_de[c.round_pos];
      ++c.round_pos;
      const double T = c.t0 + (c.t1 - c.t0) * (static_cast<double>(c.steps) /
                                               static_cast<double>(maximum_steps));
      const double beta = 1.0 / kBoltzmann / T;
      bool acc = de < 0;
      if (!acc) {
        uint32_t r[4];
        Draw(c.seed, c.steps, c.replica, 0xFFFFFFFEu, r);
        acc = UnitOpenRight(Bits64(r[0], r[1])) < std::exp(-de * beta);
      }
      if (acc) {
        std::swap(c.occ[p], c.occ[q]);
        c.energy += de;
        c.abs_energy += de;
        ++c.accepted;
      }
      if (tr_p) tr_p[i] = static_cast<int32_t>(p);
      if (tr_q) tr_q[i] = static_cast<int32_t>(q);
      if (tr_de) tr_de[i] = de;
      if (tr_acc) tr_acc[i] = acc ? 1 : 0;
      ++c.steps;
    }
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}
// Events of the current round not yet tried.
int orc_cmc_round_left(const void *cv) {
  const Cmc &c = *static_cast<const Cmc *>(cv);
  return static_cast<int>(c.round.size() - c.round_pos);
}

void orc_cmc_state(const void *cv, uint64_t *steps, double *energy, double *abs_energy,
                   uint64_t *accepted) {
  const Cmc &c = *static_cast<const Cmc *>(cv);
  *steps = c.steps;
  *energy = c.energy;
  *abs_energy = c.abs_energy;
  *accepted = c.accepted;
}
void orc_cmc_occupancy(const void *cv, uint8_t *occ) {
  const Cmc &c = *static_cast<const Cmc *>(cv);
  std::copy(c.occ.begin(), c.occ.end(), occ);
}

double orc_kmc_baseline(const void *bm, const void *em, const void *lat, const uint8_t *occ,
                        double temperature, uint64_t seed, int replicas, int threads_per_replica,
                        int lru_capacity, uint64_t steps, uint64_t *lru_hits, int chain) {
  const Model &m = *static_cast<const Model *>(bm);
  const Lattice &L = *static_cast<const Lattice *>(lat);
  // Build every pair table up front, as the predictor constructor does (...Quartic.cpp:64-100).
#pragma omp parallel for schedule(dynamic, 64)
  for (long long s = 0; s < static_cast<long long>(L.N); ++s)
    for (size_t nbr : L.nn[0][static_cast<size_t>(s)]) m.EnsurePair(static_cast<size_t>(s), nbr);
  std::vector<std::unique_ptr<Kmc>> ks;
  for (int r = 0; r < replicas; ++r) {
    orc_kmc_params p{};
    p.temperature = temperature;
    p.seed = seed;
    p.replica = static_cast<uint32_t>(r);
    p.lru_capacity = lru_capacity;
    p.omp_threads = threads_per_replica;
    p.chain = chain;
    ks.emplace_back(NewKmc(&m, static_cast<const Model *>(em), &L, occ, &p));
  }
#ifdef _OPENMP
  omp_set_max_active_levels(2);
#endif
  const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for num_threads(replicas) schedule(static, 1) if (replicas > 1)
  for (int r = 0; r < replicas; ++r)
    for (uint64_t i = 0; i < steps; ++i) ks[static_cast<size_t>(r)]->Step(~0ull - 1, nullptr, nullptr, nullptr);
  const auto t1 = std::chrono::steady_clock::now();
  uint64_t hits = 0;
  for (auto &k : ks) hits += k->lru ? k->lru->hits : 0;
  if (lru_hits) *lru_hits = hits;
  return std::chrono::duration<double>(t1 - t0).count();
}

// The CanonicalMc / SimulatedAnnealing trials of `replicas` independent replicas of occ, one
// thread each, `trials` trials per replica (maximum_steps for the ramp and the schedule);
// returns the wall seconds of the timed loop (tables built before it).
double orc_cmc_baseline(const void *pm, const void *lat, const uint8_t *occ, double t_start,
                        uint64_t seed, int replicas, uint64_t trials, uint64_t maximum_steps,
                        int anneal) {
  const Model &m = *static_cast<const Model *>(pm);
  m.PrepareSites();
  std::vector<void *> cs;
  for (int r = 0; r < replicas; ++r) {
    orc_cmc_params p{};
    p.t_start = t_start;
    p.t_end = t_start;
    p.seed = seed;
    p.replica = static_cast<uint32_t>(r);
    p.anneal = anneal;
    void *c = orc_cmc_new(pm, nullptr, lat, occ, &p, nullptr, 0);
    if (!c) {
      for (void *d : cs) orc_cmc_free(d);
      return -1.0;
    }
    cs.push_back(c);
  }
  const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for num_threads(replicas) schedule(static, 1) if (replicas > 1)
  for (int r = 0; r < replicas; ++r)
    (void)orc_cmc_run(cs[static_cast<size_t>(r)], trials, maximum_steps, nullptr, nullptr, nullptr,
                      nullptr);
  const auto t1 = std::chrono::steady_clock::now();
  for (void *c : cs) orc_cmc_free(c);
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
