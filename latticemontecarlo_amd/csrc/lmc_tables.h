// Translation-invariant cluster tables of the FCC jump environment, built with exact integer
// geometry. Coordinates are in units of a/2 (FCC sites are the integer points with even
// coordinate sum). One table set serves every lattice size and every site: the environment of a
// jump in direction d is the image of a canonical environment under the lattice rotation that
// maps the canonical frame onto d's frame.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace lmc {

constexpr int kEnvSlots = 60;   // sites within 3NN of either endpoint of a 1NN pair
constexpr int kSiteSlots = 43;  // a site and its 1NN/2NN/3NN shells
constexpr int kMaxSpecies = 8;  // species codes incl. the vacancy

struct I3 {
  int x, y, z;
};

// 12 first-neighbour jump vectors, fixed order; 6 second- and 24 third-neighbour vectors.
extern const I3 kNN1[12];
extern const I3 kNN2[6];
extern const I3 kNN3[24];

// Shell of an offset: 1/2/3 for first/second/third neighbours, 0 otherwise (includes the origin).
int ShellOf(const I3 &d);

// Site id <-> coordinates for an n x n x n cubic FCC box with the reference id convention
// (lmc/cfg/src/Config.cpp:466-470: ids are the x-major rank of the fractional positions):
// id = X * 2n^2 + Y * n + floor(Z / 2), 0 <= X,Y,Z < 2n, X+Y+Z even.
inline int64_t SiteId(int n, int X, int Y, int Z) {
  return static_cast<int64_t>(X) * 2 * n * n + static_cast<int64_t>(Y) * n + (Z >> 1);
}
inline I3 SiteXYZ(int n, int64_t id) {
  const int64_t plane = 2ll * n * n;
  const int X = static_cast<int>(id / plane);
  const int64_t r = id % plane;
  const int Y = static_cast<int>(r / n);
  const int Zh = static_cast<int>(r % n);
  return {X, Y, 2 * Zh + ((X + Y) & 1)};
}
int64_t SiteAdd(int n, int64_t id, const I3 &d);

// One cluster of the jump environment (state tables, lmc/pred/src/EnergyUtility.cpp:393-499).
struct StateCluster {
  int label;           // 0 singlet, 1..3 pair shell, 4..7 triplet type
  int slots[3];        // env slots, -1 padded
  int size;
};

// One cluster of the MMM/MM2 orbit tables (EnergyUtility.cpp:169-259), over the 58
// non-endpoint slots. s0 precedes s1 in the reference's within-cluster order.
struct OrbitCluster {
  int orbit;
  int s0, s1;  // s1 = -1 for singlets
  bool sym;
};

struct OrbitTable {
  std::vector<OrbitCluster> clusters;  // grouped by orbit, orbits in reference order
  std::vector<int> orbit_size;         // clusters per orbit
  std::vector<int> orbit_kind;         // 0 singlet, 1 pair, 2 symmetric pair
  std::vector<int> orbit_shell;        // 0 singlet, 1..3 pair shell
  // Feature layout for `ne` elements (X excluded): base offset and length of each orbit.
  void Layout(int ne, std::vector<int> &base, std::vector<int> &len, int &total) const;
};

struct JumpTables {
  // Canonical slot keys (Xk, Yk, Zk) = (R.d, R.p, R.q) with R = 2*(site - v) - d in the frame
  // (d, p, q) of the jump; slots are sorted in the reference's state order (x, then y, then z).
  std::array<I3, kEnvSlots> key{};
  int v_slot = -1, n_slot = -1;
  // Offsets (site - vacancy) of every slot for each of the 12 jump directions.
  std::array<std::array<I3, kEnvSlots>, 12> off{};
  // State clusters containing an endpoint, and the subsets containing exactly one endpoint.
  std::vector<StateCluster> state;
  std::vector<StateCluster> state_v_only, state_n_only;
  OrbitTable mmm, mm2;
  // mirror[j]: slot holding, in the frame of the reversed jump, the position of slot j
  // (a half turn about the frame's third axis) -- the backward MM2 environment.
  std::array<int, kEnvSlots> mirror{};
  // Slots within the first three shells of the vacancy (IsEscaped neighbourhood).
  std::vector<int> vac_shell_slots;
};

struct SiteTables {
  std::array<I3, kSiteSlots> off{};     // slot offsets from the centre, centre = slot 0
  std::vector<StateCluster> clusters;   // clusters containing slot 0 (247)
};

const JumpTables &GetJumpTables();
const SiteTables &GetSiteTables();

// ECI key space (EnergyUtility.cpp:314-343 ordered by ElementCluster::operator<,
// lmc/cfg/include/ElementCluster.hpp:40-50) for species names sorted by string incl. "X".
struct EciSpace {
  std::vector<std::string> species;   // sorted, includes "X"
  int x_code = -1;
  struct Key {
    int label;
    std::vector<int> codes;  // sorted species codes (string order == code order)
  };
  std::vector<Key> keys;      // index order
  // Lookup tables by species codes: -1 = key absent.
  std::vector<int> t1;        // [S]
  std::vector<int> t2;        // [3][S][S]
  std::vector<int> t3;        // [4][S][S][S]
  int S() const { return static_cast<int>(species.size()); }
  int Find(int label, std::vector<int> codes) const;
};
EciSpace BuildEciSpace(const std::vector<std::string> &elements_without_x);

}  // namespace lmc
