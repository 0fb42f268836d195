// Device helpers shared by the kernels: site id <-> coordinates (reference id convention,
// lmc/cfg/src/Config.cpp:466-470), periodic neighbour arithmetic and packed offsets.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lmc {

constexpr int kThreads = 256;
constexpr double kBoltzmann = 8.617333262145e-5;  // lmc/cfg/include/Constants.hpp:31
constexpr double kPrefactor = 1e13;               // Constants.hpp:32
constexpr double kHalfLattice = 4.046 * 0.5;      // a/2 in Angstrom (Constants.hpp:6)

struct XYZ {
  int X, Y, Z;
};

// id = X * 2n^2 + Y * n + Z/2 with 0 <= X,Y,Z < 2n, X+Y+Z even (a/2 units).
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
// Offsets of the jump environments (|component| <= 4 in a/2 units), 4 bits each, biased by 8.
__device__ __forceinline__ void Unpack16(uint32_t v, int &x, int &y, int &z) {
  x = static_cast<int>(v & 15) - 8;
  y = static_cast<int>((v >> 4) & 15) - 8;
  z = static_cast<int>((v >> 8) & 15) - 8;
}
// Wave priority of step `it` of the calling workgroup (s_setprio takes an immediate). The SIMD
// arbiter favours the oldest waves, so among the workgroups resident on one CU the first
// dispatched would run fastest and the last would set the end of a launch (measured: 1.5x
// between them with four per CU). Co-resident workgroups are 256 block indices apart on a
// 256-CU GPU, so (block / 256 + it) mod 4 gives them distinct priorities that rotate step by
// step, and each gets the same share of the CU over a launch. The two younger ones (blocks
// 512-1023) also move up one level on odd steps: the memory and LDS queues, which s_setprio
// does not reach, still favour the older pair (measured: 6 % slower without the bias, 2 % with
// it).
__device__ __forceinline__ void RotatePriority(uint64_t it) {
  const unsigned base = static_cast<unsigned>((blockIdx.x >> 8) + it) & 3u;
  const unsigned young = (blockIdx.x >> 9) & 1u;
  switch (base + (young & static_cast<unsigned>(it) & (base < 3u ? 1u : 0u))) {
    case 0: __builtin_amdgcn_s_setprio(0); break;
    case 1: __builtin_amdgcn_s_setprio(1); break;
    case 2: __builtin_amdgcn_s_setprio(2); break;
    default: __builtin_amdgcn_s_setprio(3); break;
  }
}

// A replica's lattice size and a site id are usable (uniform over a block).
__device__ __forceinline__ bool ValidSite(int n, int64_t site) {
  return n >= 4 && site >= 0 && site < 4ll * n * n * n;
}

// Image change of one axis for a displacement of d grid steps (a/2) from lhs to rhs on the 2n
// axis: Config::LatticeJump's floor(shift + 0.5) for |shift| > 0.5 + 1e-8 with shift = d / 2n
// (Config.cpp:438-444), exact for |d| < 2n.
__device__ __forceinline__ int ImageChange(int d, int n) { return (d > n) - (d < -n); }

// Config::LatticeJump (Config.cpp:431-456) on the atom ids and map_shift of one replica: the
// atoms at lattice ids lhs and rhs swap places; the lhs atom's map_shift loses the image change
// (icx, icy, icz) of the displacement lhs -> rhs and the rhs atom's gains it.
__device__ __forceinline__ void AtomJump(int32_t *atom, int32_t *shift, int64_t lhs, int64_t rhs,
                                         int icx, int icy, int icz) {
  const int32_t al = atom[lhs], ar = atom[rhs];
  atom[lhs] = ar;
  atom[rhs] = al;
  if (icx | icy | icz) {
    int32_t *sl = shift + 3 * static_cast<int64_t>(al), *sr = shift + 3 * static_cast<int64_t>(ar);
    sl[0] -= icx;
    sl[1] -= icy;
    sl[2] -= icz;
    sr[0] += icx;
    sr[1] += icy;
    sr[2] += icz;
  }
}

// v of lane l - kShift within the lane's row of 16 (0 past the row start): DPP row_shr on both
// halves of the double.
template <int kShift>
__device__ __forceinline__ double RowShr(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x110 + kShift, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x110 + kShift, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double ReadLane(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Sum over the 64 lanes of a full wave, returned to every lane: shifted adds within each row of
// 16 leave the row sum in its last lane; the four row sums are then added in order.
__device__ __forceinline__ double WaveSum(double v) {
  v += RowShr<1>(v);
  v += RowShr<2>(v);
  v += RowShr<4>(v);
  v += RowShr<8>(v);
  return (ReadLane(v, 15) + ReadLane(v, 31)) + (ReadLane(v, 47) + ReadLane(v, 63));
}

}  // namespace lmc
