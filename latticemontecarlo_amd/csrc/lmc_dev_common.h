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
// A replica's lattice size and a site id are usable (uniform over a block).
__device__ __forceinline__ bool ValidSite(int n, int64_t site) {
  return n >= 4 && site >= 0 && site < 4ll * n * n * n;
}

}  // namespace lmc
