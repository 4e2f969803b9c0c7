// Mode, hop and lattice helpers shared by the kernels and by the host emulation of the image
// kernels (tests/host): plain C++ with __host__ __device__ qualifiers, which the host build
// defines away.
#pragma once

#include <stdint.h>

namespace nh {

constexpr int32_t kHopNone = -1;     // item is the configuration itself
constexpr int32_t kHopInvalid = -2;  // unused item slot: every kernel skips it

// hop packing: from | to << 16 (modes < 2^14, so bit 30 stays clear)
__host__ __device__ inline int32_t pack_hop(int from, int to) { return from | (to << 16); }
__host__ __device__ inline int hop_from(int32_t h) { return h & 0xffff; }
__host__ __device__ inline int hop_to(int32_t h) { return (h >> 16) & 0xffff; }

// site exchange: the occupations of sites i and j trade places in both spin blocks (an up or
// down electron hops, a doublon and a hole or an up and a down electron swap); packed as
// kExchangeBit | i | j << 15 (sites < 2^14).  Only the Metropolis proposals make them.
constexpr int32_t kExchangeBit = 1 << 30;
__host__ __device__ inline int32_t pack_exchange(int i, int j) { return kExchangeBit | i | (j << 15); }
__host__ __device__ inline bool is_exchange(int32_t h) { return h >= 0 && (h & kExchangeBit); }
__host__ __device__ inline int exchange_i(int32_t h) { return h & 0x7fff; }
__host__ __device__ inline int exchange_j(int32_t h) { return (h >> 15) & 0x7fff; }

// Image of mode m under group element g = fx | fy << 1 | si << 2 (every g is an involution).
__host__ __device__ inline int group_map(int m, int g, int N, int LX, int LY) {
  int spin = m >= N;
  int site = m - spin * N;
  int x = site / LY, y = site - (site / LY) * LY;
  if (g & 1) x = (x == 0) ? 0 : LX - x;
  if (g & 2) y = (y == 0) ? 0 : LY - y;
  spin ^= (g >> 2) & 1;
  return spin * N + x * LY + y;
}

// Occupation (+1/-1) of mode m of N-site configuration b with an optional hop or site
// exchange applied.
__host__ __device__ inline int8_t item_occ(const int8_t* __restrict__ s_b, int32_t hop, int m, int N) {
  if (hop >= 0) {
    if (hop & kExchangeBit) {
      const int off = m >= N ? N : 0, site = m - off;
      if (site == exchange_i(hop)) return s_b[off + exchange_j(hop)];
      if (site == exchange_j(hop)) return s_b[off + exchange_i(hop)];
      return s_b[m];
    }
    if (m == hop_from(hop)) return -1;
    if (m == hop_to(hop)) return 1;
  }
  return s_b[m];
}

// periodic displacement index r(p - q) of two sites
__host__ __device__ inline int disp_index(int sp, int sq, int LX, int LY) {
  int xp = sp / LY, yp = sp - xp * LY;
  int xq = sq / LY, yq = sq - xq * LY;
  int dx = xp - xq, dy = yp - yq;
  dx += (dx < 0) ? LX : 0;
  dy += (dy < 0) ? LY : 0;
  return dx * LY + dy;
}

}  // namespace nh
