// Mode, hop and lattice helpers shared by the kernels and by the host emulation of the image
// kernels (tests/host): plain C++ with __host__ __device__ qualifiers, which the host build
// defines away.
#pragma once

#include <stdint.h>

namespace nh {

constexpr int32_t kHopNone = -1;     // item is the configuration itself
constexpr int32_t kHopInvalid = -2;  // unused item slot: every kernel skips it

// hop packing: from | to << 16 (modes < 2^15)
__host__ __device__ inline int32_t pack_hop(int from, int to) { return from | (to << 16); }
__host__ __device__ inline int hop_from(int32_t h) { return h & 0xffff; }
__host__ __device__ inline int hop_to(int32_t h) { return (h >> 16) & 0xffff; }

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

// Occupation (+1/-1) of mode m in configuration b with an optional hop applied.
__host__ __device__ inline int8_t item_occ(const int8_t* __restrict__ s_b, int32_t hop, int m) {
  if (hop >= 0) {
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
