// wx_hilbert.h - Hilbert keys of blocks (host code): getHilbertKey and
// HKey3 (source/amr.f90:2031-2061, source/hilbert.f90:38-127), the key of a
// block's lowest finest-level cell along the 3-D Hilbert curve of the
// reference, used by its load balance (HilbertOrder, amr.f90:1930-1996) and
// here as a block processing order.
#pragma once

#include <algorithm>
#include <cmath>

namespace wx {

// The 12-state machine of the 3-D curve (the published table of Lawder's
// state-diagram method, as the reference tabulates it): for current state s
// and octant digit d = 4*x_bit + 2*y_bit + z_bit, the next state and the key
// digit.
constexpr unsigned char HILBERT_NEXT[12][8] = {
    {1, 2, 3, 2, 4, 5, 3, 5},    {2, 6, 0, 7, 8, 8, 0, 7},   {0, 9, 10, 9, 1, 1, 11, 11},
    {6, 0, 6, 11, 9, 0, 9, 8},   {11, 11, 0, 7, 5, 9, 0, 7}, {4, 4, 8, 8, 0, 6, 10, 6},
    {5, 7, 5, 3, 1, 1, 11, 11},  {6, 1, 6, 10, 9, 4, 9, 10}, {10, 3, 1, 1, 10, 3, 5, 9},
    {4, 4, 8, 8, 2, 7, 2, 3},    {7, 2, 11, 2, 7, 5, 8, 5},  {10, 3, 2, 6, 10, 3, 4, 4}};
constexpr unsigned char HILBERT_DIGIT[12][8] = {
    {0, 1, 3, 2, 7, 6, 4, 5}, {0, 7, 1, 6, 3, 4, 2, 5}, {0, 3, 7, 4, 1, 2, 6, 5},
    {2, 3, 1, 0, 5, 4, 6, 7}, {4, 3, 5, 2, 7, 0, 6, 1}, {6, 5, 1, 2, 7, 4, 0, 3},
    {4, 7, 3, 0, 5, 6, 2, 1}, {6, 7, 5, 4, 1, 0, 2, 3}, {2, 5, 3, 4, 1, 6, 0, 7},
    {2, 1, 5, 6, 3, 0, 4, 7}, {4, 5, 7, 6, 3, 2, 0, 1}, {6, 1, 7, 0, 5, 2, 4, 3}};

// key of cell (x, y, z) (0-based) on a curve of 2^order cells per edge
inline unsigned hilbert_key3(unsigned x, unsigned y, unsigned z, int order) {
  unsigned key = 0;
  int state = 0;
  for (int i = order - 1; i >= 0; i--) {
    const int d = (int)((((x >> i) & 1u) << 2) | (((y >> i) & 1u) << 1) | ((z >> i) & 1u));
    key |= (unsigned)HILBERT_DIGIT[state][d] << (3 * i);
    state = HILBERT_NEXT[state][d];
  }
  return key;
}

// curve order of a mesh: ceiling(log(maxnb)/log(2)) in double precision,
// maxnb the largest block count along an axis at the finest level
inline int hilbert_order(const int nroot[3], int maxlev) {
  long maxnb = (long)std::max(nroot[0], std::max(nroot[1], nroot[2])) << (maxlev - 1);
  return (int)std::ceil(std::log(1.0 * (double)maxnb) / std::log(2.0));
}

// key of the block at `level` with 1-based block coordinates c
inline unsigned block_hilbert_key(const int c[3], int level, int maxlev, int order) {
  const int sh = maxlev - level;
  return hilbert_key3((unsigned)(c[0] - 1) << sh, (unsigned)(c[1] - 1) << sh, (unsigned)(c[2] - 1) << sh, order);
}

}  // namespace wx
