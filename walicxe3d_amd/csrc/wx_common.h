// wx_common.h - shared constants and device data structures.
#pragma once

#include <cstdint>

namespace wx {

constexpr int NC = 16;             // cells per block edge (ncells_block)
constexpr int NG = 2;              // ghost layers (nghost)
constexpr int NB = NC + 2 * NG;    // 20: stored edge with ghosts
constexpr int NCELL = NB * NB * NB;
constexpr int NINT = NC * NC * NC; // 4096 interior cells
constexpr int MAXV = 8;

// source/constants.f90:52-104
enum { LEFT = 0, RIGHT, FRONT, BACK, BOTTOM, TOP };     // 0-based face index
enum { BC_REFLECTIVE = 1, BC_FREE_FLOW = 2, BC_PERIODIC = 3 };
enum { NEIGH_SAME = 1, NEIGH_COARSER = 2, NEIGH_FINER = 3, NEIGH_BOUNDARY = 4 };
enum { SOLVER_HLL1 = 2, SOLVER_HLL = 3, SOLVER_HLLC = 4, SOLVER_HLLE = 5, SOLVER_HLLD = 6 };

// Index checks of the CPU emulation of the kernels (the emulator build
// defines WX_EMU_CHECKS): a violated range aborts with its source line.
// Nothing in the GPU build.
#ifdef WX_EMU_CHECKS
void wx_emu_check_failed(const char *what, const char *file, int line);
#define WX_EMU_CHECK(c) ((c) ? (void)0 : wx_emu_check_failed(#c, __FILE__, __LINE__))
#else
#define WX_EMU_CHECK(c) ((void)0)
#endif

// Index of cell (i,j,k), i,j,k in [-1, 18] (reference bounds nxmin..nxmax)
__host__ __device__ constexpr inline int cid(int i, int j, int k) {
  WX_EMU_CHECK(i >= -1 && i <= NC + 2 && j >= -1 && j <= NC + 2 && k >= -1 && k <= NC + 2);
  return ((k + 1) * NB + (j + 1)) * NB + (i + 1);
}

// One face of one block in the neighbour table.  For NEIGH_FINER the four
// sources are stored by transverse quadrant q = t1 + 2*t2 (t1, t2 the
// sibling coordinates of the source along the face's first and second
// transverse axes); for the other types only src[0] is used.  A periodic
// domain face is stored as NEIGH_SAME with the wrapped partner.
struct Face {
  int type;
  int src[4];
};

// Per-block record
struct BlockInfo {
  int level;       // 1-based
  int sib[3];      // siblingCoords (0/1), -1 at level 1
  int bid;         // reference block ID
  int coords[3];   // 1-based block coordinates at its level
  Face face[6];
};

}  // namespace wx
