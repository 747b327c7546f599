// wx_io.h - the reference's output files, host side (source/output.f90,
// source/init.f90:527-623).  Arrays are in the library's block layout
// [blk][nv][16][16][16] (i fastest); the files hold what the reference's
// writers put there, byte for byte.
#pragma once

#include <string>

namespace wx {

// writeBin (output.f90:156-214): stream file, int32 block count, then per
// block its int32 bID and U(nb, 1:neqtot, 1:16, 1:16, 1:16) in Fortran order
// (variable fastest, then i, j, k); neqtot = nv + npassive, the passive
// columns (not carried by the library) written as zero.
int bin_write(const char *path, int nblk, int nv, int npassive, const int *bids, const double *u,
              std::string &err);
// warmstart's read of the same file (init.f90:582-610); passive columns
// skipped.  *nblk is set even when the file holds more than cap blocks.
int bin_read(const char *path, int nv, int npassive, int cap, int *nblk, int *bids, double *u,
             std::string &err);
// writeState (output.f90:219-268): '(es22.15,i8,i5)' time (physical units),
// it, noutput, then '(a)' datadir.
int state_write(const char *path, double time_phys, int it, int noutput, const char *datadir, std::string &err);
// warmstart's read of it (init.f90:545-560)
int state_read(const char *path, double *time_phys, int *it, int *noutput, char *datadir, int len,
               std::string &err);

// The mesh and scales the VTK writers need.
struct VtkInfo {
  int nblk = 0, nv = 5, npassive = 0, maxlev = 1, eos = 0;
  int phys_units = 0;           // units_type == PHYS_UNITS
  const int *bids = nullptr;    // [nblk]
  const int *levels = nullptr;  // [nblk]
  const int *coords = nullptr;  // [nblk][3], 1-based at the block's level
  const double *dxyz = nullptr; // dx(1..maxlev), dy(...), dz(...)
  int nroot[3] = {1, 1, 1};
  double l_sc = 1, d_sc = 1, v_sc = 1, p_sc = 1, b_sc = 1;
  double temp_mu0 = 1, temp_mui = 1, ion_thres = 0;   // calcTemp
  double mu0 = 1, kb = 1;       // calcTemp, EOS_H_RATE: mu0 and Boltzmann's constant (cgs)
  int ncarried = 0;             // passive scalars prim holds (after the flow variables)
};

// write3DVTKBlocks (output.f90:273-750): every cell a hexahedron of its own
// 8 points; rho, vel, P, B (8 variables), T and the passive scalars (those
// prim carries, zero for the rest) as FP32, big-endian; prim [blk][nv][16^3].
int vtk_blocks_write(const char *path, const VtkInfo &m, const double *prim, std::string &err);
// write3DVTKGrid (output.f90:755-895): one hexahedron per block with its
// rank (0), bID and Hilbert key.
int vtk_grid_write(const char *path, const VtkInfo &m, std::string &err);
// master.visit of write3DVTKBlocks (output.f90:305-340), one rank:
// replaced at output 0, appended to afterwards.
int visit_write(const char *path, const char *blocks_name, int noutput, bool replace, std::string &err);

}  // namespace wx
