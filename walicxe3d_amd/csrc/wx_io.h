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

}  // namespace wx
