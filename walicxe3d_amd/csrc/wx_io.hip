// wx_io.hip - the reference's checkpoint files (host code; see wx_io.h).
#include "wx_io.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "wx_common.h"

namespace wx {

namespace {

struct File {
  FILE *f = nullptr;
  File(const char *path, const char *mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

}  // namespace

int bin_write(const char *path, int nblk, int nv, int npassive, const int *bids, const double *u,
              std::string &err) {
  File out(path, "wb");
  if (!out.f) {
    err = std::string("cannot open ") + path;
    return 1;
  }
  const int neq = nv + npassive;
  const int32_t n = nblk;
  bool ok = std::fwrite(&n, 4, 1, out.f) == 1;
  std::vector<double> rec((size_t)neq * NINT, 0.0);
  for (int b = 0; b < nblk && ok; b++) {
    const int32_t id = bids[b];
    ok = std::fwrite(&id, 4, 1, out.f) == 1;
    // [v][cell] -> [cell][v]: the variable runs fastest in the file
    for (int v = 0; v < nv; v++) {
      const double *U = u + ((size_t)b * nv + v) * NINT;
      for (int q = 0; q < NINT; q++) rec[(size_t)q * neq + v] = U[q];
    }
    ok = ok && std::fwrite(rec.data(), sizeof(double), rec.size(), out.f) == rec.size();
  }
  if (!ok) err = std::string("write failed: ") + path;
  return ok ? 0 : 1;
}

int bin_read(const char *path, int nv, int npassive, int cap, int *nblk, int *bids, double *u,
             std::string &err) {
  File in(path, "rb");
  if (!in.f) {
    err = std::string("cannot open ") + path;
    return 1;
  }
  const int neq = nv + npassive;
  int32_t n = 0;
  if (std::fread(&n, 4, 1, in.f) != 1 || n < 0) {
    err = "bad block count";
    return 1;
  }
  std::fseek(in.f, 0, SEEK_END);
  const long size = std::ftell(in.f);
  const long expect = 4 + (long)n * (4 + (long)neq * NINT * 8);
  if (size != expect) {
    err = "file size " + std::to_string(size) + " does not match " + std::to_string(n) + " blocks of " +
          std::to_string(neq) + " variables (" + std::to_string(expect) + " bytes)";
    return 1;
  }
  *nblk = n;
  if (n > cap) {
    err = "file holds " + std::to_string(n) + " blocks, more than the capacity";
    return 2;
  }
  std::fseek(in.f, 4, SEEK_SET);
  std::vector<double> rec((size_t)neq * NINT);
  for (int b = 0; b < n; b++) {
    int32_t id = 0;
    if (std::fread(&id, 4, 1, in.f) != 1 || std::fread(rec.data(), sizeof(double), rec.size(), in.f) != rec.size()) {
      err = "short read";
      return 1;
    }
    bids[b] = id;
    for (int v = 0; v < nv; v++) {
      double *U = u + ((size_t)b * nv + v) * NINT;
      for (int q = 0; q < NINT; q++) U[q] = rec[(size_t)q * neq + v];
    }
  }
  return 0;
}

int state_write(const char *path, double time_phys, int it, int noutput, const char *datadir, std::string &err) {
  File out(path, "w");
  if (!out.f) {
    err = std::string("cannot open ") + path;
    return 1;
  }
  // es22.15 with a two-digit exponent: C's %22.15E prints the same text
  // while |exponent| < 100
  const int ok = std::fprintf(out.f, "%22.15E%8d%5d\n%s\n", time_phys, it, noutput, datadir ? datadir : "") > 0;
  if (!ok) err = std::string("write failed: ") + path;
  return ok ? 0 : 1;
}

int state_read(const char *path, double *time_phys, int *it, int *noutput, char *datadir, int len,
               std::string &err) {
  File in(path, "r");
  if (!in.f) {
    err = std::string("cannot open ") + path;
    return 1;
  }
  char line[512];
  if (!std::fgets(line, sizeof line, in.f) || std::strlen(line) < 35) {
    err = "bad state record";
    return 1;
  }
  // fixed-width fields of (es22.15, i8, i5)
  char f1[23], f2[9], f3[6];
  std::memcpy(f1, line, 22);
  f1[22] = 0;
  std::memcpy(f2, line + 22, 8);
  f2[8] = 0;
  std::memcpy(f3, line + 30, 5);
  f3[5] = 0;
  *time_phys = std::strtod(f1, nullptr);
  *it = (int)std::strtol(f2, nullptr, 10);
  *noutput = (int)std::strtol(f3, nullptr, 10);
  if (datadir && len > 0) {
    datadir[0] = 0;
    if (std::fgets(line, sizeof line, in.f)) {
      size_t l = std::strcspn(line, "\r\n");
      while (l > 0 && line[l - 1] == ' ') l--;   // '(a)' into a blank-padded variable
      if ((int)l >= len) l = len - 1;
      std::memcpy(datadir, line, l);
      datadir[l] = 0;
    }
  }
  return 0;
}

}  // namespace wx
