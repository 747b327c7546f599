// wx_io.hip - the reference's checkpoint files (host code; see wx_io.h).
#include "wx_io.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "wx_common.h"
#include "wx_hilbert.h"

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

// ---------------------------------------------------------------------------
// VTK: the reference opens the files with convert="BIG_ENDIAN" and writes
// default reals as FP32 (real(..., 4)), integers as int32, and text records
// followed by char(10).
// ---------------------------------------------------------------------------
namespace {

class BeWriter {
 public:
  explicit BeWriter(const char *path) : f_(path, "wb") { buf_.reserve(1 << 20); }
  bool ok() const { return f_.f != nullptr && good_; }
  void text(const char *s) {
    put(s, std::strlen(s));
    lf();
  }
  void lf() { put("\n", 1); }
  void f32(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    be32(u);
  }
  void i32(int32_t x) { be32((uint32_t)x); }
  bool close() {
    flush();
    return ok();
  }

 private:
  void be32(uint32_t u) {
    const char b[4] = {(char)(u >> 24), (char)(u >> 16), (char)(u >> 8), (char)u};
    put(b, 4);
  }
  void put(const char *p, size_t n) {
    buf_.insert(buf_.end(), p, p + n);
    if (buf_.size() >= (1u << 20)) flush();
  }
  void flush() {
    if (f_.f && !buf_.empty() && std::fwrite(buf_.data(), 1, buf_.size(), f_.f) != buf_.size()) good_ = false;
    buf_.clear();
  }
  File f_;
  std::vector<char> buf_;
  bool good_ = true;
};

// getRefCorner (amr.f90:1007-1025): (x-1)*ncells*dx(level)
void ref_corner(const VtkInfo &m, int b, double xb[3]) {
  const int l = m.levels[b];
  for (int a = 0; a < 3; a++) xb[a] = (double)((m.coords[3 * b + a] - 1) * NC) * m.dxyz[a * m.maxlev + l - 1];
}

// calcTemp (hydro_core.f90:231-270), EOS_ADIABATIC / SINGLE_SPECIE / TWOTEMP,
// and EOS_H_RATE (:258-264) with the neutral hydrogen density nhi
double calc_temp_host(const VtkInfo &m, double rho, double p, double nhi) {
  if (m.eos == 3) return p / (2.0 * rho / m.mu0 - nhi) * m.p_sc / m.kb;
  double t = p / rho * m.temp_mu0;
  if (m.eos == 2 && t > m.ion_thres) t = p / rho * m.temp_mui;
  return t;
}

void cells_header(BeWriter &w, long ncells, int npercell) {
  char line[64];
  std::snprintf(line, sizeof line, "CELLS %12ld %12ld", ncells, ncells * (npercell + 1));
  w.text(line);
  for (long c = 0, p = 0; c < ncells; c++, p += npercell) {
    w.i32(npercell);
    for (int q = 0; q < npercell; q++) w.i32((int32_t)(p + q));
  }
  w.lf();
  std::snprintf(line, sizeof line, "CELL_TYPES %12ld", ncells);
  w.text(line);
  for (long c = 0; c < ncells; c++) w.i32(11);
  w.lf();
  std::snprintf(line, sizeof line, "CELL_DATA%12ld", ncells);
  w.text(line);
}

void vtk_head(BeWriter &w, long npoints) {
  char line[64];
  w.text("# vtk DataFile Version 4.2");
  w.text("Output from Walicxe3D");
  w.text("BINARY");
  w.text("DATASET UNSTRUCTURED_GRID");
  std::snprintf(line, sizeof line, "POINTS %12ld float", npoints);
  w.text(line);
}

}  // namespace

int vtk_blocks_write(const char *path, const VtkInfo &m, const double *prim, std::string &err) {
  BeWriter w(path);
  if (!w.ok()) {
    err = std::string("cannot open ") + path;
    return 1;
  }
  const long ncells = (long)m.nblk * NINT;
  const bool ph = m.phys_units != 0;
  vtk_head(w, ncells * 8);
  for (int b = 0; b < m.nblk; b++) {
    double xb[3];
    ref_corner(m, b, xb);
    const int l = m.levels[b] - 1;
    const double dx = m.dxyz[l], dy = m.dxyz[m.maxlev + l], dz = m.dxyz[2 * m.maxlev + l];
    for (int k = 1; k <= NC; k++)
      for (int j = 1; j <= NC; j++)
        for (int i = 1; i <= NC; i++)
          for (int a = 0; a < 2; a++)
            for (int bb = 0; bb < 2; bb++)
              for (int c = 0; c < 2; c++) {
                double x = xb[0] + (double)(i + c - 1) * dx, y = xb[1] + (double)(j + bb - 1) * dy,
                       z = xb[2] + (double)(k + a - 1) * dz;
                if (ph) {
                  x = x * m.l_sc;
                  y = y * m.l_sc;
                  z = z * m.l_sc;
                }
                w.f32((float)x);
                w.f32((float)y);
                w.f32((float)z);
              }
  }
  w.lf();
  cells_header(w, ncells, 8);
  const int nv = m.nv;
  auto P = [&](int b, int v, int q) { return prim[((size_t)b * nv + v) * NINT + q]; };
  w.text("SCALARS rho float 1");
  w.text("LOOKUP_TABLE default");
  for (int b = 0; b < m.nblk; b++)
    for (int q = 0; q < NINT; q++) w.f32((float)(ph ? P(b, 0, q) * m.d_sc : P(b, 0, q)));
  w.lf();
  w.text("VECTORS vel float");
  for (int b = 0; b < m.nblk; b++)
    for (int q = 0; q < NINT; q++)
      for (int v = 1; v <= 3; v++) w.f32((float)(ph ? P(b, v, q) * m.v_sc : P(b, v, q)));
  w.lf();
  w.text("SCALARS P float 1");
  w.text("LOOKUP_TABLE default");
  for (int b = 0; b < m.nblk; b++)
    for (int q = 0; q < NINT; q++) w.f32((float)(ph ? P(b, 4, q) * m.p_sc : P(b, 4, q)));
  w.lf();
  if (nv == 8) {
    w.text("VECTORS B float");
    for (int b = 0; b < m.nblk; b++)
      for (int q = 0; q < NINT; q++)
        for (int v = 5; v <= 7; v++) w.f32((float)(ph ? P(b, v, q) * m.b_sc : P(b, v, q)));
    w.lf();
  }
  w.text("SCALARS T float 1");
  w.text("LOOKUP_TABLE default");
  for (int b = 0; b < m.nblk; b++)
    for (int q = 0; q < NINT; q++)
      w.f32((float)calc_temp_host(m, P(b, 0, q), P(b, 4, q), m.ncarried ? P(b, 5, q) : 0.0));
  w.lf();
  for (int pas = 1; pas <= m.npassive; pas++) {
    char line[64];
    std::snprintf(line, sizeof line, "SCALARS passive%d float 1", pas);
    w.text(line);
    w.text("LOOKUP_TABLE default");
    if (pas <= m.ncarried) {
      for (int b = 0; b < m.nblk; b++)
        for (int q = 0; q < NINT; q++) w.f32((float)P(b, 4 + pas, q));
    } else {
      for (long c = 0; c < ncells; c++) w.f32(0.0f);
    }
    w.lf();
  }
  if (!w.close()) {
    err = std::string("write failed: ") + path;
    return 1;
  }
  return 0;
}

int vtk_grid_write(const char *path, const VtkInfo &m, std::string &err) {
  BeWriter w(path);
  if (!w.ok()) {
    err = std::string("cannot open ") + path;
    return 1;
  }
  const bool ph = m.phys_units != 0;
  vtk_head(w, (long)m.nblk * 8);
  for (int b = 0; b < m.nblk; b++) {
    double xb[3];
    ref_corner(m, b, xb);
    const int l = m.levels[b] - 1;
    const double dx = m.dxyz[l], dy = m.dxyz[m.maxlev + l], dz = m.dxyz[2 * m.maxlev + l];
    for (int k = 0; k < 2; k++)
      for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
          double x = xb[0] + (double)(i * NC) * dx, y = xb[1] + (double)(j * NC) * dy, z = xb[2] + (double)(k * NC) * dz;
          if (ph) {
            x = x * m.l_sc;
            y = y * m.l_sc;
            z = z * m.l_sc;
          }
          w.f32((float)x);
          w.f32((float)y);
          w.f32((float)z);
        }
  }
  w.lf();
  cells_header(w, m.nblk, 8);
  w.text("SCALARS rank integer 1");
  w.text("LOOKUP_TABLE default");
  for (int b = 0; b < m.nblk; b++) w.i32(0);
  w.lf();
  w.text("SCALARS bID integer 1");
  w.text("LOOKUP_TABLE default");
  for (int b = 0; b < m.nblk; b++) w.i32(m.bids[b]);
  w.lf();
  // the Hilbert key goes out as real(hk, 4) under an "integer" header, as
  // the reference writes it
  w.text("SCALARS hkey integer 1");
  w.text("LOOKUP_TABLE default");
  const int order = hilbert_order(m.nroot, m.maxlev);
  for (int b = 0; b < m.nblk; b++)
    w.f32((float)(int)block_hilbert_key(&m.coords[3 * b], m.levels[b], m.maxlev, order));
  w.lf();
  if (!w.close()) {
    err = std::string("write failed: ") + path;
    return 1;
  }
  return 0;
}

int visit_write(const char *path, const char *blocks_name, int noutput, bool replace, std::string &err) {
  File f(path, replace ? "w" : "a");
  if (!f.f) {
    err = std::string("cannot open ") + path;
    return 1;
  }
  bool ok = true;
  if (noutput == 0) ok = std::fprintf(f.f, "!NBLOCKS %5d\n", 1) > 0;
  ok = ok && std::fprintf(f.f, "%s\n", blocks_name) > 0;
  if (!ok) err = std::string("write failed: ") + path;
  return ok ? 0 : 1;
}

}  // namespace wx
