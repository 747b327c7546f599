// wx_api.cpp - the C ABI (include/walicxe_hip.h): context, mesh topology,
// state transfer and the step sequence of both numerical forms.
#include "../../include/walicxe_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "wx_device.h"
#include "wx_fused.h"
#include "wx_hilbert.h"
#include "wx_io.h"
#include "wx_kernels.h"

using namespace wx;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(WX_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// Device scratch that is freed on every return path.
template <typename T>
struct DevBuf {
  T *p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc(&p, n * sizeof(T)); }
  T *release() {
    T *q = p;
    p = nullptr;
    return q;
  }
};

// bID -> index with open addressing (linear probing); one allocation per
// map, where std::unordered_map allocates per entry (the mesh's block
// index and the admesh plan)
struct BidMap {
  std::vector<long> key;
  std::vector<int> val;
  size_t mask = 0;
  BidMap() : BidMap(1) {}
  explicit BidMap(size_t n) {
    size_t cap = 16;
    while (cap < 2 * n) cap <<= 1;
    key.assign(cap, -1);
    val.assign(cap, -1);
    mask = cap - 1;
  }
  static size_t hash(long k) { return (size_t)(((uint64_t)k * 0x9E3779B97F4A7C15ull) >> 20); }
  bool insert(long k, int v) {
    size_t i = hash(k) & mask;
    for (; key[i] != -1; i = (i + 1) & mask)
      if (key[i] == k) return false;
    key[i] = k;
    val[i] = v;
    return true;
  }
  int find(long k) const {
    for (size_t i = hash(k) & mask; key[i] != -1; i = (i + 1) & mask)
      if (key[i] == k) return val[i];
    return -1;
  }
  bool count(long k) const { return find(k) >= 0; }
};

struct Ctx {
  wx_config cfg{};
  std::vector<double> tabT, tabC, winds;
  PlaneWind pw{};                    // wx_set_plane_wind (plain form); plane 0: none
  UserSource us{};                   // wx_set_uniform_gravity (plain form); on 0: none
  DevCfg dc{};
  int device = 0;
  hipStream_t st = nullptr;
  // mesh
  int nblk = 0, maxlev = 0, nroot[3] = {0, 0, 0};
  std::vector<double> dxyz;          // grid spacings per level (wx_set_mesh)
  std::vector<int> bids;
  std::vector<BlockInfo> info;
  BidMap index;                      // bID -> position in bids
  // device memory
  BlockInfo *d_info = nullptr;
  double *d_tabT = nullptr, *d_tabC = nullptr;
  double *dU = nullptr, *dUP = nullptr, *dP = nullptr, *dF = nullptr, *dTmp = nullptr;
  unsigned long long *d_scal = nullptr;  // [0] dt bits, [1] floor count
  bool have_state = false;
  bool ghosts_recorded = false;      // fused form: ghost layer of wx_record_ghosts in d_ghosts
  double *d_ghosts = nullptr;        // [blk][rho, p][6][256]
  double *d_mark = nullptr;          // [blk] largest gradient (wx_mark_blocks)
  FusedState fused;
  // staging of the reference's Fortran layout (fortran_stage_*): a chunk of
  // [v, k] planes in the order [plane][j][i][slot], on the device and in
  // pinned host memory, and the slot -> block map
  double *d_stage = nullptr, *h_stage = nullptr;
  size_t stage_cap = 0;              // doubles
  int *d_slotmap = nullptr;
  int slotmap_cap = 0;
};

void free_mesh(Ctx *c) {
  hipFree(c->d_info);
  hipFree(c->dU);
  hipFree(c->dUP);
  hipFree(c->dP);
  hipFree(c->dF);
  hipFree(c->dTmp);
  c->d_info = nullptr;
  c->dU = c->dUP = c->dP = c->dF = c->dTmp = nullptr;
  fused_free(c->fused);
  hipFree(c->d_ghosts);
  hipFree(c->d_mark);
  c->d_ghosts = nullptr;
  c->d_mark = nullptr;
  c->have_state = false;
  c->ghosts_recorded = false;
}

// the plane wind as the kernels take it (root blocks along z of the mesh)
static PlaneWind plane_wind(const Ctx *c) {
  PlaneWind pw = c->pw;
  pw.nroot_z = c->nroot[2];
  return pw;
}

// ---------------------------------------------------------------------------
// Topology: the face neighbour table (source/amr.f90:439-631 neighborLevel and
// neighbors, :831-909 siblingCoords and children, :916-997 father and
// bcoords), built once per block list with a bID -> index map instead of the
// reference's linear scan of the global block list per query.
// ---------------------------------------------------------------------------
struct Numbering {
  int nroot[3];
  int maxlev;
  std::vector<long> minid, maxid;
  void init(const int nr[3], int ml) {
    for (int d = 0; d < 3; d++) nroot[d] = nr[d];
    maxlev = ml;
    long n = (long)nr[0] * nr[1] * nr[2];
    minid.assign(ml, 0);
    maxid.assign(ml, 0);
    minid[0] = 1;
    maxid[0] = n;
    for (int l = 1; l < ml; l++) {
      minid[l] = maxid[l - 1] + 1;
      maxid[l] = maxid[l - 1] + n * (1L << (3 * l));
    }
  }
  int level(long bid) const {
    for (int l = 0; l < maxlev; l++)
      if (bid >= minid[l] && bid <= maxid[l]) return l + 1;
    return -1;
  }
  void nb(int l, long n[3]) const {
    for (int d = 0; d < 3; d++) n[d] = (long)nroot[d] << (l - 1);
  }
  void coords(long bid, int l, int c[3]) const {
    long n[3];
    nb(l, n);
    long loc = bid - minid[l - 1];
    c[0] = (int)(loc % n[0]) + 1;
    c[1] = (int)((loc / n[0]) % n[1]) + 1;
    c[2] = (int)(loc / (n[0] * n[1])) + 1;
  }
  long bid(int l, const int c[3]) const {
    long n[3];
    nb(l, n);
    return 1 + (c[0] - 1) + (long)(c[1] - 1) * n[0] + (long)(c[2] - 1) * n[0] * n[1] + minid[l - 1] - 1;
  }
};

int build_topology(Ctx *c, const int *bids) {
  Numbering N;
  N.init(c->nroot, c->maxlev);
  c->index = BidMap(c->nblk);
  c->bids.assign(bids, bids + c->nblk);
  for (int b = 0; b < c->nblk; b++) {
    if (N.level(bids[b]) < 0) return fail(WX_ETOPO, "block ID outside the mesh: " + std::to_string(bids[b]));
    if (!c->index.insert(bids[b], b)) return fail(WX_ETOPO, "duplicate block ID " + std::to_string(bids[b]));
  }
  c->info.assign(c->nblk, BlockInfo{});
  for (int b = 0; b < c->nblk; b++) {
    BlockInfo &B = c->info[b];
    B.bid = bids[b];
    B.level = N.level(bids[b]);
    N.coords(bids[b], B.level, B.coords);
    for (int d = 0; d < 3; d++) B.sib[d] = B.level > 1 ? (B.coords[d] - 1) % 2 : -1;
    long nbl[3];
    N.nb(B.level, nbl);
    for (int f = 0; f < 6; f++) {
      Face &F = B.face[f];
      F.type = NEIGH_BOUNDARY;
      for (int q = 0; q < 4; q++) F.src[q] = -1;
      int n = f >> 1, low = (f & 1) == 0;
      int p[3] = {B.coords[0], B.coords[1], B.coords[2]};
      p[n] += low ? -1 : 1;
      bool wrapped = false;
      if (p[n] < 1 || p[n] > nbl[n]) {
        if (c->cfg.bc[f] != BC_PERIODIC) continue;
        p[n] = low ? (int)nbl[n] : 1;  // periodic wrap
        wrapped = true;
      }
      long nid = N.bid(B.level, p);
      const int same = c->index.find(nid);
      if (same >= 0) {
        F.type = NEIGH_SAME;
        F.src[0] = same;
        continue;
      }
      // the reference fills no periodic ghosts (source/boundary.f90:594); the
      // periodic fill it needs copies the block across the face at the same
      // level, and has nothing to copy from when that block is refined or
      // coarsened
      if (wrapped)
        return fail(WX_ETOPO, "periodic face " + std::to_string(f) + " of block " + std::to_string(B.bid) +
                                  ": the block across it is not active at the same level");
      if (B.level > 1) {
        int fp[3] = {(p[0] + 1) / 2, (p[1] + 1) / 2, (p[2] + 1) / 2};
        const int coarser = c->index.find(N.bid(B.level - 1, fp));
        if (coarser >= 0) {
          F.type = NEIGH_COARSER;
          F.src[0] = coarser;
          continue;
        }
      }
      if (B.level >= c->maxlev)
        return fail(WX_ETOPO, "no neighbour for block " + std::to_string(B.bid) + " face " + std::to_string(f));
      int t1 = (n == 0) ? 1 : 0, t2 = (n == 2) ? 1 : 2;
      F.type = NEIGH_FINER;
      for (int s2 = 0; s2 < 2; s2++)
        for (int s1 = 0; s1 < 2; s1++) {
          int ch[3];
          ch[n] = 2 * p[n] - 1 + (low ? 1 : 0);
          ch[t1] = 2 * p[t1] - 1 + s1;
          ch[t2] = 2 * p[t2] - 1 + s2;
          const int finer = c->index.find(N.bid(B.level + 1, ch));
          if (finer < 0)
            return fail(WX_ETOPO, "2:1 level balance violated at block " + std::to_string(B.bid) + " face " +
                                      std::to_string(f));
          F.src[s1 + 2 * s2] = finer;
        }
    }
  }
  return WX_OK;
}

void fill_devcfg(Ctx *c, const double *dxyz) {
  DevCfg &d = c->dc;
  const wx_config &g = c->cfg;
  // the physics takes the flow variables (neq); d.nv counts the stored ones
  // (neqtot), passive scalars included
  const int neq = g.nv == 8 ? 8 : 5;
  d.pd = Phys<double>{neq, g.mhd, g.solver, g.limiter, g.gamma, g.cv};
  d.pf = Phys<float>{neq, g.mhd, g.solver, g.limiter, (float)g.gamma, (float)g.cv};
  d.nv = g.nv;
  d.eight_wave = g.eight_wave;
  d.cooling = g.cooling;
  d.eos = g.eos;
  d.nwind = g.nwind;
  d.ntab = g.ntab;
  for (int f = 0; f < 6; f++) d.bc[f] = g.bc[f];
  d.visc_eta = g.visc_eta;
  d.cfl = g.cfl;
  d.cooling_limit = g.cooling_limit;
  d.ion_thres = g.ion_thres;
  d.d_sc = g.d_sc;
  d.e_sc = g.e_sc;
  d.t_sc = g.t_sc;
  d.p_sc = g.p_sc;
  d.l_sc = g.l_sc;
  d.v_sc = g.v_sc;
  d.temp_mu0 = g.mu0 * AMU * g.p_sc / g.d_sc / KB;
  d.temp_mui = g.mui * AMU * g.p_sc / g.d_sc / KB;
  d.mui_amu = g.mui * AMU;
  for (int w = 0; w < g.nwind; w++)
    for (int q = 0; q < 8; q++) d.winds[w][q] = c->winds[8 * w + q];
  if (dxyz) {
    d.maxlev = c->maxlev;
    for (int a = 0; a < 3; a++)
      for (int l = 0; l < c->maxlev; l++) d.dx[a][l] = dxyz[a * c->maxlev + l];
  }
  d.nblk = c->nblk;
  d.info = c->d_info;
  d.tabT = c->d_tabT;
  d.tabC = c->d_tabC;
}

__global__ void k_scatter(int nv, int nblk, const double *src, double *dst, int zero_ghosts) {
  // one thread per stored cell: dst [blk][nv][20^3] <- src [blk][nv][16^3]
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t total = (size_t)nblk * nv * NCELL;
  if (t >= total) return;
  int q = (int)(t % NCELL);
  size_t bv = t / NCELL;
  int i = q % NB - 1, j = (q / NB) % NB - 1, k = q / (NB * NB) - 1;
  bool in = i >= 1 && i <= NC && j >= 1 && j <= NC && k >= 1 && k <= NC;
  if (in)
    dst[t] = src[bv * NINT + ((k - 1) * NC + (j - 1)) * NC + (i - 1)];
  else if (zero_ghosts)
    dst[t] = 0.0;
}

__global__ void k_gather(int nv, int nblk, const double *src, double *dst) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t total = (size_t)nblk * nv * NINT;
  if (t >= total) return;
  int q = (int)(t % NINT);
  size_t bv = t / NINT;
  int i = q % NC + 1, j = (q / NC) % NC + 1, k = q / (NC * NC) + 1;
  dst[t] = src[bv * NCELL + cid(i, j, k)];
}

int launch_check(const char *what);

// The reference's U(nbmax, nv, -1:18, -1:18, -1:18) keeps the slot index
// fastest, so one block's values are nbmax * nv doubles apart: the device
// reorders a chunk of [v, k] planes between the block layout [blk][nv][16^3]
// and [plane][j][i][slot], and the host moves, for each interior cell and
// variable, the contiguous runs of occupied slots (one memcpy per run)
// between that staging and the Fortran array.
constexpr size_t STAGE_CHUNK = size_t(1) << 23;   // doubles per chunk (64 MB)

__global__ void k_blocks_to_stage(int nv, int nbmax, int plane0, int nplanes, const int *slot2blk,
                                  const double *src, double *stage) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nplanes * NC * NC * nbmax) return;
  const int s = (int)(t % nbmax);
  const size_t rest = t / nbmax;
  const int cell = (int)(rest % (NC * NC));                 // (j - 1) * NC + (i - 1)
  const int pl = plane0 + (int)(rest / (NC * NC));          // v * NC + (k - 1)
  const int v = pl / NC, k0 = pl % NC;
  const int b = slot2blk[s];
  stage[t] = b >= 0 ? src[((size_t)b * nv + v) * NINT + (size_t)k0 * NC * NC + cell] : 0.0;
}

__global__ void k_stage_to_blocks(int nv, int nbmax, int plane0, int nplanes, const int *slot2blk,
                                  const double *stage, double *dst) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nplanes * NC * NC * nbmax) return;
  const int s = (int)(t % nbmax);
  const size_t rest = t / nbmax;
  const int cell = (int)(rest % (NC * NC));
  const int pl = plane0 + (int)(rest / (NC * NC));
  const int v = pl / NC, k0 = pl % NC;
  const int b = slot2blk[s];
  if (b >= 0) dst[((size_t)b * nv + v) * NINT + (size_t)k0 * NC * NC + cell] = stage[t];
}

// slot -> block map on the device and the runs of occupied slots; staging
// buffers grown to one chunk
static int stage_prepare(Ctx *c, int nbmax, const int *lb, std::vector<std::pair<int, int>> &runs, int &planes) {
  std::vector<int> map(nbmax);
  runs.clear();
  for (int s = 0; s < nbmax; s++) {
    map[s] = lb[s] == -1 ? -1 : c->index.find(lb[s]);
    if (map[s] < 0) continue;
    if (!runs.empty() && runs.back().second == s) runs.back().second = s + 1;
    else runs.push_back({s, s + 1});
  }
  const size_t plane = (size_t)NC * NC * nbmax;
  planes = (int)std::max<size_t>(1, std::min<size_t>((size_t)c->cfg.nv * NC, STAGE_CHUNK / plane));
  const size_t need = plane * planes;
  if (need > c->stage_cap) {
    hipFree(c->d_stage);
    hipHostFree(c->h_stage);
    c->d_stage = nullptr;
    c->h_stage = nullptr;
    c->stage_cap = 0;
    HIPCHK(hipMalloc(&c->d_stage, need * sizeof(double)));
    HIPCHK(hipHostMalloc(&c->h_stage, need * sizeof(double), hipHostMallocDefault));
    c->stage_cap = need;
  }
  if (nbmax > c->slotmap_cap) {
    hipFree(c->d_slotmap);
    c->d_slotmap = nullptr;
    c->slotmap_cap = 0;
    HIPCHK(hipMalloc(&c->d_slotmap, nbmax * sizeof(int)));
    c->slotmap_cap = nbmax;
  }
  HIPCHK(hipMemcpy(c->d_slotmap, map.data(), nbmax * sizeof(int), hipMemcpyHostToDevice));
  return WX_OK;
}

// offset of U(s, v, i, j, k) in the Fortran array (0-based slot s)
static inline size_t fortran_offset(int nbmax, int nv, int s, int v, int i, int j, int k) {
  return (size_t)s + (size_t)nbmax * (v + (size_t)nv * ((i + 1) + NB * ((j + 1) + (size_t)NB * (k + 1))));
}

// device block layout -> Fortran array on the host (interior cells of the
// occupied slots; nothing else of the array is written)
static int fortran_stage_out(Ctx *c, const double *d_src, double *out, int nbmax, const int *lb) {
  const int nv = c->cfg.nv;
  std::vector<std::pair<int, int>> runs;
  int planes;
  int rc = stage_prepare(c, nbmax, lb, runs, planes);
  if (rc != WX_OK) return rc;
  for (int p0 = 0; p0 < nv * NC; p0 += planes) {
    const int np = std::min(planes, nv * NC - p0);
    const size_t n = (size_t)np * NC * NC * nbmax;
    hipLaunchKernelGGL(k_blocks_to_stage, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, nv, nbmax, p0, np,
                       c->d_slotmap, d_src, c->d_stage);
    HIPCHK(hipMemcpyAsync(c->h_stage, c->d_stage, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for (int pl = 0; pl < np; pl++) {
      const int v = (p0 + pl) / NC, k = (p0 + pl) % NC + 1;
      for (int j = 1; j <= NC; j++)
        for (int i = 1; i <= NC; i++) {
          const double *row = c->h_stage + ((size_t)(pl * NC + (j - 1)) * NC + (i - 1)) * nbmax;
          double *dst = out + fortran_offset(nbmax, nv, 0, v, i, j, k);
          for (const auto &r : runs) std::memcpy(dst + r.first, row + r.first, (r.second - r.first) * sizeof(double));
        }
    }
  }
  return launch_check("Fortran layout");
}

// Fortran array on the host -> device block layout (blocks of the context
// with no slot in lb are zero)
static int fortran_stage_in(Ctx *c, const double *in, double *d_dst, int nbmax, const int *lb) {
  const int nv = c->cfg.nv;
  std::vector<std::pair<int, int>> runs;
  int planes;
  int rc = stage_prepare(c, nbmax, lb, runs, planes);
  if (rc != WX_OK) return rc;
  HIPCHK(hipMemsetAsync(d_dst, 0, (size_t)c->nblk * nv * NINT * sizeof(double), c->st));
  for (int p0 = 0; p0 < nv * NC; p0 += planes) {
    const int np = std::min(planes, nv * NC - p0);
    const size_t n = (size_t)np * NC * NC * nbmax;
    // the previous chunk's copy out of h_stage is done before it is refilled
    HIPCHK(hipStreamSynchronize(c->st));
    for (int pl = 0; pl < np; pl++) {
      const int v = (p0 + pl) / NC, k = (p0 + pl) % NC + 1;
      for (int j = 1; j <= NC; j++)
        for (int i = 1; i <= NC; i++) {
          double *row = c->h_stage + ((size_t)(pl * NC + (j - 1)) * NC + (i - 1)) * nbmax;
          const double *src = in + fortran_offset(nbmax, nv, 0, v, i, j, k);
          for (const auto &r : runs) std::memcpy(row + r.first, src + r.first, (r.second - r.first) * sizeof(double));
        }
    }
    HIPCHK(hipMemcpyAsync(c->d_stage, c->h_stage, n * sizeof(double), hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL(k_stage_to_blocks, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, nv, nbmax, p0, np,
                       c->d_slotmap, c->d_stage, d_dst);
  }
  HIPCHK(hipStreamSynchronize(c->st));
  return launch_check("Fortran layout");
}

int launch_check(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(WX_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return WX_OK;
}

// ------------------------- plain-form sequences ----------------------------
int plain_timestep_raw(Ctx *c, double *dt) {
  HIPCHK(hipMemsetAsync(c->d_scal, 0xFF, sizeof(unsigned long long), c->st));
  plain_timestep(c->dc, c->dP, c->d_scal, c->st);
  unsigned long long bits = 0;
  HIPCHK(hipMemcpyAsync(&bits, c->d_scal, sizeof(bits), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  double m;
  std::memcpy(&m, &bits, sizeof(m));
  // the reference's dt_loc = dt_hydro*CFL (or min(dt_hydro, dt_cool)*CFL,
  // dt_cool = 1e30 unused)
  if (!c->cfg.cooling)
    *dt = m * c->cfg.cfl;
  else
    *dt = std::fmin(m, 1.0e30) * c->cfg.cfl;
  return WX_OK;
}

int plain_godunov(Ctx *c, int order, double dt) {
  double dtp = order == 2 ? dt / 2.0 : dt;
  plain_boundary(c->dc, plane_wind(c), c->dU, 1, c->st);
  plain_prims(c->dc, c->dU, c->dP, WX_CELLS_GHOST, nullptr, c->st);
  plain_fluxes(c->dc, c->dP, c->dF, 1, c->st);
  plain_upwind(c->dc, c->us, c->dU, c->dP, c->dF, c->dUP, dtp, c->st);
  if (order == 2) {
    plain_boundary(c->dc, plane_wind(c), c->dUP, 2, c->st);
    plain_prims(c->dc, c->dUP, c->dP, WX_CELLS_ALL, nullptr, c->st);
    plain_fluxes(c->dc, c->dP, c->dF, 2, c->st);
    plain_upwind(c->dc, c->us, c->dU, c->dP, c->dF, c->dUP, dt, c->st);
  }
  return launch_check("godunov");
}

int update_prims(Ctx *c, long *bad) {
  HIPCHK(hipMemsetAsync(c->d_scal + 1, 0, sizeof(unsigned long long), c->st));
  plain_prims(c->dc, c->dU, c->dP, WX_CELLS_ALL, c->d_scal + 1, c->st);
  if (bad) {
    unsigned long long n = 0;
    HIPCHK(hipMemcpyAsync(&n, c->d_scal + 1, sizeof(n), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    *bad = (long)n;
  }
  return launch_check("prims");
}

Ctx *as_ctx(void *p) { return static_cast<Ctx *>(p); }
bool is_fused(const Ctx *c) { return c->cfg.form != WX_FORM_PLAIN_FP64; }

int need_state(Ctx *c) {
  if (!c) return fail(WX_EINVAL, "null context");
  if (!c->d_info) return fail(WX_ESTATE, "no mesh: call wx_set_mesh first");
  if (!c->have_state) return fail(WX_ESTATE, "no state: call wx_upload_state first");
  return WX_OK;
}

}  // namespace


// ---------------------------------------------------------------------------
// Remesh on the device: carries the interior state from the current block
// list to a new one, with the data motion of the reference's refineBlock
// (source/amr.f90:284-384: the father's octant copied into each child) and
// coarsenBlock (:1430-1650: each father cell the mean of its 2x2x2 children
// cells), without a round trip through host memory.
// ---------------------------------------------------------------------------
struct RemapSrc {
  int kind;          // 0 copy, 1 injection from an ancestor, 2 mean of 8 children
  int src;           // old block index (copy, injection)
  int shift;         // levels between the new block and its ancestor
  int f0[3];         // (coords-1)*16 of the new block at its level
  int a0[3];         // (coords-1)*16 of the ancestor at its level
  int kids[8];       // old indices of the children, x fastest then y, z
};

__global__ void __launch_bounds__(256) k_remap(int nv, const RemapSrc *map, const double *src, double *dst) {
  const int b = blockIdx.x, v = blockIdx.y;
  const int i = threadIdx.x & 15, j = threadIdx.x >> 4;
  const RemapSrc m = map[b];
  double *D = dst + ((size_t)b * nv + v) * NINT;
  for (int k = 0; k < NC; k++) {
    const int q = (k * NC + j) * NC + i;
    double val;
    if (m.kind == 0) {
      val = src[((size_t)m.src * nv + v) * NINT + q];
    } else if (m.kind == 1) {
      const int ai = ((m.f0[0] + i) >> m.shift) - m.a0[0];
      const int aj = ((m.f0[1] + j) >> m.shift) - m.a0[1];
      const int ak = ((m.f0[2] + k) >> m.shift) - m.a0[2];
      val = src[((size_t)m.src * nv + v) * NINT + (ak * NC + aj) * NC + ai];
    } else {
      const int c = (i >> 3) + 2 * (j >> 3) + 4 * (k >> 3);
      const double *S = src + ((size_t)m.kids[c] * nv + v) * NINT;
      const int i0 = 2 * (i & 7), j0 = 2 * (j & 7), k0 = 2 * (k & 7);
      // z outermost, x innermost
      double acc = 0.0;
      for (int dk = 0; dk < 2; dk++)
        for (int dj = 0; dj < 2; dj++)
          for (int di = 0; di < 2; di++) acc = acc + S[((k0 + dk) * NC + (j0 + dj)) * NC + (i0 + di)];
      val = acc / 8.0;
    }
    D[q] = val;
  }
}

// Source of every new block; fails when a new block has neither an old
// copy, an old ancestor, nor all 8 children in the old list.
static int remap_table(Ctx *c, int nblk, const int *bids, std::vector<RemapSrc> &out) {
  Numbering N;
  N.init(c->nroot, c->maxlev);
  out.assign(nblk, RemapSrc{});
  for (int n = 0; n < nblk; n++) {
    RemapSrc &m = out[n];
    const long b = bids[n];
    const int lev = N.level(b);
    if (lev < 0) return fail(WX_ETOPO, "block ID outside the mesh: " + std::to_string(b));
    const int same = c->index.find(b);
    if (same >= 0) {
      m.kind = 0;
      m.src = same;
      continue;
    }
    int bc[3];
    N.coords(b, lev, bc);
    // nearest old ancestor
    int p[3] = {bc[0], bc[1], bc[2]};
    bool found = false;
    for (int l = lev - 1, d = 1; l >= 1; l--, d++) {
      for (int a = 0; a < 3; a++) p[a] = (p[a] + 1) / 2;
      const int anc = c->index.find(N.bid(l, p));
      if (anc >= 0) {
        m.kind = 1;
        m.src = anc;
        m.shift = d;
        for (int a = 0; a < 3; a++) {
          m.f0[a] = (bc[a] - 1) * NC;
          m.a0[a] = (p[a] - 1) * NC;
        }
        found = true;
        break;
      }
    }
    if (found) continue;
    if (lev >= c->maxlev) return fail(WX_ETOPO, "no source for block " + std::to_string(b));
    m.kind = 2;
    for (int q = 0; q < 8; q++) {
      int ch[3] = {2 * bc[0] - 1 + (q & 1), 2 * bc[1] - 1 + ((q >> 1) & 1), 2 * bc[2] - 1 + ((q >> 2) & 1)};
      const int kid = c->index.find(N.bid(lev + 1, ch));
      if (kid < 0)
        return fail(WX_ETOPO, "block " + std::to_string(b) + " needs a restriction over more than one level");
      m.kids[q] = kid;
    }
  }
  return WX_OK;
}

// PRIM of the new blocks that the reference carries through a mesh change:
// copied blocks keep theirs, refined children take their father's by the
// same injection as U (refineBlock, source/amr.f90:355-365); new coarse
// blocks (kind 2) keep the PRIM computed from their averaged U.
__global__ void __launch_bounds__(256) k_carry_prims(int nv, const RemapSrc *map, const double *src, double *P) {
  const int b = blockIdx.x, v = blockIdx.y;
  if (map[b].kind == 2) return;
  const int i = threadIdx.x & 15, j = threadIdx.x >> 4;
  const double *S = src + ((size_t)b * nv + v) * NINT;
  double *D = P + ((size_t)b * nv + v) * NCELL;
  for (int k = 0; k < NC; k++) D[cid(i + 1, j + 1, k + 1)] = S[(k * NC + j) * NC + i];
}

// ---------------------------------------------------------------------------
// Mesh adaptation, admesh (source/amr.f90:44-133) on one rank: the flags
// come from the device (k_mark); the proximity rule, the sibling families and
// the new block list are host work on block IDs, with a bID -> index map in
// place of the reference's scans of the global block list.
// ---------------------------------------------------------------------------
enum { FLAG_NONE = 0, FLAG_REFINE = 1, FLAG_COARSE = 2 };   // constants.f90:76-78


static long amr_father(const Numbering &N, long bid) {
  const int lev = N.level(bid);
  if (lev <= 1) return -1;
  int c[3];
  N.coords(bid, lev, c);
  int f[3] = {(c[0] + 1) / 2, (c[1] + 1) / 2, (c[2] + 1) / 2};
  return N.bid(lev - 1, f);
}

// children (source/amr.f90:882-909): x fastest, then y, then z
static long amr_child(const Numbering &N, long bid, int q) {
  const int lev = N.level(bid);
  int c[3];
  N.coords(bid, lev, c);
  int ch[3] = {2 * c[0] - 1 + (q & 1), 2 * c[1] - 1 + ((q >> 1) & 1), 2 * c[2] - 1 + ((q >> 2) & 1)};
  return N.bid(lev + 1, ch);
}

// neighbors (source/amr.f90:539-631) over the active set: NEIGH_BOUNDARY at
// every domain face (neighborLevel, :439-523, does not wrap periodic faces);
// for NEIGH_SAME and NEIGH_COARSER *id is the neighbour, for NEIGH_FINER the
// inactive same-level block whose children face bid.
static int amr_neighbour(const Numbering &N, const BidMap &act, long bid, int dir, long *id) {
  const int lev = N.level(bid);
  int c[3];
  N.coords(bid, lev, c);
  long n[3];
  N.nb(lev, n);
  const int ax = dir >> 1;
  c[ax] += (dir & 1) ? 1 : -1;
  if (c[ax] < 1 || c[ax] > n[ax]) return NEIGH_BOUNDARY;
  const long nid = N.bid(lev, c);
  *id = nid;
  if (act.count(nid)) return NEIGH_SAME;
  const long fid = amr_father(N, nid);
  if (fid != -1 && act.count(fid)) {
    *id = fid;
    return NEIGH_COARSER;
  }
  return NEIGH_FINER;
}

// checkProximity(maxlev - 1) (source/amr.f90:1109-1254) on `flag`, then the
// block list after refineBlock (:284-384) of every block flagged for
// refinement and coarseBlocks (:1430-1650) of every family whose 8 members
// are flagged for coarsening.  The new list keeps the old order: the 8
// children in place of their father, a father in place of the first of its
// children in the list.  Within one level the proximity result does not
// depend on the order of the blocks: a block's own test reads only flags of
// finer levels, which are final by then, and the same-level changes only
// turn COARSE into NONE.
static int amr_plan(const Numbering &N, const std::vector<int> &bids, std::vector<int> &flag, std::vector<int> &out) {
  const size_t nb = bids.size();
  BidMap act(nb);
  std::vector<std::vector<int>> by_level(N.maxlev + 1);
  for (size_t n = 0; n < nb; n++) {
    const int lev = N.level(bids[n]);
    if (lev < 0) return fail(WX_ETOPO, "block ID outside the mesh: " + std::to_string(bids[n]));
    if (!act.insert(bids[n], (int)n)) return fail(WX_ETOPO, "duplicate block ID " + std::to_string(bids[n]));
    if (flag[n] < FLAG_NONE || flag[n] > FLAG_COARSE) return fail(WX_EINVAL, "unknown refinement flag");
    by_level[lev].push_back((int)n);
  }
  for (int level = N.maxlev - 1; level >= 1; level--)
    for (int n : by_level[level]) {
      const long b = bids[n];
      const int bflag = flag[n];
      if (bflag == FLAG_REFINE) {
        // same-level neighbours stop coarsening, coarser ones must refine
        for (int dir = 0; dir < 6; dir++) {
          long id = -1;
          const int t = amr_neighbour(N, act, b, dir, &id);
          if (t == NEIGH_SAME) {
            int &f = flag[act.find(id)];
            if (f == FLAG_COARSE) f = FLAG_NONE;
          } else if (t == NEIGH_COARSER) {
            flag[act.find(id)] = FLAG_REFINE;
          }
        }
      } else if (bflag == FLAG_COARSE) {
        // no coarsening next to a finer family with a member not coarsening
        bool inhibited = false;
        for (int dir = 0; dir < 6 && !inhibited; dir++) {
          long id = -1;
          if (amr_neighbour(N, act, b, dir, &id) != NEIGH_FINER) continue;
          for (int q = 0; q < 8 && !inhibited; q++) {
            const int m = act.find(amr_child(N, id, q));
            if (m >= 0 && flag[m] != FLAG_COARSE) {
              flag[n] = FLAG_NONE;
              inhibited = true;
            }
          }
        }
      }
    }
  BidMap family(nb);                       // father -> slot of its count
  std::vector<int> flagged;                 // members flagged, per family
  for (size_t n = 0; n < nb; n++)
    if (flag[n] == FLAG_COARSE) {
      const long f = amr_father(N, bids[n]);
      if (f == -1) continue;
      int s = family.find(f);
      if (s < 0) {
        s = (int)flagged.size();
        family.insert(f, s);
        flagged.push_back(0);
      }
      flagged[s]++;
    }
  out.clear();
  out.reserve(nb);
  std::vector<char> placed(flagged.size(), 0);
  for (size_t n = 0; n < nb; n++) {
    const long b = bids[n];
    if (flag[n] == FLAG_REFINE) {
      if (N.level(b) >= N.maxlev)
        return fail(WX_ETOPO, "block " + std::to_string(b) + " is flagged for refinement at the maximum level");
      for (int q = 0; q < 8; q++) out.push_back((int)amr_child(N, b, q));
      continue;
    }
    if (flag[n] == FLAG_COARSE) {
      const long f = amr_father(N, b);
      const int s = f == -1 ? -1 : family.find(f);
      if (s >= 0 && flagged[s] == 8) {
        if (!placed[s]) {
          placed[s] = 1;
          out.push_back((int)f);
        }
        continue;
      }
    }
    out.push_back((int)b);
  }
  return WX_OK;
}

// face table of a block list without touching a context (2:1 balance and
// the periodic partners checked as wx_set_mesh does)
static int check_mesh(const Ctx *c, const std::vector<int> &bids, Ctx &t) {
  t.cfg = c->cfg;
  t.maxlev = c->maxlev;
  for (int d = 0; d < 3; d++) t.nroot[d] = c->nroot[d];
  t.nblk = (int)bids.size();
  return build_topology(&t, bids.data());
}

// markByPhysical flags of every block of the context.  The fused form keeps
// no ghost layers: its kernel forms the depth-1 ghosts of the current state
// (or reads the layer wx_record_ghosts kept).
static int mark_blocks(Ctx *c, double refine_thres, double coarse_thres, int *flags, double *maxgrad) {
  std::vector<double> m(c->nblk);
  if (!c->d_mark) HIPCHK(hipMalloc(&c->d_mark, m.size() * sizeof(double)));
  if (is_fused(c))
    fused_mark(c->dc, c->fused.U, c->ghosts_recorded ? c->d_ghosts : nullptr, c->d_mark, c->st);
  else
    plain_mark(c->dc, c->dP, c->d_mark, c->st);
  int rc = launch_check("mark");
  if (rc != WX_OK) return rc;
  HIPCHK(hipMemcpyAsync(m.data(), c->d_mark, m.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int n = 0; n < c->nblk; n++) {
    const int lev = c->info[n].level;
    if (m[n] >= refine_thres)
      flags[n] = lev < c->maxlev ? FLAG_REFINE : FLAG_NONE;
    else
      flags[n] = (m[n] <= coarse_thres && lev > 1) ? FLAG_COARSE : FLAG_NONE;
    if (maxgrad) maxgrad[n] = m[n];
  }
  return WX_OK;
}

extern "C" {

const char *wx_last_error(void) { return g_err.c_str(); }
const char *wx_version(void) { return "walicxe3d-amd 0.1 (gfx950)"; }

int wx_init(const wx_config *cfg, void **out) {
  if (!cfg || !out) return fail(WX_EINVAL, "null argument");
  if (cfg->nv < 5 || cfg->nv > 8) return fail(WX_EINVAL, "nv must be 5 or 8, or 6..7 (hydro with passive scalars)");
  if (cfg->solver < 2 || cfg->solver > 6) return fail(WX_EINVAL, "unsupported solver");
  // passive scalars (neqtot = 5 + npassive, hydro_core.f90:307,372,429,
  // hllc.f90:262-285): the plain form's hydro kernels carry one or two
  if (cfg->nv == 6 || cfg->nv == 7) {
    if (cfg->form != WX_FORM_PLAIN_FP64)
      return fail(WX_EINVAL, "passive scalars need the plain FP64 form (the fused forms carry the flow variables)");
    if (cfg->mhd) return fail(WX_EINVAL, "passive scalars are carried for hydro (nv = 5 + npassive) only");
    if (cfg->nwind)
      return fail(WX_EINVAL, "wind sources with passive scalars: the reference leaves the passives of the wind state unset");
  }
  if ((cfg->solver == 5 || cfg->solver == 6) && cfg->nv != 8) return fail(WX_EINVAL, "HLLE/HLLD need nv = 8");
  if (cfg->limiter < -1 || cfg->limiter > 6) return fail(WX_EINVAL, "unknown limiter");
  if (cfg->cooling && (cfg->ntab < 2 || cfg->ntab > MAXTAB || !cfg->tab_logT || !cfg->tab_loglam))
    return fail(WX_EINVAL, "cooling needs a table of 2..512 points");
  if (cfg->nwind < 0 || cfg->nwind > MAXWIND || (cfg->nwind && !cfg->winds))
    return fail(WX_EINVAL, "up to 8 wind sources");
  if (cfg->form != WX_FORM_PLAIN_FP64 && cfg->form != WX_FORM_FUSED_FP32 && cfg->form != WX_FORM_FUSED_FP64)
    return fail(WX_EINVAL, "unknown form");
  if (cfg->form != WX_FORM_PLAIN_FP64 && cfg->solver == SOLVER_HLL1)
    return fail(WX_EINVAL, "the fused form runs the two-stage scheme (SOLVER_HLL1 needs the plain form)");
  // the fused form instantiates HLL/HLLC for the 5 hydro variables and
  // HLLE/HLLD for the 8 MHD variables only
  if (cfg->form != WX_FORM_PLAIN_FP64 && cfg->nv == 8 && (cfg->solver != SOLVER_HLLE && cfg->solver != SOLVER_HLLD))
    return fail(WX_EINVAL, "the fused form with nv = 8 runs HLLE or HLLD");
  if (cfg->form != WX_FORM_PLAIN_FP64 && cfg->nv == 8 && !cfg->mhd)
    return fail(WX_EINVAL, "the fused form with nv = 8 needs mhd = 1");
  // calcTemp (source/hydro_core.f90:231-270): EOS_H_RATE (3) takes the
  // neutral hydrogen density from the first passive scalar (plain form)
  if (cfg->eos < 0 || cfg->eos > 3)
    return fail(WX_EINVAL, "eos must be EOS_ADIABATIC, EOS_SINGLE_SPECIE, EOS_TWOTEMP or EOS_H_RATE");
  if (cfg->eos == 3 && cfg->nv != 6 && cfg->nv != 7)
    return fail(WX_EINVAL, "EOS_H_RATE reads the neutral hydrogen density from the first passive scalar "
                           "(hydro, nv = 6 or 7, plain FP64 form)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(WX_EHIP, "no HIP device");
  Ctx *c = new Ctx();
  c->cfg = *cfg;
  if (cfg->cooling) {
    c->tabT.assign(cfg->tab_logT, cfg->tab_logT + cfg->ntab);
    c->tabC.assign(cfg->tab_loglam, cfg->tab_loglam + cfg->ntab);
    for (int i = 1; i < cfg->ntab; i++)
      if (!(c->tabT[i] > c->tabT[i - 1])) {
        delete c;
        return fail(WX_EINVAL, "cooling table temperatures must increase");
      }
  }
  if (cfg->nwind) c->winds.assign(cfg->winds, cfg->winds + 8 * cfg->nwind);
  c->cfg.tab_logT = c->tabT.empty() ? nullptr : c->tabT.data();
  c->cfg.tab_loglam = c->tabC.empty() ? nullptr : c->tabC.data();
  c->cfg.winds = c->winds.empty() ? nullptr : c->winds.data();
  hipGetDevice(&c->device);
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(WX_EHIP, "stream creation failed");
  }
  if (hipMalloc(&c->d_scal, 4 * sizeof(unsigned long long)) != hipSuccess) {
    hipStreamDestroy(c->st);
    delete c;
    return fail(WX_ENOMEM, "device allocation failed");
  }
  if (cfg->cooling) {
    hipMalloc(&c->d_tabT, sizeof(double) * cfg->ntab);
    hipMalloc(&c->d_tabC, sizeof(double) * cfg->ntab);
    hipMemcpy(c->d_tabT, c->tabT.data(), sizeof(double) * cfg->ntab, hipMemcpyHostToDevice);
    hipMemcpy(c->d_tabC, c->tabC.data(), sizeof(double) * cfg->ntab, hipMemcpyHostToDevice);
  }
  fill_devcfg(c, nullptr);
  *out = c;
  return WX_OK;
}

int wx_finalize(void *ctx) {
  Ctx *c = as_ctx(ctx);
  if (!c) return WX_OK;
  hipStreamSynchronize(c->st);
  free_mesh(c);
  hipFree(c->d_scal);
  hipFree(c->d_tabT);
  hipFree(c->d_tabC);
  hipFree(c->d_stage);
  hipHostFree(c->h_stage);
  hipFree(c->d_slotmap);
  hipStreamDestroy(c->st);
  delete c;
  return WX_OK;
}

// wx_set_mesh; `pre`, when given, holds the face table of this block list
// already built (check_mesh with the context's nroot, maxlev and boundaries),
// which is taken over instead of built again
static int set_mesh_impl(Ctx *c, const int nroot[3], int maxlev, const double *dxyz, int nblk, const int *bids,
                         Ctx *pre) {
  HIPCHK(hipStreamSynchronize(c->st));
  free_mesh(c);
  c->nblk = nblk;
  c->maxlev = maxlev;
  for (int d = 0; d < 3; d++) c->nroot[d] = nroot[d];
  if (dxyz != c->dxyz.data()) c->dxyz.assign(dxyz, dxyz + 3 * maxlev);
  int rc = WX_OK;
  if (pre) {
    c->bids = std::move(pre->bids);
    c->info = std::move(pre->info);
    c->index = std::move(pre->index);
  } else {
    rc = build_topology(c, bids);
  }
  if (rc != WX_OK) return rc;
  size_t n = (size_t)nblk * c->cfg.nv * NCELL;
  HIPCHK(hipMalloc(&c->d_info, sizeof(BlockInfo) * nblk));
  HIPCHK(hipMemcpy(c->d_info, c->info.data(), sizeof(BlockInfo) * nblk, hipMemcpyHostToDevice));
  if (!is_fused(c)) {
    // the plain form's arrays with ghost layers; the fused form keeps its
    // state without ghosts (FusedState) and needs none of them
    HIPCHK(hipMalloc(&c->dU, n * sizeof(double)));
    HIPCHK(hipMalloc(&c->dP, n * sizeof(double)));
    HIPCHK(hipMemset(c->dP, 0, n * sizeof(double)));
    HIPCHK(hipMalloc(&c->dUP, n * sizeof(double)));
    HIPCHK(hipMemset(c->dUP, 0, n * sizeof(double)));
    HIPCHK(hipMalloc(&c->dF, plain_flux_elems(nblk, c->cfg.nv) * sizeof(double)));
  }
  fill_devcfg(c, dxyz);
  if (is_fused(c)) {
    c->fused.fp64 = c->cfg.form == WX_FORM_FUSED_FP64;
    for (int d = 0; d < 3; d++) c->fused.nroot[d] = c->nroot[d];
    int frc = fused_alloc(c->fused, c->dc, c->info, c->st);
    if (frc != 0) return fail(WX_ENOMEM, "fused-form allocation failed");
  }
  return WX_OK;
}

int wx_set_mesh(void *ctx, const int nroot[3], int maxlev, const double *dxyz, int nblk, const int *bids) {
  Ctx *c = as_ctx(ctx);
  if (!c || !nroot || !dxyz || !bids) return fail(WX_EINVAL, "null argument");
  if (maxlev < 1 || maxlev > MAXLEV || nblk < 1) return fail(WX_EINVAL, "bad mesh size");
  return set_mesh_impl(c, nroot, maxlev, dxyz, nblk, bids, nullptr);
}

// Interior state [nblk][nv][16^3] already on the device into the context:
// the fused state and its first dt candidate (copied, or with adopt the
// buffer itself, which the context then owns), or the plain form's ghosted
// U with zero ghosts and its PRIM.
static int upload_device(Ctx *c, double *d_in, bool adopt = false) {
  int nv = c->cfg.nv;
  c->ghosts_recorded = false;
  if (is_fused(c)) {
    if (fused_upload(c->fused, c->dc, d_in, adopt, c->st)) return fail(WX_EHIP, fused_error());
    c->have_state = true;
    return WX_OK;
  }
  size_t total = (size_t)c->nblk * nv * NCELL;
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->st, nv, c->nblk, d_in, c->dU, 1);
  int rc = launch_check("upload");
  HIPCHK(hipStreamSynchronize(c->st));
  if (adopt) hipFree(d_in);
  if (rc != WX_OK) return rc;
  c->have_state = true;
  long bad = 0;
  rc = update_prims(c, &bad);
  if (rc != WX_OK) return rc;
  return (int)bad;
}

int wx_upload_state(void *ctx, const double *u, int layout, int nbmax, const int *lb) {
  Ctx *c = as_ctx(ctx);
  if (!c || !u) return fail(WX_EINVAL, "null argument");
  if (!c->d_info) return fail(WX_ESTATE, "no mesh");
  int nv = c->cfg.nv;
  size_t ni = (size_t)c->nblk * nv * NINT;
  if (layout == WX_LAYOUT_FORTRAN) {
    if (!lb || nbmax < 1) return fail(WX_EINVAL, "Fortran layout needs nbmax and localBlocks");
  } else if (layout != WX_LAYOUT_BLOCKS) {
    return fail(WX_EINVAL, "unknown layout");
  }
  DevBuf<double> d_in;
  HIPCHK(d_in.alloc(ni));
  if (layout == WX_LAYOUT_FORTRAN) {
    int rc = fortran_stage_in(c, u, d_in.p, nbmax, lb);
    if (rc != WX_OK) return rc;
  } else {
    HIPCHK(hipMemcpyAsync(d_in.p, u, ni * sizeof(double), hipMemcpyHostToDevice, c->st));
  }
  return upload_device(c, d_in.p);
}

int wx_upload_prims(void *ctx, const double *p, int layout, int nbmax, const int *lb) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (!p) return fail(WX_EINVAL, "null argument");
  if (is_fused(c)) return fail(WX_EINVAL, "plain-form entry point");
  int nv = c->cfg.nv;
  size_t ni = (size_t)c->nblk * nv * NINT;
  if (layout == WX_LAYOUT_FORTRAN) {
    if (!lb || nbmax < 1) return fail(WX_EINVAL, "Fortran layout needs nbmax and localBlocks");
  } else if (layout != WX_LAYOUT_BLOCKS) {
    return fail(WX_EINVAL, "unknown layout");
  }
  DevBuf<double> d_in;
  HIPCHK(d_in.alloc(ni));
  if (layout == WX_LAYOUT_FORTRAN) {
    rc = fortran_stage_in(c, p, d_in.p, nbmax, lb);
    if (rc != WX_OK) return rc;
  } else {
    HIPCHK(hipMemcpyAsync(d_in.p, p, ni * sizeof(double), hipMemcpyHostToDevice, c->st));
  }
  size_t total = (size_t)c->nblk * nv * NCELL;
  const double *in = d_in.p;
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->st, nv, c->nblk, in, c->dP, 0);
  rc = launch_check("upload_prims");
  HIPCHK(hipStreamSynchronize(c->st));
  return rc;
}

int wx_download_state(void *ctx, double *u, double *prim, int layout, int nbmax, const int *lb) {
  if (layout == WX_LAYOUT_FORTRAN) {
    if (!lb || nbmax < 1) return fail(WX_EINVAL, "Fortran layout needs nbmax and localBlocks");
  } else if (layout != WX_LAYOUT_BLOCKS) {
    return fail(WX_EINVAL, "unknown layout");
  }
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  int nv = c->cfg.nv;
  size_t ni = (size_t)c->nblk * nv * NINT;
  DevBuf<double> d_out;
  for (int w = 0; w < 2; w++) {
    double *dst = w == 0 ? u : prim;
    if (!dst) continue;
    const double *from;
    if (is_fused(c) && w == 0) {
      // the fused state is already in the block layout of the interface
      from = c->fused.U;
    } else {
      if (!d_out.p) HIPCHK(d_out.alloc(ni));
      double *out = d_out.p;
      if (is_fused(c)) {
        if (fused_prims(c->fused, c->dc, out, c->st)) return fail(WX_EHIP, fused_error());
      } else {
        const double *g = w == 0 ? c->dU : c->dP;
        hipLaunchKernelGGL(k_gather, dim3((unsigned)((ni + 255) / 256)), dim3(256), 0, c->st, nv, c->nblk, g, out);
      }
      from = out;
    }
    if (layout == WX_LAYOUT_FORTRAN) {
      rc = fortran_stage_out(c, from, dst, nbmax, lb);
      if (rc != WX_OK) return rc;
    } else {
      HIPCHK(hipMemcpyAsync(dst, from, ni * sizeof(double), hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
    }
  }
  return launch_check("download");
}

int wx_get_timestep(void *ctx, double *dt_raw) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (!dt_raw) return fail(WX_EINVAL, "null dt");
  if (is_fused(c)) return fused_timestep(c->fused, c->dc, c->st, dt_raw) ? WX_EHIP : WX_OK;
  return plain_timestep_raw(c, dt_raw);
}

int wx_godunov(void *ctx, int order, double dt) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (order != 1 && order != 2) return fail(WX_EINVAL, "order must be 1 or 2");
  if (is_fused(c)) return fail(WX_EINVAL, "wx_godunov is a plain-form entry point");
  return plain_godunov(c, order, dt);
}

int wx_viscous_step(void *ctx) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (is_fused(c)) return fail(WX_EINVAL, "plain-form entry point");
  plain_viscous(c->dc, c->dUP, c->dU, c->st);
  return launch_check("viscous");
}

int wx_update_prims(void *ctx) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (is_fused(c)) return fail(WX_EINVAL, "plain-form entry point");
  long bad = 0;
  rc = update_prims(c, &bad);
  return rc != WX_OK ? rc : (int)bad;
}

int wx_cooling(void *ctx, double dt) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (is_fused(c)) return fail(WX_EINVAL, "plain-form entry point");
  if (!c->cfg.cooling) return WX_OK;
  plain_cooling(c->dc, c->dU, c->dP, dt, c->cfg.mu0, c->st);
  return launch_check("cooling");
}

int wx_step(void *ctx, int *it, double *dt, double *time, int *nextout, double dtout, double t_sc) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (!it || !dt || !time || !nextout) return fail(WX_EINVAL, "null argument");
  if (is_fused(c)) return wx_run(ctx, 1, it, dt, time, nextout, dtout, t_sc);
  *it += 1;
  double dtl;
  rc = plain_timestep_raw(c, &dtl);
  if (rc != WX_OK) return rc;
  // start-up ramp and output clamp (source/hydro_solver.f90:191-207)
  if (*it <= 10) dtl = std::ldexp(1.0, -(11 - *it)) * dtl;
  double tout = (double)(*nextout) * dtout / t_sc;
  bool dump = false;
  if (*time + dtl >= tout) {
    dtl = tout - *time;
    dump = true;
  }
  *dt = dtl;
  rc = plain_godunov(c, c->cfg.solver == SOLVER_HLL1 ? 1 : 2, dtl);
  if (rc != WX_OK) return rc;
  *time = *time + dtl;
  plain_viscous(c->dc, c->dUP, c->dU, c->st);
  long bad = 0;
  rc = update_prims(c, &bad);
  if (rc != WX_OK) return rc;
  if (c->cfg.cooling) plain_cooling(c->dc, c->dU, c->dP, dtl, c->cfg.mu0, c->st);
  if (dump) *nextout += 1;
  rc = launch_check("step");
  return rc != WX_OK ? rc : (int)bad;
}

int wx_run(void *ctx, int nsteps, int *it, double *dt, double *time, int *nextout, double dtout, double t_sc) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (!it || !dt || !time || !nextout || nsteps < 0) return fail(WX_EINVAL, "bad argument");
  if (!is_fused(c)) {
    for (int s = 0; s < nsteps; s++) {
      rc = wx_step(ctx, it, dt, time, nextout, dtout, t_sc);
      if (rc < 0) return rc;
    }
    return WX_OK;
  }
  int frc = fused_run(c->fused, c->dc, c->st, nsteps, it, dt, time, nextout, dtout, t_sc);
  if (frc) return fail(WX_EHIP, std::string("fused run failed: ") + fused_error());
  return WX_OK;
}

int wx_boundary(void *ctx, int depth, int which) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (is_fused(c)) return fail(WX_EINVAL, "plain-form entry point");
  if (depth != 1 && depth != 2) return fail(WX_EINVAL, "depth must be 1 or 2");
  plain_boundary(c->dc, plane_wind(c), which == WX_UP ? c->dUP : c->dU, depth, c->st);
  return launch_check("boundary");
}

int wx_calc_prims(void *ctx, int which, int cells) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (is_fused(c)) return fail(WX_EINVAL, "plain-form entry point");
  if (cells < 0 || cells > 2) return fail(WX_EINVAL, "bad cell range");
  plain_prims(c->dc, which == WX_UP ? c->dUP : c->dU, c->dP, cells, nullptr, c->st);
  return launch_check("calc_prims");
}

int wx_fluxes_update(void *ctx, int order, double dtp) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (is_fused(c)) return fail(WX_EINVAL, "plain-form entry point");
  plain_fluxes(c->dc, c->dP, c->dF, order, c->st);
  plain_upwind(c->dc, c->us, c->dU, c->dP, c->dF, c->dUP, dtp, c->st);
  return launch_check("fluxes_update");
}

int wx_copy_raw(void *ctx, int which, double *out) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (is_fused(c)) return fail(WX_EINVAL, "plain-form entry point");
  const double *src = which == 0 ? c->dU : which == 1 ? c->dUP : c->dP;
  size_t n = (size_t)c->nblk * c->cfg.nv * NCELL;
  HIPCHK(hipMemcpyAsync(out, src, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return WX_OK;
}

int wx_set_tuning(void *ctx, int kind, int value) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(WX_EINVAL, "null context");
  return fused_set_tuning(c->fused, kind, value) ? fail(WX_EINVAL, "unknown tuning option") : WX_OK;
}

int wx_time_stages(void *ctx, int reps) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (!is_fused(c)) return fail(WX_EINVAL, "fused-form entry point");
  if (reps < 1) return fail(WX_EINVAL, "reps must be positive");
  if (fused_time_stages(c->fused, c->dc, c->st, reps)) return fail(WX_EHIP, fused_error());
  return WX_OK;
}

int wx_last_timing(void *ctx, double *s1, double *s2, double *tot) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(WX_EINVAL, "null context");
  fused_last_timing(c->fused, s1, s2, tot);
  return WX_OK;
}

// wx_remesh on a context with state; `pre`, when given, is the new list's
// face table that check_mesh already built
static int remesh_impl(Ctx *c, int nblk, const int *bids, Ctx *pre) {
  std::vector<RemapSrc> table;
  int rc = remap_table(c, nblk, bids, table);
  Ctx checked;
  if (rc == WX_OK && !pre) {
    rc = check_mesh(c, std::vector<int>(bids, bids + nblk), checked);
    pre = &checked;
  }
  if (rc != WX_OK) return rc;              // the current mesh and state are untouched
  const int nv = c->cfg.nv;
  HIPCHK(hipStreamSynchronize(c->st));
  size_t nold = (size_t)c->nblk * nv * NINT, nnew = (size_t)nblk * nv * NINT;
  if (is_fused(c)) {
    // the fused state has the block layout k_remap reads and writes: the new
    // state is remapped straight from it and becomes the new fused state
    DevBuf<double> d_new;
    DevBuf<RemapSrc> d_map;
    HIPCHK(d_new.alloc(nnew));
    HIPCHK(d_map.alloc(table.size()));
    HIPCHK(hipMemcpyAsync(d_map.p, table.data(), table.size() * sizeof(RemapSrc), hipMemcpyHostToDevice, c->st));
    const RemapSrc *map = d_map.p;
    double *to = d_new.p;
    hipLaunchKernelGGL(k_remap, dim3(nblk, nv), dim3(256), 0, c->st, nv, map, c->fused.U, to);
    rc = launch_check("remesh");
    HIPCHK(hipStreamSynchronize(c->st));
    if (rc == WX_OK) {
      const std::vector<double> dxyz = c->dxyz;
      int nroot[3] = {c->nroot[0], c->nroot[1], c->nroot[2]};
      rc = set_mesh_impl(c, nroot, c->maxlev, dxyz.data(), nblk, bids, pre);
    }
    if (rc != WX_OK) return rc;
    return upload_device(c, d_new.release(), true);   // the fused state owns it now
  }
  // the plain form also carries PRIM (its interior pressure is the one
  // cooling left, not the one U gives back)
  const bool carry = true;
  DevBuf<double> d_old, d_new, d_newp;
  DevBuf<RemapSrc> d_map;
  HIPCHK(d_old.alloc(nold));
  HIPCHK(d_new.alloc(nnew));
  if (carry) HIPCHK(d_newp.alloc(nnew));
  HIPCHK(d_map.alloc(table.size()));
  HIPCHK(hipMemcpyAsync(d_map.p, table.data(), table.size() * sizeof(RemapSrc), hipMemcpyHostToDevice, c->st));
  const RemapSrc *map = d_map.p;
  double *old = d_old.p, *unew = d_new.p, *pnew = d_newp.p;
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((nold + 255) / 256)), dim3(256), 0, c->st, nv, c->nblk, c->dU, old);
  hipLaunchKernelGGL(k_remap, dim3(nblk, nv), dim3(256), 0, c->st, nv, map, old, unew);
  if (carry) {
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((nold + 255) / 256)), dim3(256), 0, c->st, nv, c->nblk, c->dP, old);
    hipLaunchKernelGGL(k_remap, dim3(nblk, nv), dim3(256), 0, c->st, nv, map, old, pnew);
  }
  rc = launch_check("remesh");
  HIPCHK(hipStreamSynchronize(c->st));
  if (rc == WX_OK) {
    const std::vector<double> dxyz = c->dxyz;
    int nroot[3] = {c->nroot[0], c->nroot[1], c->nroot[2]};
    rc = set_mesh_impl(c, nroot, c->maxlev, dxyz.data(), nblk, bids, pre);
  }
  if (rc >= WX_OK) rc = upload_device(c, d_new.p);
  if (rc >= WX_OK && carry) {
    hipLaunchKernelGGL(k_carry_prims, dim3(nblk, nv), dim3(256), 0, c->st, nv, map, pnew, c->dP);
    int rc2 = launch_check("remesh prims");
    if (rc2 == WX_OK && hipStreamSynchronize(c->st) != hipSuccess) rc2 = fail(WX_EHIP, "remesh prims failed");
    if (rc2 != WX_OK) rc = rc2;
  }
  return rc;
}

int wx_remesh(void *ctx, int nblk, const int *bids) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (!bids || nblk < 1) return fail(WX_EINVAL, "bad block list");
  return remesh_impl(c, nblk, bids, nullptr);
}

int wx_admesh_plan(const int nroot[3], int maxlev, int nblk, const int *bids, const int *flags, int *flags_out,
                   int cap, int *nblk_new, int *bids_new) {
  if (!nroot || !bids || !flags || !nblk_new || nblk < 1 || maxlev < 1 || maxlev > MAXLEV)
    return fail(WX_EINVAL, "bad argument");
  Numbering N;
  N.init(nroot, maxlev);
  std::vector<int> b(bids, bids + nblk), f(flags, flags + nblk), out;
  int rc = amr_plan(N, b, f, out);
  if (rc != WX_OK) return rc;
  if (flags_out) std::copy(f.begin(), f.end(), flags_out);
  *nblk_new = (int)out.size();
  if ((int)out.size() > cap || !bids_new)
    return fail(WX_ENOMEM, "new block list (" + std::to_string(out.size()) + " blocks) exceeds the capacity");
  std::copy(out.begin(), out.end(), bids_new);
  return WX_OK;
}

// Argument checks of wx_eval_physics and its allocation-free variant; the
// record widths (doubles per record) of op
static int eval_check(const wx_config *cfg, int op, int arg, int n, const double *in, double *out, int *nin,
                      int *nout) {
  if (!cfg || !in || !out || n < 0) return fail(WX_EINVAL, "null argument");
  if (cfg->nv != 5 && cfg->nv != 8) return fail(WX_EINVAL, "nv must be 5 or 8");
  if (op < WX_EVAL_FLOW2PRIM || op > WX_EVAL_SPEEDS) return fail(WX_EINVAL, "unknown op");
  if (op == WX_EVAL_PRIM2FLUXES && (arg < 1 || arg > 3)) return fail(WX_EINVAL, "direction must be 1, 2 or 3");
  if (op == WX_EVAL_LIMITER && (arg < -1 || arg > 6)) return fail(WX_EINVAL, "unknown limiter");
  if (op == WX_EVAL_RIEMANN && (arg < SOLVER_HLL1 || arg > SOLVER_HLLD || (arg >= SOLVER_HLLE && cfg->nv != 8)))
    return fail(WX_EINVAL, "unsupported solver");
  if (op == WX_EVAL_RIEMANN_SEL && cfg->nv != 8) return fail(WX_EINVAL, "HLLD needs nv = 8");
  const int nv = cfg->nv;
  *nin = op == WX_EVAL_LIMITER ? 4 * nv : (op == WX_EVAL_RIEMANN || op == WX_EVAL_RIEMANN_SEL ? 2 * nv : nv);
  *nout = op == WX_EVAL_FLOW2PRIM ? nv + 1 : (op == WX_EVAL_LIMITER ? 2 * nv : (op == WX_EVAL_SPEEDS ? 5 : nv));
  return WX_OK;
}

int wx_eval_physics(const wx_config *cfg, int op, int arg, int n, const double *in, double *out) {
  int nin, nout;
  int rc = eval_check(cfg, op, arg, n, in, out, &nin, &nout);
  if (rc != WX_OK) return rc;
  if (n == 0) return WX_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(WX_EHIP, "no HIP device");
  Phys<double> P{cfg->nv, cfg->mhd, cfg->solver, cfg->limiter, cfg->gamma, cfg->cv};
  DevBuf<double> din, dout;
  HIPCHK(din.alloc((size_t)n * nin));
  HIPCHK(dout.alloc((size_t)n * nout));
  HIPCHK(hipMemcpy(din.p, in, (size_t)n * nin * sizeof(double), hipMemcpyHostToDevice));
  plain_eval(P, op, arg, n, din.p, dout.p, nullptr);
  rc = launch_check("eval");
  if (rc != WX_OK) return rc;
  HIPCHK(hipMemcpy(out, dout.p, (size_t)n * nout * sizeof(double), hipMemcpyDeviceToHost));
  return WX_OK;
}

// ---- allocation-free variant ------------------------------------------------
// The staging of wx_eval_physics_noalloc: inputs then outputs of one call,
// in a fixed array that each device holds from the library's load to the
// process's end.  Host code never passes the array's name to a kernel or a
// copy (on the host it is only a stand-in): it uses the device address
// resolved per device by eval_scratch().
static __device__ __align__(256) double scratch[WX_EVAL_NOALLOC_DOUBLES];

constexpr int WX_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[WX_MAX_DEVICES];

// the device address of `scratch` on the current device (resolved once per
// device, with that device current)
static int eval_scratch(double **ptr) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return fail(WX_EHIP, "no current HIP device");
  if (dev >= WX_MAX_DEVICES) return fail(WX_EINVAL, "device index beyond the supported count");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(WX_EHIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *ptr = g_scratch_ptr[dev];
  return WX_OK;
}

int wx_eval_physics_noalloc(const wx_config *cfg, int op, int arg, int n, const double *in, double *out) {
  int nin, nout;
  int rc = eval_check(cfg, op, arg, n, in, out, &nin, &nout);
  if (rc != WX_OK) return rc;
  if ((long long)n * (nin + nout) > (long long)WX_EVAL_NOALLOC_DOUBLES)
    return fail(WX_EINVAL, "n = " + std::to_string(n) + " records of " + std::to_string(nin + nout) +
                               " doubles exceed the fixed staging of " + std::to_string(WX_EVAL_NOALLOC_DOUBLES) +
                               " doubles");
  if (n == 0) return WX_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(WX_EHIP, "no HIP device");
  double *ptr = nullptr;
  rc = eval_scratch(&ptr);
  if (rc != WX_OK) return rc;
  // every element the kernel reads is written by the copy first, and it
  // writes every output element: nothing in the array needs clearing
  double *din = ptr, *dout = ptr + (size_t)n * nin;
  Phys<double> P{cfg->nv, cfg->mhd, cfg->solver, cfg->limiter, cfg->gamma, cfg->cv};
  HIPCHK(hipMemcpy(din, in, (size_t)n * nin * sizeof(double), hipMemcpyHostToDevice));
  plain_eval(P, op, arg, n, din, dout, nullptr);
  rc = launch_check("eval");
  if (rc != WX_OK) return rc;
  HIPCHK(hipMemcpy(out, dout, (size_t)n * nout * sizeof(double), hipMemcpyDeviceToHost));
  return WX_OK;
}

int wx_hilbert_keys(const int nroot[3], int maxlev, int nblk, const int *bids, int *keys) {
  if (!nroot || !bids || !keys || nblk < 0 || maxlev < 1 || maxlev > MAXLEV) return fail(WX_EINVAL, "bad argument");
  Numbering N;
  N.init(nroot, maxlev);
  const int order = hilbert_order(nroot, maxlev);
  if (3 * order > 31) return fail(WX_EINVAL, "mesh too large for 31-bit Hilbert keys");
  for (int n = 0; n < nblk; n++) {
    const int lev = N.level(bids[n]);
    if (lev < 0) return fail(WX_ETOPO, "block ID outside the mesh: " + std::to_string(bids[n]));
    int c[3];
    N.coords(bids[n], lev, c);
    keys[n] = (int)block_hilbert_key(c, lev, maxlev, order);
  }
  return WX_OK;
}

// Passive scalars the context carries (nv 6, 7: hydro with one or two).
static int carried_passives(const Ctx *c) { return (c->cfg.nv == 6 || c->cfg.nv == 7) ? c->cfg.nv - 5 : 0; }

int wx_write_bin(void *ctx, const char *path, int npassive) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  // the file holds neqtot = flow variables + npassive columns: the passives
  // the context carries, then zero columns for the rest
  const int carried = carried_passives(c);
  if (!path || npassive < carried) return fail(WX_EINVAL, "bad argument (npassive below the passives carried)");
  npassive -= carried;
  std::vector<double> u((size_t)c->nblk * c->cfg.nv * NINT);
  rc = wx_download_state(ctx, u.data(), nullptr, WX_LAYOUT_BLOCKS, 0, nullptr);
  if (rc < 0) return rc;
  std::string err;
  if (bin_write(path, c->nblk, c->cfg.nv, npassive, c->bids.data(), u.data(), err)) return fail(WX_EIO, err);
  return WX_OK;
}

int wx_read_bin(const char *path, int nv, int npassive, int cap, int *nblk, int *bids, double *u) {
  if (!path || !nblk || (nv != 5 && nv != 8) || npassive < 0) return fail(WX_EINVAL, "bad argument");
  std::string err;
  const int r = bin_read(path, nv, npassive, cap, nblk, bids, u, err);
  if (r == 2) return fail(WX_ENOMEM, err);
  if (r) return fail(WX_EIO, err);
  return WX_OK;
}

int wx_write_state(const char *path, double time, double t_sc, int it, int noutput, const char *datadir) {
  if (!path) return fail(WX_EINVAL, "null path");
  std::string err;
  if (state_write(path, time * t_sc, it, noutput, datadir, err)) return fail(WX_EIO, err);
  return WX_OK;
}

int wx_read_state(const char *path, double t_sc, double *time, int *it, int *nextout, char *datadir, int len) {
  if (!path || !time || !it || !nextout) return fail(WX_EINVAL, "null argument");
  std::string err;
  double tp = 0.0;
  int nout = 0;
  if (state_read(path, &tp, it, &nout, datadir, len, err)) return fail(WX_EIO, err);
  *time = tp / t_sc;
  *nextout = nout + 1;
  return WX_OK;
}

int wx_write_vtk(void *ctx, const char *datadir, int noutput, int npassive, int units) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (!datadir || npassive < carried_passives(c) || noutput < 0 || noutput > 9999 || (units != 1 && units != 2))
    return fail(WX_EINVAL, "bad argument");
  const int nv = c->cfg.nv;
  std::vector<double> u((size_t)c->nblk * nv * NINT), prim(u.size());
  rc = wx_download_state(ctx, u.data(), nullptr, WX_LAYOUT_BLOCKS, 0, nullptr);
  if (rc < 0) return rc;
  // writeOutput calls updatePrims before the VTK files: PRIM from U
  for (int b = 0; b < c->nblk; b++)
    for (int q = 0; q < NINT; q++) {
      double uu[MAXV], pp[MAXV];
      for (int v = 0; v < nv; v++) uu[v] = u[((size_t)b * nv + v) * NINT + q];
      flow2prim(c->dc.pd, uu, pp);
      for (int v = c->dc.pd.nv; v < nv; v++) pp[v] = uu[v];   // passive scalars
      for (int v = 0; v < nv; v++) prim[((size_t)b * nv + v) * NINT + q] = pp[v];
    }
  std::vector<int> levels(c->nblk), coords(3 * (size_t)c->nblk);
  for (int b = 0; b < c->nblk; b++) {
    levels[b] = c->info[b].level;
    for (int a = 0; a < 3; a++) coords[3 * b + a] = c->info[b].coords[a];
  }
  VtkInfo m;
  m.nblk = c->nblk;
  m.nv = nv;
  m.npassive = npassive;
  m.maxlev = c->maxlev;
  m.eos = c->cfg.eos;
  m.phys_units = units == 1;
  m.bids = c->bids.data();
  m.levels = levels.data();
  m.coords = coords.data();
  m.dxyz = c->dxyz.data();
  for (int d = 0; d < 3; d++) m.nroot[d] = c->nroot[d];
  m.l_sc = c->cfg.l_sc;
  m.d_sc = c->cfg.d_sc;
  m.v_sc = c->cfg.v_sc;
  m.p_sc = c->cfg.p_sc;
  m.b_sc = std::sqrt(4.0 * PI_REF * c->cfg.d_sc) * c->cfg.v_sc;   // parameters.f90: b_sc
  m.temp_mu0 = c->dc.temp_mu0;
  m.temp_mui = c->dc.temp_mui;
  m.ion_thres = c->cfg.ion_thres;
  m.mu0 = c->cfg.mu0;
  m.kb = KB;
  m.ncarried = carried_passives(c);
  std::string dir(datadir);
  if (dir.empty() || dir.back() != '/') dir += '/';
  char name[64];
  std::snprintf(name, sizeof name, "Blocks000.%04d.vtk", noutput);
  std::string err;
  if (visit_write((dir + "master.visit").c_str(), name, noutput, noutput == 0, err) ||
      vtk_blocks_write((dir + name).c_str(), m, prim.data(), err))
    return fail(WX_EIO, err);
  std::snprintf(name, sizeof name, "Grid.%04d.vtk", noutput);
  if (vtk_grid_write((dir + name).c_str(), m, err)) return fail(WX_EIO, err);
  return WX_OK;
}

int wx_set_plane_wind(void *ctx, int plane, double rho, double vel, double temp, double mu) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(WX_EINVAL, "null context");
  if (plane < 0 || plane > 6) return fail(WX_EINVAL, "plane must be 0 (none) or PLANE_LEFT=1 .. PLANE_TOP=6");
  if (plane == 0) {
    c->pw = PlaneWind{};
    return WX_OK;
  }
  if (is_fused(c))
    return fail(WX_EINVAL, "a plane wind needs the plain FP64 form (the fused forms impose spherical sources only)");
  if (c->cfg.nv == 6 || c->cfg.nv == 7)
    return fail(WX_EINVAL, "a plane wind with passive scalars: the reference leaves the passives of the wind state unset");
  if (c->cfg.nv != 5)
    return fail(WX_EINVAL, "a plane wind is hydro only (the reference leaves the field of its state unset)");
  if (!(rho > 0.0) || !(temp > 0.0) || !(mu > 0.0) || !std::isfinite(rho) || !std::isfinite(vel) ||
      !std::isfinite(temp) || !std::isfinite(mu))
    return fail(WX_EINVAL, "plane wind: rho, temp and mu must be positive and finite, vel finite");
  c->pw = PlaneWind{plane, 0, rho, vel, temp, mu};
  return WX_OK;
}

int wx_set_uniform_gravity(void *ctx, int on, double gx, double gy, double gz) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(WX_EINVAL, "null context");
  if (on == 0) {
    c->us = UserSource{};
    return WX_OK;
  }
  if (on != 1) return fail(WX_EINVAL, "on must be 0 (none) or 1");
  if (is_fused(c))
    return fail(WX_EINVAL, "a user source term needs the plain FP64 form (the fused forms carry none)");
  if (!std::isfinite(gx) || !std::isfinite(gy) || !std::isfinite(gz))
    return fail(WX_EINVAL, "uniform gravity: gx, gy and gz must be finite");
  c->us = UserSource{1, {gx, gy, gz}};
  return WX_OK;
}

int wx_record_ghosts(void *ctx) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (!is_fused(c)) return WX_OK;   // the plain form's PRIM keeps its ghost layer
  if (!c->d_ghosts) HIPCHK(hipMalloc(&c->d_ghosts, (size_t)c->nblk * 2 * 6 * 256 * sizeof(double)));
  fused_ghost_prims(c->dc, c->fused.U, c->d_ghosts, c->st);
  rc = launch_check("ghost record");
  if (rc == WX_OK && hipStreamSynchronize(c->st) != hipSuccess) rc = fail(WX_EHIP, "ghost record failed");
  c->ghosts_recorded = rc == WX_OK;
  return rc;
}

int wx_mark_blocks(void *ctx, double refine_thres, double coarse_thres, int *flags, double *maxgrad) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (!flags) return fail(WX_EINVAL, "null flags");
  return mark_blocks(c, refine_thres, coarse_thres, flags, maxgrad);
}

int wx_admesh(void *ctx, double refine_thres, double coarse_thres, int cap, int *nblk_new, int *bids_new) {
  Ctx *c = as_ctx(ctx);
  int rc = need_state(c);
  if (rc != WX_OK) return rc;
  if (!nblk_new) return fail(WX_EINVAL, "null argument");
  std::vector<int> flags(c->nblk), out;
  rc = mark_blocks(c, refine_thres, coarse_thres, flags.data(), nullptr);
  if (rc != WX_OK) return rc;
  Numbering N;
  N.init(c->nroot, c->maxlev);
  rc = amr_plan(N, c->bids, flags, out);
  if (rc != WX_OK) return rc;
  *nblk_new = (int)out.size();
  if ((int)out.size() > cap || !bids_new)
    return fail(WX_ENOMEM, "new block list (" + std::to_string(out.size()) + " blocks) exceeds the capacity");
  if (out == c->bids) {   // no change: the mesh, state and step graph stay
    std::copy(out.begin(), out.end(), bids_new);
    return WX_OK;
  }
  Ctx checked;
  rc = check_mesh(c, out, checked);
  if (rc != WX_OK) return rc;
  std::copy(out.begin(), out.end(), bids_new);
  return remesh_impl(c, (int)out.size(), out.data(), &checked);
}

}  // extern "C"
