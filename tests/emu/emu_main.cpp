// Test infrastructure: drives the library's C ABI (include/walicxe_hip.h) on
// the CPU build of tests/emu for the cases named on the command line: each
// pair is an input file written by tests/test_emulated_kernels.py and a step
// count; the final state goes to <input>.out.  A remesh count of -1 in the
// input asks for wx_admesh after every step instead (final block list to
// <input>.blocks).  With EMU_VTK_DIR set, writeOutput's VTK files of the
// initial and the final state (output numbers 0 and nsteps, one passive
// column, units_type EMU_VTK_UNITS, default CODE_UNITS) go to that folder.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <cstring>
#include <string>
#include "walicxe_hip.h"
static int run_case(const char *path, int nsteps);
int main(int argc, char **argv) {
  for (int a = 1; a + 1 < argc; a += 2) {
    printf("case %s steps %s\n", argv[a], argv[a + 1]); fflush(stdout);
    if (run_case(argv[a], atoi(argv[a + 1]))) return 1;
  }
  return 0;
}
static int run_case(const char *path, int nsteps) {
  FILE *f = fopen(path, "rb");
  int hdr[8];
  fread(hdr, 4, 8, f);
  int nv = hdr[0], nblk = hdr[1], ntab = hdr[2], nwind = hdr[3], maxlev = hdr[4];
  int nroot[3] = {hdr[5], hdr[6], hdr[7]};
  int clen; fread(&clen, 4, 1, f);
  wx_config cfg; if (clen != (int)sizeof(cfg)) { printf("cfg size %d vs %zu\n", clen, sizeof(cfg)); return 1; }
  fread(&cfg, clen, 1, f);
  std::vector<double> tT(ntab), tC(ntab), wnd(8 * nwind), dx(3 * maxlev), U((size_t)nblk * nv * 4096);
  std::vector<int> bids(nblk);
  fread(tT.data(), 8, ntab, f); fread(tC.data(), 8, ntab, f); fread(wnd.data(), 8, 8 * nwind, f);
  fread(dx.data(), 8, 3 * maxlev, f); fread(bids.data(), 4, nblk, f); fread(U.data(), 8, U.size(), f);
  double dtout_tsc[2]; fread(dtout_tsc, 8, 2, f);
  // optional: a second block list to move the state to (wx_remesh) after the run
  int nb2 = 0;
  std::vector<int> bids2;
  double thres[2] = {0.3, 0.03};
  if (fread(&nb2, 4, 1, f) == 1 && nb2 == -1) {
    fread(thres, 8, 2, f);
  } else if (nb2 > 0) {
    bids2.resize(nb2);
    fread(bids2.data(), 4, nb2, f);
  }
  fclose(f);
  cfg.tab_logT = ntab ? tT.data() : nullptr; cfg.tab_loglam = ntab ? tC.data() : nullptr;
  cfg.winds = nwind ? wnd.data() : nullptr;
  void *ctx;
  int rc = wx_init(&cfg, &ctx); if (rc < 0) { printf("init %d %s\n", rc, wx_last_error()); return 1; }
  rc = wx_set_mesh(ctx, nroot, maxlev, dx.data(), nblk, bids.data()); if (rc < 0) { printf("mesh %d %s\n", rc, wx_last_error()); return 1; }
  rc = wx_upload_state(ctx, U.data(), WX_LAYOUT_BLOCKS, 0, nullptr); if (rc < 0) { printf("upload %d %s\n", rc, wx_last_error()); return 1; }
  if (const char *g = getenv("EMU_GRAVITY")) {   // "gx gy gz": a uniform acceleration
    double gx = 0, gy = 0, gz = 0;
    if (sscanf(g, "%lf %lf %lf", &gx, &gy, &gz) != 3) { printf("EMU_GRAVITY needs three numbers\n"); return 1; }
    rc = wx_set_uniform_gravity(ctx, 1, gx, gy, gz); if (rc < 0) { printf("gravity %d %s\n", rc, wx_last_error()); return 1; }
  }
  if (getenv("EMU_NSLAB") && cfg.form != WX_FORM_PLAIN_FP64) {
    rc = wx_set_tuning(ctx, 4, atoi(getenv("EMU_NSLAB"))); if (rc < 0) { printf("slabs %d %s\n", rc, wx_last_error()); return 1; }
  }
  const char *vtkdir = getenv("EMU_VTK_DIR");
  const int npv = (cfg.nv == 6 || cfg.nv == 7) ? cfg.nv - 5 : 1;   // passive fields: the carried ones, or one zero
  const int vtkunits = getenv("EMU_VTK_UNITS") ? atoi(getenv("EMU_VTK_UNITS")) : 2;
  if (vtkdir && wx_write_vtk(ctx, vtkdir, 0, npv, vtkunits) < 0) { printf("vtk %s\n", wx_last_error()); return 1; }
  int it = 0, nextout = 1; double dt = 0, time = 0;
  if (nb2 == -1) {
    std::vector<int> nb(8 * 4096);
    for (int n = 0; n < nsteps && rc >= 0; n++) {
      // fused form: the recorded ghost layer on even steps, the current
      // state's exchange on odd ones
      if (cfg.form != WX_FORM_PLAIN_FP64 && n % 2 == 0) {
        rc = wx_record_ghosts(ctx);
        if (rc < 0) { printf("record %d %s\n", rc, wx_last_error()); return 1; }
      }
      if (cfg.form == WX_FORM_PLAIN_FP64) rc = wx_step(ctx, &it, &dt, &time, &nextout, dtout_tsc[0], dtout_tsc[1]);
      else rc = wx_run(ctx, 1, &it, &dt, &time, &nextout, dtout_tsc[0], dtout_tsc[1]);
      if (rc < 0) break;
      rc = wx_admesh(ctx, thres[0], thres[1], (int)nb.size(), &nblk, nb.data());
      if (rc < 0) { printf("admesh %d %s\n", rc, wx_last_error()); return 1; }
      printf("step %d blocks %d\n", n + 1, nblk); fflush(stdout);
    }
    bids.assign(nb.begin(), nb.begin() + nblk);
    U.resize((size_t)nblk * nv * 4096);
    std::string fn = std::string(path) + ".blocks";
    FILE *fb = fopen(fn.c_str(), "wb"); fwrite(bids.data(), 4, nblk, fb); fclose(fb);
  } else if (cfg.form == WX_FORM_PLAIN_FP64) {
    for (int n = 0; n < nsteps && rc >= 0; n++)
      rc = wx_step(ctx, &it, &dt, &time, &nextout, dtout_tsc[0], dtout_tsc[1]);
  } else {
    rc = wx_run(ctx, nsteps, &it, &dt, &time, &nextout, dtout_tsc[0], dtout_tsc[1]);
  }
  if (rc < 0) { printf("run %d %s\n", rc, wx_last_error()); return 1; }
  if (nb2 > 0) {
    rc = wx_remesh(ctx, nb2, bids2.data());
    if (rc < 0) { printf("remesh %d %s\n", rc, wx_last_error()); return 1; }
    rc = wx_remesh(ctx, nblk, bids.data());
    if (rc < 0) { printf("remesh back %d %s\n", rc, wx_last_error()); return 1; }
    if (cfg.form == WX_FORM_PLAIN_FP64) rc = wx_step(ctx, &it, &dt, &time, &nextout, dtout_tsc[0], dtout_tsc[1]);
    else rc = wx_run(ctx, 1, &it, &dt, &time, &nextout, dtout_tsc[0], dtout_tsc[1]);
    if (rc < 0) { printf("run after remesh %d %s\n", rc, wx_last_error()); return 1; }
  }
  std::vector<double> Uo(U.size());
  wx_download_state(ctx, Uo.data(), nullptr, WX_LAYOUT_BLOCKS, 0, nullptr);
  { std::string fn = std::string(path) + ".out"; FILE *fo = fopen(fn.c_str(), "wb"); fwrite(Uo.data(), 8, Uo.size(), fo); fclose(fo); }
  if (vtkdir && wx_write_vtk(ctx, vtkdir, nsteps, npv, vtkunits) < 0) { printf("vtk %s\n", wx_last_error()); return 1; }
  // the reference's checkpoint file of the final state (one zero passive
  // column, as the oracle builds of the reference carry one passive scalar)
  // (hydro with passives: the file's passive columns hold the carried ones)
  { const int np = (cfg.nv == 6 || cfg.nv == 7) ? cfg.nv - 5 : 1;
    std::string fn = std::string(path) + ".wxbin"; if (wx_write_bin(ctx, fn.c_str(), np) < 0) { printf("write_bin %s\n", wx_last_error()); return 1; } }
  double s = 0; for (double x : Uo) s += x;
  printf("it %d dt %.17g time %.17g sum %.17g\n", it, dt, time, s);
  wx_finalize(ctx);
  return 0;
}
