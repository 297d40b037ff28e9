/*
 * ref_llr_harness.cpp -- test infrastructure, not product code.
 *
 * Links against the reference's OWN LLR functions (llr.cpp compiled with main
 * renamed, lattice.c, group.c) and dumps the state of its ladder window by
 * window, so that the oracle's serial restatement (oracle_llr.c, ogl_ref_*)
 * and the device replay of llr can be checked against the reference bit for
 * bit.  Built by oracle/ref.mk into oracle/_ref/ref_llr_harness.
 *
 *   ref_llr_harness ladder <group> D Nt Nx axis bg Etop Ebot step hw a0 seed K NRM <outdir>
 *
 * Repeats main's set-up (llr.cpp:296-321: per-site streams seeded iseed + i,
 * hot start from site 0's stream, calibrate_rect_mult), then one repeat
 * (R = 1) of the two ladders as run_pass walks them (llr.cpp:256-274:
 * streams reseeded iseed + i, hot start walking down, cold start walking
 * up), calling the reference's seed_window and rm_solve for each window.
 * Writes
 *   init.f64      RECT_MULT
 *   windows.f64   per window visited, 9 doubles: pass (0 down, 1 up), window
 *                 index n, reached, S1 and S2 after the drive-in, a_n (0 when
 *                 not reached), S1 and S2 after rm_solve, max drift so far
 *   start_<p>.i32 the links at the start of pass p
 *   links_<p>_<k>.i32  the links after the k-th window of pass p
 */
extern "C" {
#include "group.h"
#include "lattice.h"
}
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <string>
#include <vector>

extern std::default_random_engine *rnd;
extern std::vector<group_t> smallgroup;
extern double S1, S2, BG, g_maxdrift, RECT_MULT;
extern int AXIS;
double recompute_S1();
double recompute_S2();
void calibrate_rect_mult();
bool seed_window(double Elo, double Ehi, double a_guess);
double rm_solve(double Ewin, double hw, double a0, unsigned K, unsigned NRM);

static void write_file(const std::string &path, const void *p, size_t n) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) { perror(path.c_str()); exit(2); }
  if (n && fwrite(p, 1, n, f) != n) { perror(path.c_str()); exit(2); }
  fclose(f);
}

static int mode_ladder(int argc, char **argv) {
  if (argc < 17) { fprintf(stderr, "ladder: bad args\n"); return 1; }
  D = atoi(argv[3]); Nt = atoi(argv[4]); Nx = atoi(argv[5]);
  AXIS = atoi(argv[6]);
  BG = atof(argv[7]);
  const double Etop = atof(argv[8]), Ebot = atof(argv[9]), stepE = atof(argv[10]), hw = atof(argv[11]);
  const double a0 = atof(argv[12]);
  const int iseed = atoi(argv[13]);
  const unsigned K = (unsigned)atoi(argv[14]), NRM = (unsigned)atoi(argv[15]);
  const std::string out = argv[16];
  const int M = (int)((Etop - Ebot) / stepE + 0.5) + 1;
  V = Nt;
  for (unsigned d = 1; d < D; ++d) V *= Nx;
  rnd = new std::default_random_engine[V];
  load_group(argv[2]);
  std::uniform_int_distribution<> randgrp(0, P - 1);
  // llr.cpp:306-309 (the pool rule)
  double min_retr = 0, nn_retr = 10;
  for (unsigned i = 0; i < P; ++i) if (ReTr[i] < min_retr) min_retr = ReTr[i];
  for (unsigned i = 0; i < P; ++i) if (ReTr[i] > min_retr && ReTr[i] < nn_retr) nn_retr = ReTr[i];
  smallgroup.clear();
  for (unsigned i = 0; i < P; ++i) if (ReTr[i] < nn_retr + 1e-6 && ReTr[i] > min_retr) smallgroup.push_back(i);
  a = (group_t *)malloc(sizeof(*a) * V * D);
  for (unsigned i = 0; i < V; ++i) rnd[i].seed(iseed + i);
  for (unsigned i = 0; i < V * D; ++i) a[i] = randgrp(rnd[0]);
  step(0, 0, 1);
  calibrate_rect_mult();
  write_file(out + "/init.f64", &RECT_MULT, sizeof(double));
  std::vector<double> rec;
  for (unsigned i = 0; i < V; ++i) rnd[i].seed(iseed + 0 * 1000003u + i);
  for (int pass = 0; pass < 2; ++pass) {
    const int dir = pass == 0 ? +1 : -1;
    if (dir > 0) for (unsigned i = 0; i < V * D; ++i) a[i] = randgrp(rnd[0]);
    else         for (unsigned i = 0; i < V * D; ++i) a[i] = id;
    S1 = recompute_S1();
    S2 = recompute_S2();
    write_file(out + "/start_" + std::to_string(pass) + ".i32", a, sizeof(int) * V * D);
    double a_prev = a0;
    for (int s = 0; s < M; ++s) {
      const int n = dir > 0 ? s : M - 1 - s;
      const double Ewin = Etop - n * stepE, Elo = Ewin - hw, Ehi = Ewin + hw;
      const bool ok = seed_window(Elo, Ehi, a_prev);
      const double s1 = S1, s2 = S2;
      double an = 0;
      if (ok) an = rm_solve(Ewin, hw, a_prev, K, NRM);
      rec.insert(rec.end(), {(double)pass, (double)n, ok ? 1.0 : 0.0, s1, s2, an, S1, S2, g_maxdrift});
      if (!ok) break;
      write_file(out + "/links_" + std::to_string(pass) + "_" + std::to_string(s) + ".i32", a,
                 sizeof(int) * V * D);
      a_prev = an;
    }
  }
  write_file(out + "/windows.f64", rec.data(), rec.size() * sizeof(double));
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 2) { fprintf(stderr, "usage: ref_llr_harness ladder ...\n"); return 1; }
  if (!strcmp(argv[1], "ladder")) return mode_ladder(argc, argv);
  fprintf(stderr, "unknown mode %s\n", argv[1]);
  return 1;
}
