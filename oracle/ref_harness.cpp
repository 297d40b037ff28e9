/*
 * ref_harness.cpp -- test infrastructure, not product code.
 *
 * Links against the reference's OWN functions (dym-mod-metro.cpp compiled with
 * main renamed, lattice.c, group.c) and dumps what they compute, so that the
 * oracle restatement and the HIP path can be checked against the reference
 * itself.  Built by oracle/ref.mk into oracle/_ref/ref_harness.
 *
 *   ref_harness loops  <group> D Nt Nx <cfg.i32> <outdir>
 *       For the link configuration in cfg.i32 (int32, reference order
 *       a[i*D+mu]) write, per link, the 2(D-1) plaquette staples and 6(D-1)
 *       rectangle staples of get_staples / get_rect_loops
 *       (dym-mod-metro.cpp:63-142), r1 and m1 as the update's hit loop forms
 *       them (:166-169), and the loop elements of plaquette / wilson /
 *       polyakov (lattice.c:76-124), plus the measurement-block means
 *       (dym-mod-metro.cpp:282-318) at one thread.
 *
 *   ref_harness traj   <group> D Nt Nx beta0 beta1 beta2 seed nsweeps hot|cold <outdir> [every]
 *       Repeat main's initialisation (dym-mod-metro.cpp:222-268) and call the
 *       reference update() nsweeps times; write the link array after every
 *       `every` sweeps (default: only at the end) and the hit/acc counters.
 */
extern "C" {
#include "group.h"
#include "lattice.h"
}
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include <random>

extern std::default_random_engine *rnd;
extern std::vector<group_t> smallgroup;
extern char *site_parity;
extern unsigned long hit, acc;
void get_staples(unsigned int i, unsigned int d, group_t *st);
void get_rect_loops(unsigned int i, unsigned int d, group_t *st);
void update();
extern "C" void getpos(unsigned idx, unsigned int *pos);

static void write_file(const std::string &path, const void *p, size_t n) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) { perror(path.c_str()); exit(2); }
  if (n && fwrite(p, 1, n, f) != n) { perror(path.c_str()); exit(2); }
  fclose(f);
}

static void set_volume() {
  V = Nt;
  for (unsigned d = 1; d < D; ++d) V *= Nx;
}

static void build_pool() {
  // same rule as dym-mod-metro.cpp:239-243
  double min_retr = 0, nn_retr = 10;
  for (unsigned i = 0; i < P; ++i) if (ReTr[i] < min_retr) min_retr = ReTr[i];
  for (unsigned i = 0; i < P; ++i) if (ReTr[i] > min_retr && ReTr[i] < nn_retr) nn_retr = ReTr[i];
  smallgroup.clear();
  for (unsigned i = 0; i < P; ++i)
    if (ReTr[i] < nn_retr + 1e-6 && ReTr[i] > min_retr) smallgroup.push_back(i);
}

static int mode_loops(int argc, char **argv) {
  if (argc < 8) { fprintf(stderr, "loops: bad args\n"); return 1; }
  D = atoi(argv[3]); Nt = atoi(argv[4]); Nx = atoi(argv[5]);
  set_volume();
  load_group(argv[2]);
  std::string out = argv[7];
  a = (group_t *)malloc(sizeof(group_t) * V * D);
  FILE *f = fopen(argv[6], "rb");
  if (!f || fread(a, sizeof(group_t), (size_t)V * D, f) != (size_t)V * D) { fprintf(stderr, "cfg read\n"); return 1; }
  fclose(f);
  step(0, 0, 1);
  const unsigned ns = 2 * (D - 1), nr = 6 * (D - 1);
  std::vector<int> st((size_t)V * D * ns), re((size_t)V * D * nr);
  std::vector<double> r1((size_t)V * D), m1((size_t)V * D);
  for (unsigned i = 0; i < V; ++i)
    for (unsigned d = 0; d < D; ++d) {
      size_t l = (size_t)i * D + d;
      get_staples(i, d, &st[l * ns]);
      get_rect_loops(i, d, &re[l * nr]);
      group_t b = a[l];
      double s1 = 0, s2 = 0;
      for (unsigned j = 0; j < ns; ++j) s1 += ReTr[mult[b][st[l * ns + j]]];
      for (unsigned k = 0; k < nr; ++k) s2 += ReTr[mult[b][re[l * nr + k]]];
      r1[l] = s1; m1[l] = s2;
    }
  write_file(out + "/staples.i32", st.data(), st.size() * 4);
  write_file(out + "/rect.i32", re.data(), re.size() * 4);
  write_file(out + "/r1.f64", r1.data(), r1.size() * 8);
  write_file(out + "/m1.f64", m1.data(), m1.size() * 8);

  // loop elements
  std::vector<int> plaq, rect21, poly, wil;
  for (unsigned i = 0; i < V; ++i)
    for (unsigned d1 = 0; d1 < D; ++d1)
      for (unsigned d2 = d1 + 1; d2 < D; ++d2) plaq.push_back(plaquette(i, d1, d2));
  for (unsigned i = 0; i < V; ++i)
    for (unsigned d1 = 0; d1 < D; ++d1)
      for (unsigned d2 = 0; d2 < D; ++d2)
        if (d1 != d2) rect21.push_back(wilson(i, d1, d2, 2, 1));
  for (unsigned i = 0; i < V; ++i) {
    unsigned pos[16]; getpos(i, pos);
    if (pos[0] == 0) poly.push_back(polyakov(i));
  }
  // Wilson grid, k outer, l inner, then site, then spatial direction
  for (unsigned k = 0; k < Nx; ++k)
    for (unsigned l = 0; l < Nt; ++l)
      for (unsigned i = 0; i < V; ++i)
        for (unsigned d1 = 1; d1 < D; ++d1) wil.push_back(wilson(i, d1, 0, k + 1, l + 1));
  write_file(out + "/plaq.i32", plaq.data(), plaq.size() * 4);
  write_file(out + "/rect21.i32", rect21.data(), rect21.size() * 4);
  write_file(out + "/poly.i32", poly.data(), poly.size() * 4);
  write_file(out + "/wil.i32", wil.data(), wil.size() * 4);

  // measurement-block means in the order main computes them (one thread)
  std::vector<double> gm;
  double rep, imp; getpoly(&rep, &imp);
  double simpleplaq = 0;
  for (unsigned i = 0; i < V; ++i)
    for (unsigned d1 = 0; d1 < D; ++d1)
      for (unsigned d2 = d1 + 1; d2 < D; ++d2) simpleplaq += ReTr[plaquette(i, d1, d2)];
  simpleplaq /= V * D * (D - 1) / 2;
  double simplerect = 0;
  for (unsigned i = 0; i < V; ++i)
    for (unsigned d1 = 0; d1 < D; ++d1)
      for (unsigned d2 = 0; d2 < D; ++d2)
        if (d1 != d2) simplerect += ReTr[wilson(i, d1, d2, 2, 1)];
  simplerect /= V * D * (D - 1);
  gm.push_back(rep); gm.push_back(imp); gm.push_back(simpleplaq); gm.push_back(simplerect);
  for (unsigned k = 0; k < Nx; ++k)
    for (unsigned l = 0; l < Nt; ++l) {
      double lw = 0;
      for (unsigned i = 0; i < V; ++i)
        for (unsigned d1 = 1; d1 < D; ++d1) lw += ReTr[wilson(i, d1, 0, k + 1, l + 1)];
      gm.push_back(lw / (V * (D - 1)));
    }
  write_file(out + "/means.f64", gm.data(), gm.size() * 8);
  return 0;
}

static int mode_traj(int argc, char **argv) {
  if (argc < 13) { fprintf(stderr, "traj: bad args\n"); return 1; }
  D = atoi(argv[3]); Nt = atoi(argv[4]); Nx = atoi(argv[5]);
  beta0 = atof(argv[6]); beta1 = atof(argv[7]); beta2 = atof(argv[8]);
  int iseed = atoi(argv[9]);
  unsigned nsweeps = (unsigned)atoi(argv[10]);
  bool cold = (argv[11][0] == 'c' || argv[11][0] == 'C');
  std::string out = argv[12];
  unsigned every = (argc > 13) ? (unsigned)atoi(argv[13]) : 0;
  set_volume();
  // dym-mod-metro.cpp:227-228
  rnd = new std::default_random_engine[V];
  for (int i = 0; i < (int)V; ++i) rnd[i].seed(iseed + i);
  load_group(argv[2]);
  std::uniform_int_distribution<> randgrp(0, P - 1);
  build_pool();
  a = (group_t *)malloc(sizeof(*a) * V * D);
  for (unsigned i = 0; i < V * D; i++) a[i] = cold ? id : randgrp(rnd[0]);
  step(0, 0, 1);
  site_parity = (char *)malloc(V);
  for (unsigned i = 0; i < V; ++i) {
    unsigned pos[16]; getpos(i, pos);
    int s = 0;
    for (unsigned d = 0; d < D; ++d) s += pos[d];
    site_parity[i] = s & 1;
  }
  write_file(out + "/links_0.i32", a, sizeof(int) * V * D);
  std::vector<unsigned long> counters;
  for (unsigned s = 1; s <= nsweeps; ++s) {
    update();
    if ((every && s % every == 0) || s == nsweeps) {
      char name[64]; snprintf(name, sizeof name, "/links_%u.i32", s);
      write_file(out + name, a, sizeof(int) * V * D);
      counters.push_back(s); counters.push_back(hit); counters.push_back(acc);
    }
  }
  write_file(out + "/counters.u64", counters.data(), counters.size() * 8);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 2) { fprintf(stderr, "usage: ref_harness loops|traj ...\n"); return 1; }
  if (!strcmp(argv[1], "loops")) return mode_loops(argc, argv);
  if (!strcmp(argv[1], "traj")) return mode_traj(argc, argv);
  fprintf(stderr, "unknown mode %s\n", argv[1]);
  return 1;
}
