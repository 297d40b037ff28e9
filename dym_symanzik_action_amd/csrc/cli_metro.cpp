// cli_metro.cpp -- drop-in for the reference's dym-mod-metro binary.
//
//   dym-mod-metro <group> D Nt Nx beta0 beta1 beta2 seed [K] [N] [Ntherm] [hot|cold]
//                 [--rng philox|minstd] [--order colour|reference]
//                 [--sum tree|ordered] [--wgrid KxL] [--replay]
//
// Positional arguments, defaults (K = N = 1000, Ntherm = 0, NHIT = 20, hot)
// and the stdout lines follow dym-mod-metro.cpp:199-337, lattice.c:57 and
// timer.cpp:14,24; usage goes to stderr with exit status 1 when fewer than
// eight positional arguments are given.  The options change only what they
// name: --replay (= --rng minstd --order reference --sum ordered) reproduces
// the reference run at one thread bit for bit; --wgrid measures a KxL Wilson
// grid instead of the reference's Nx x Nt.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include <string>
#include <vector>

#include "../../include/dsa.h"

namespace {

struct Timer {
  timeval t0;
  std::string msg;
  void start(const char *m) {
    msg = m;
    printf("TM_BEGIN: %-40s ...\n", msg.c_str());
    gettimeofday(&t0, nullptr);
  }
  void stop() {
    timeval t1;
    gettimeofday(&t1, nullptr);
    double t = (double)(t1.tv_sec - t0.tv_sec) + (t1.tv_usec - t0.tv_usec) / 1000000.0;
    printf("TM_END:   %-20s %-20s... done in %.3e sec.\n", msg.c_str(), "", t);
    fflush(stdout);
  }
};

int die(const char *what) {
  fprintf(stderr, "dym-mod-metro: %s: %s\n", what, dsa_last_error());
  return 2;
}

}  // namespace

int main(int argc, char **argv) {
  std::vector<char *> pos;
  int rng = DSA_RNG_PHILOX, order = DSA_ORDER_COLOUR, sum = DSA_SUM_TREE;
  int wk = -1, wl = -1;
  for (int i = 1; i < argc; ++i) {
    const char *a = argv[i];
    if (strncmp(a, "--", 2) != 0) {
      pos.push_back(argv[i]);
      continue;
    }
    const char *v = (i + 1 < argc) ? argv[i + 1] : nullptr;
    if (!strcmp(a, "--replay")) {
      rng = DSA_RNG_MINSTD;
      order = DSA_ORDER_REFERENCE;
      sum = DSA_SUM_ORDERED;
    } else if (!strcmp(a, "--rng") && v) {
      if (!strcmp(v, "philox")) rng = DSA_RNG_PHILOX;
      else if (!strcmp(v, "minstd")) rng = DSA_RNG_MINSTD;
      else { fprintf(stderr, "unknown --rng %s\n", v); return 1; }
      ++i;
    } else if (!strcmp(a, "--order") && v) {
      if (!strcmp(v, "colour")) order = DSA_ORDER_COLOUR;
      else if (!strcmp(v, "reference")) order = DSA_ORDER_REFERENCE;
      else { fprintf(stderr, "unknown --order %s\n", v); return 1; }
      ++i;
    } else if (!strcmp(a, "--sum") && v) {
      if (!strcmp(v, "tree")) sum = DSA_SUM_TREE;
      else if (!strcmp(v, "ordered")) sum = DSA_SUM_ORDERED;
      else { fprintf(stderr, "unknown --sum %s\n", v); return 1; }
      ++i;
    } else if (!strcmp(a, "--wgrid") && v) {
      if (sscanf(v, "%dx%d", &wk, &wl) != 2 || wk < 0 || wl < 0) { fprintf(stderr, "bad --wgrid %s\n", v); return 1; }
      ++i;
    } else {
      fprintf(stderr, "unknown option %s\n", a);
      return 1;
    }
  }
  if (pos.size() < 8) {
    fprintf(stderr, "usage: %s group D Nt Nx beta0 beta1 beta2 seed [K] [N] [Ntherm]\n", argv[0]);
    return 1;
  }
  const char *gf = pos[0];
  dsa_chain_params p{};
  p.D = (uint32_t)atoi(pos[1]);
  p.Nt = (uint32_t)atoi(pos[2]);
  p.Nx = (uint32_t)atoi(pos[3]);
  p.beta0 = atof(pos[4]);
  p.beta1 = atof(pos[5]);
  p.beta2 = atof(pos[6]);
  p.seed = atoi(pos[7]);
  const unsigned K = pos.size() > 8 ? (unsigned)atoi(pos[8]) : 1000;
  const unsigned N = pos.size() > 9 ? (unsigned)atoi(pos[9]) : 1000;
  const unsigned NT = pos.size() > 10 ? (unsigned)atoi(pos[10]) : 0;
  const char *sk = pos.size() > 11 ? pos[11] : "hot";
  const bool cold = sk[0] == 'c' || sk[0] == 'C';
  p.start = cold ? DSA_START_COLD : DSA_START_HOT;
  p.rng = rng;
  p.order = order;
  if (wk < 0) {
    wk = (int)p.Nx;
    wl = (int)p.Nt;
  }

  printf("PARAMS(grp,D,Nt,Nx,beta0,beta1,beta2,seed): %s %d %d %d %e %e %e %d\n", gf, (int)p.D, (int)p.Nt,
         (int)p.Nx, p.beta0, p.beta1, p.beta2, p.seed);
  printf("RUN(K,N,Ntherm,NHIT): %u %u %u %d\n", K, N, NT, 20);
  fflush(stdout);
  dsa_group *g = nullptr;
  if (dsa_group_load(gf, &g)) return die("loading the group");
  dsa_group_info gi;
  dsa_group_info_get(g, &gi);
  std::vector<int32_t> mult((size_t)gi.P * gi.P);
  dsa_group_tables(g, mult.data(), nullptr, nullptr, nullptr, nullptr);
  printf("read the group\n");
  fflush(stdout);
  printf("id: %i\n", gi.id);
  {
    const uint32_t P = gi.P;
    auto m = [&](uint32_t x, uint32_t y) { return mult[(size_t)(x % P) * P + (y % P)]; };
    printf("check stuff: (a*b)*c: %d a*(b*c): %d\n", m(m(1, 2), 3), m(1, m(2, 3)));
  }
  printf("min_retr: %e nn_retr: %e \n", gi.min_retr, gi.nn_retr);
  printf("small group size: %lu\n", (unsigned long)gi.npool);
  printf("START: %s\n", cold ? "cold" : "hot");
  printf("Initializing nearest neighbours [%d:%d]... \n", (int)p.Nt, (int)p.Nx);
  dsa_chain *c = nullptr;
  if (dsa_chain_create(g, &p, &c)) return die("creating the chain");
  printf("init done\n");
  fflush(stdout);
  if (NT && dsa_chain_sweep(c, NT)) return die("thermalising");
  printf("thermo done\n");
  fflush(stdout);

  std::vector<double> m(4 + (size_t)wk * wl);
  for (unsigned n = 0; n < N; ++n) {
    Timer tm;
    tm.start("update");
    if (K && dsa_chain_sweep(c, K)) return die("sweeping");
    tm.stop();
    tm.start("meas");
    if (dsa_chain_measure(c, (uint32_t)wk, (uint32_t)wl, sum, m.data())) return die("measuring");
    printf("GMES: %e %e %e %e", 999.0, m[0], m[1], m[2]);
    for (int q = 0; q < wk * wl; ++q) printf(" %e", m[4 + q]);
    printf("\n");
    printf("RMES: %e %e\n", m[2], m[3]);
    fflush(stdout);
    uint64_t hit = 0, acc = 0;
    if (dsa_chain_counters(c, &hit, &acc)) return die("reading counters");
    printf("ACC: %f\n", ((double)acc) / hit);
    tm.stop();
  }
  dsa_chain_destroy(c);
  dsa_group_free(g);
  return 0;
}
