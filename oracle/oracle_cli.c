/*
 * oracle_cli.c -- the oracle driven like dym-mod-metro (TEST INFRASTRUCTURE).
 *
 *   oracle_metro <group> D Nt Nx beta0 beta1 beta2 seed [K] [N] [Ntherm] [hot|cold]
 *                [--rng minstd|philox] [--order reference|colour] [--sum ordered|hist]
 *                [--wgrid KxL]
 *
 * Prints the reference's stdout lines (dym-mod-metro.cpp:217-329) except the
 * TM_BEGIN/TM_END timer lines.  Defaults reproduce the reference at one thread.
 */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv) {
  int rng = OG_RNG_MINSTD, order = OG_ORDER_REFERENCE, sum = 0;
  int wk = -1, wl = -1;
  char *pos[32];
  int np = 0;
  for (int i = 1; i < argc; ++i) {
    if (!strcmp(argv[i], "--rng") && i + 1 < argc) {
      ++i;
      rng = !strcmp(argv[i], "philox") ? OG_RNG_PHILOX : OG_RNG_MINSTD;
    } else if (!strcmp(argv[i], "--order") && i + 1 < argc) {
      ++i;
      order = !strcmp(argv[i], "colour") ? OG_ORDER_COLOUR : OG_ORDER_REFERENCE;
    } else if (!strcmp(argv[i], "--sum") && i + 1 < argc) {
      ++i;
      sum = !strcmp(argv[i], "hist") ? 1 : 0;
    } else if (!strcmp(argv[i], "--wgrid") && i + 1 < argc) {
      ++i;
      if (sscanf(argv[i], "%dx%d", &wk, &wl) != 2) return 1;
    } else if (np < 32) {
      pos[np++] = argv[i];
    }
  }
  if (np < 8) {
    fprintf(stderr, "usage: %s group D Nt Nx beta0 beta1 beta2 seed [K] [N] [Ntherm] [hot|cold]\n", argv[0]);
    return 1;
  }
  const char *gf = pos[0];
  unsigned D = atoi(pos[1]), Nt = atoi(pos[2]), Nx = atoi(pos[3]);
  double beta0 = atof(pos[4]), beta1 = atof(pos[5]), beta2 = atof(pos[6]);
  int iseed = atoi(pos[7]);
  unsigned K = np > 8 ? (unsigned)atoi(pos[8]) : 1000;
  unsigned N = np > 9 ? (unsigned)atoi(pos[9]) : 1000;
  unsigned NT = np > 10 ? (unsigned)atoi(pos[10]) : 0;
  const char *sk = np > 11 ? pos[11] : "hot";
  int cold = sk[0] == 'c' || sk[0] == 'C';
  if (wk < 0) { wk = (int)Nx; wl = (int)Nt; }
  printf("PARAMS(grp,D,Nt,Nx,beta0,beta1,beta2,seed): %s %d %d %d %e %e %e %d\n", gf, D, Nt, Nx,
         beta0, beta1, beta2, iseed);
  printf("RUN(K,N,Ntherm,NHIT): %u %u %u %d\n", K, N, NT, OG_NHIT);
  og_group g;
  if (og_group_load(gf, &g)) { fprintf(stderr, "cannot load group %s\n", gf); return 2; }
  const unsigned P = g.P;
  printf("read the group\n");
  printf("id: %i\n", g.id);
  printf("check stuff: (a*b)*c: %d a*(b*c): %d\n", g.mult[(size_t)g.mult[1 * P + 2] * P + 3],
         g.mult[1 * P + g.mult[2 * P + 3]]);
  printf("min_retr: %e nn_retr: %e \n", g.min_retr, g.nn_retr);
  printf("small group size: %lu\n", (unsigned long)g.npool);
  printf("START: %s\n", cold ? "cold" : "hot");
  og_chain c;
  if (og_chain_init(&c, &g, D, Nt, Nx, beta0, beta1, beta2, iseed, cold, rng, order)) return 3;
  printf("Initializing nearest neighbours [%d:%d]... \n", Nt, Nx);
  printf("init done\n");
  og_chain_sweep(&c, NT);
  printf("thermo done\n");
  double *m = (double *)malloc(sizeof(double) * (4 + (size_t)wk * wl));
  for (unsigned n = 0; n < N; ++n) {
    og_chain_sweep(&c, K);
    og_measure(&c.L, (unsigned)wk, (unsigned)wl, sum, m);
    printf("GMES: %e %e %e %e", 999.0, m[0], m[1], m[2]);
    for (int q = 0; q < wk * wl; ++q) printf(" %e", m[4 + q]);
    printf("\n");
    printf("RMES: %e %e\n", m[2], m[3]);
    printf("ACC: %f\n", ((double)c.acc) / c.hit);
    fflush(stdout);
  }
  free(m);
  og_chain_free(&c);
  og_group_free(&g);
  return 0;
}
