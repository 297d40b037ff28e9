/*
 * oracle_llr.c -- CPU restatement of the reference's 1-D LLR driver
 * (llr.cpp:46-345) at one thread, and of the device LLR algorithm.
 * TEST INFRASTRUCTURE ONLY.
 *
 *   oracle_llr group D Nt Nx axis bg Etop Ebot step hw a0 seed [K] [NRM] [R]
 * prints what the reference llr prints (llr.cpp:299-341).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_llr.h"

#define MUL(x, y) (g->mult[(size_t)(x) * g->P + (unsigned)(y)])

/* ------------------------------------------------- reference llr.cpp (serial) */

static double ref_S1(const og_lat *L) { return og_S1(L); }
static double ref_S2(const ogl_ref *R) { return og_sum_m1(&R->c.L) / R->rect_mult; }
static double Ecur(const ogl_ref *R) { return R->axis == 1 ? R->S1 : R->S2; }
static double recompute_Ewin(const ogl_ref *R) { return R->axis == 1 ? ref_S1(&R->c.L) : ref_S2(R); }
static void resync(ogl_ref *R) {
  R->S1 = ref_S1(&R->c.L);
  R->S2 = ref_S2(R);
}

/* llr.cpp:155-171 */
void ogl_ref_calibrate(ogl_ref *R) {
  og_lat *L = &R->c.L;
  const og_group *g = L->g;
  const unsigned D = L->D;
  const int nr = 6 * (int)(D - 1);
  int32_t re[6 * OG_DMAX];
  unsigned i = L->V / 2, d = 0;
  og_rect(L, i, d, re);
  int32_t b = L->a[(size_t)i * D + d];
  double m1_old = 0;
  for (int k = 0; k < nr; ++k) m1_old += g->retr[MUL(b, re[k])];
  double T_old = og_sum_m1(L);
  int32_t bnew = MUL(b, g->pool[g->npool / 2]);
  double m1_new = 0;
  for (int k = 0; k < nr; ++k) m1_new += g->retr[MUL(bnew, re[k])];
  L->a[(size_t)i * D + d] = bnew;
  double T_new = og_sum_m1(L);
  L->a[(size_t)i * D + d] = b;
  R->rect_mult = (T_new - T_old) / (m1_new - m1_old);
}

/* llr.cpp:178-219: serial over d, then i; window test on the running energy */
void ogl_ref_update_constrained(ogl_ref *R, double a_tilt, double Elo, double Ehi, int driveIn) {
  og_lat *L = &R->c.L;
  const og_group *g = L->g;
  const unsigned D = L->D, V = L->V;
  const int ns = 2 * (int)(D - 1), nr = 6 * (int)(D - 1);
  const double Emid = 0.5 * (Elo + Ehi);
  int32_t st[2 * OG_DMAX], re[6 * OG_DMAX];
  for (unsigned d = 0; d < D; d++)
    for (unsigned i = 0; i < V; i++) {
      og_staples(L, i, d, st);
      og_rect(L, i, d, re);
      const size_t l = (size_t)i * D + d;
      int32_t b = L->a[l];
      double r1 = 0, m1 = 0;
      for (int j = 0; j < ns; ++j) r1 += g->retr[MUL(b, st[j])];
      for (int k = 0; k < nr; ++k) m1 += g->retr[MUL(b, re[k])];
      for (int h = 0; h < OG_NHIT; ++h) {
        int32_t bnew = MUL(b, g->pool[(size_t)((double)g->npool * og_canonical(&R->c.ms[i]))]);
        double r1n = 0, m1n = 0;
        for (int j = 0; j < ns; ++j) r1n += g->retr[MUL(bnew, st[j])];
        for (int k = 0; k < nr; ++k) m1n += g->retr[MUL(bnew, re[k])];
        double dS1 = r1n - r1, dS2 = m1n - m1;
        double dEw = R->axis == 1 ? dS1 : dS2;
        double dEbg = R->axis == 1 ? dS2 : dS1;
        double Ec = R->axis == 1 ? R->S1 : R->S2;
        double Enew = Ec + dEw;
        int accept = 0;
        if (driveIn && (Ec < Elo || Ec > Ehi)) {
          accept = fabs(Enew - Emid) < fabs(Ec - Emid);
        } else if (Enew >= Elo && Enew <= Ehi) {
          double w = og_exp(-a_tilt * dEw - R->bg * dEbg);
          accept = w > og_canonical(&R->c.ms[i]);
        }
        R->c.hit++;
        if (accept) {
          L->a[l] = bnew;
          R->S1 += dS1;
          R->S2 += dS2;
          b = bnew;
          r1 = r1n;
          m1 = m1n;
          R->c.acc++;
        }
      }
    }
}

/* llr.cpp:223-231 */
int ogl_ref_seed_window(ogl_ref *R, double Elo, double Ehi, double a_guess) {
  unsigned drive = 0, CAP = 20000;
  while ((Ecur(R) < Elo || Ecur(R) > Ehi) && drive < CAP) {
    ogl_ref_update_constrained(R, a_guess, Elo, Ehi, 1);
    ++drive;
  }
  resync(R);
  return Ecur(R) >= Elo && Ecur(R) <= Ehi;
}

/* llr.cpp:234-250 */
double ogl_ref_rm_solve(ogl_ref *R, double Ewin, double hw, double a0, unsigned K, unsigned NRM) {
  double Elo = Ewin - hw, Ehi = Ewin + hw, W = 2 * hw, a = a0;
  for (unsigned m = 0; m < NRM; ++m) {
    double sum = 0;
    for (unsigned k = 0; k < K; ++k) {
      ogl_ref_update_constrained(R, a, Elo, Ehi, 0);
      sum += Ecur(R);
    }
    double gr = sum / K - Ewin;
    a += (12.0 / (W * W * (m + 1))) * gr;
    double drift = fabs(Ecur(R) - recompute_Ewin(R));
    if (drift > R->maxdrift) R->maxdrift = drift;
    resync(R);
  }
  return a;
}

/* llr.cpp:256-274 */
static int run_pass(ogl_ref *R, int dir, double a0, double Etop, double stepE, double hw, unsigned K, unsigned NRM,
                    int M, double *an, int *nan_, int maxrep) {
  og_lat *L = &R->c.L;
  const og_group *g = L->g;
  const size_t nl = (size_t)L->V * L->D;
  if (dir > 0)
    for (size_t i = 0; i < nl; ++i) L->a[i] = og_randint(&R->c.ms[0], 0, (int)g->P - 1);
  else
    for (size_t i = 0; i < nl; ++i) L->a[i] = g->id;
  resync(R);
  double a_prev = a0;
  int reached = 0;
  for (int s = 0; s < M; ++s) {
    int n = dir > 0 ? s : M - 1 - s;
    double Ewin = Etop - n * stepE, Elo = Ewin - hw, Ehi = Ewin + hw;
    if (!ogl_ref_seed_window(R, Elo, Ehi, a_prev)) break;
    double v = ogl_ref_rm_solve(R, Ewin, hw, a_prev, K, NRM);
    an[n * maxrep + nan_[n]++] = v;
    a_prev = v;
    ++reached;
  }
  return reached;
}

int ogl_ref_init(ogl_ref *R, const og_group *g, unsigned D, unsigned Nt, unsigned Nx, int axis, double bg,
                 int seed) {
  memset(R, 0, sizeof *R);
  if (og_chain_init(&R->c, g, D, Nt, Nx, 0, 0, 0, seed, 1, OG_RNG_MINSTD, OG_ORDER_REFERENCE)) return -1;
  /* llr.cpp:314-315: rnd[i].seed(iseed + i) with i unsigned, so the seed is
   * the 32-bit unsigned sum, zero-extended (dym-mod-metro's int i
   * sign-extends); then the hot links from site 0's stream */
  const unsigned V = R->c.L.V;
  for (unsigned i = 0; i < V; ++i) R->c.ms[i] = og_minstd_seed((long long)(uint32_t)((uint32_t)seed + i));
  for (size_t l = 0; l < (size_t)V * D; ++l) R->c.L.a[l] = og_randint(&R->c.ms[0], 0, (int)g->P - 1);
  R->axis = axis;
  R->bg = bg;
  R->rect_mult = 6.0;
  ogl_ref_calibrate(R);
  return 0;
}

int ogl_ref_main(int argc, char **argv) {
  if (argc < 13) {
    fprintf(stderr, "usage: %s group D Nt Nx axis bg Etop Ebot step hw a0 seed [K] [NRM] [R]\n", argv[0]);
    return 1;
  }
  const char *gf = argv[1];
  unsigned D = atoi(argv[2]), Nt = atoi(argv[3]), Nx = atoi(argv[4]);
  int axis = atoi(argv[5]);
  double bg = atof(argv[6]), Etop = atof(argv[7]), Ebot = atof(argv[8]), stepE = atof(argv[9]);
  double hw = atof(argv[10]), a0 = atof(argv[11]);
  int iseed = atoi(argv[12]);
  unsigned K = argc > 13 ? (unsigned)atoi(argv[13]) : 50;
  unsigned NRM = argc > 14 ? (unsigned)atoi(argv[14]) : 80;
  unsigned Rr = argc > 15 ? (unsigned)atoi(argv[15]) : 1;
  int M = (int)((Etop - Ebot) / stepE + 0.5) + 1;
  printf("LLR(grp,D,Nt,Nx,axis,bg): %s %d %d %d %d %g | Etop=%g Ebot=%g step=%g hw=%g a0=%g seed=%d K=%u NRM=%u R=%u | %d windows\n",
         gf, D, Nt, Nx, axis, bg, Etop, Ebot, stepE, hw, a0, iseed, K, NRM, Rr, M);
  unsigned V = Nt;
  for (unsigned d = 1; d < D; ++d) V *= Nx;
  printf("NPLAQ: %u\n", V * D * (D - 1) / 2);
  printf("NRECT: %u\n", V * D * (D - 1));
  printf("AXIS: %d\n", axis);
  og_group g;
  if (og_group_load(gf, &g)) return 2;
  ogl_ref R;
  printf("Initializing nearest neighbours [%d:%d]... \n", Nt, Nx);
  if (ogl_ref_init(&R, &g, D, Nt, Nx, axis, bg, iseed)) return 3;
  printf("RECT_MULT calibrated: %.6f\n", R.rect_mult);
  if (axis == 2 && fabs(R.rect_mult - 6.0) > 1e-3)
    printf("WARNING: RECT_MULT=%.3f != 6 -- the 1x2 rectangles wrap on this lattice\n"
           "  (extent too small in some direction); the S2 action is degenerate and the\n"
           "  S2 window is invalid. Use a lattice with extent >= 3 for axis=2.\n",
           R.rect_mult);
  double *an = (double *)calloc((size_t)M * 2 * Rr, sizeof(double));
  int *nan_ = (int *)calloc((size_t)M, sizeof(int));
  for (unsigned r = 0; r < Rr; ++r) {
    for (unsigned i = 0; i < V; ++i) R.c.ms[i] = og_minstd_seed((long long)(uint32_t)((uint32_t)iseed + r * 1000003u + i));
    int rd = run_pass(&R, +1, a0, Etop, stepE, hw, K, NRM, M, an, nan_, 2 * Rr);
    int ru = run_pass(&R, -1, a0, Etop, stepE, hw, K, NRM, M, an, nan_, 2 * Rr);
    printf("repeat %u: down reached %d/%d, up reached %d/%d\n", r, rd, M, ru, M);
  }
  printf("\n# Ewin   a_n        sd        nrep\n");
  for (int n = 0; n < M; ++n) {
    double Ewin = Etop - n * stepE;
    if (!nan_[n]) {
      printf("ANE: %8.2f   --(unreached)\n", Ewin);
      continue;
    }
    double mean = 0;
    for (int k = 0; k < nan_[n]; ++k) mean += an[n * 2 * Rr + k];
    mean /= nan_[n];
    double var = 0;
    for (int k = 0; k < nan_[n]; ++k) var += (an[n * 2 * Rr + k] - mean) * (an[n * 2 * Rr + k] - mean);
    double sd = nan_[n] > 1 ? sqrt(var / (nan_[n] - 1)) : 0.0;
    printf("ANE: %8.2f  %9.5f  %9.5f  %lu\n", Ewin, mean, sd, (unsigned long)nan_[n]);
  }
  printf("max |incr - recompute| E(axis%d) : %.3e\n", axis, R.maxdrift);
  free(an);
  free(nan_);
  og_chain_free(&R.c);
  og_group_free(&g);
  return 0;
}

/* ------------------------------------------- the device LLR algorithm */

long long ogl_fix(double x) { return (long long)(x * 4294967296.0 + (x >= 0.0 ? 0.5 : -0.5)); }
static double ogl_unfix(long long v) { return (double)v / 4294967296.0; }

int ogl_dev_init(ogl_dev *L, const og_group *g, unsigned D, unsigned Nt, unsigned Nx, int axis, double bg,
                 double Ewin, double hw, double a0, int seed, int cold, unsigned K, unsigned NRM, double tol,
                 unsigned cap, double Ewin2, double hw2, double a20) {
  memset(L, 0, sizeof *L);
  if (og_chain_init(&L->c, g, D, Nt, Nx, 0, 0, 0, seed, cold, OG_RNG_PHILOX, OG_ORDER_COLOUR)) return -1;
  L->axis = axis;
  L->bg = bg;
  L->Ewin = Ewin;
  L->W = 2 * hw;
  L->lo = ogl_fix(Ewin - hw);
  L->hi = ogl_fix(Ewin + hw);
  L->W2 = 1.0;
  if (axis == 3) {
    L->Ewin2 = Ewin2;
    L->W2 = 2 * hw2;
    L->lo2 = ogl_fix(Ewin2 - hw2);
    L->hi2 = ogl_fix(Ewin2 + hw2);
    L->a2 = a20;
  }
  L->K = K;
  L->NRM = NRM;
  L->tol = tol;
  L->cap = cap ? cap : 20000;
  L->a = a0;
  L->phase = OGL_SEED;
  L->resync = 2; /* initial energies: computed, not compared */
  L->old = (int32_t *)malloc(sizeof(int32_t) * L->c.L.V);
  L->d1 = (long long *)malloc(sizeof(long long) * L->c.L.V);
  L->d2 = (long long *)malloc(sizeof(long long) * L->c.L.V);
  L->na = (unsigned *)malloc(sizeof(unsigned) * L->c.L.V);
  return 0;
}

void ogl_dev_free(ogl_dev *L) {
  og_chain_free(&L->c);
  free(L->old);
  free(L->d1);
  free(L->d2);
  free(L->na);
}

static void dev_resync(ogl_dev *L) {
  if (!L->resync) return;
  const og_lat *Lt = &L->c.L;
  const og_group *g = Lt->g;
  const unsigned D = Lt->D;
  long long r1 = 0, r2 = 0;
  for (unsigned i = 0; i < Lt->V; ++i) {
    for (unsigned d1 = 0; d1 < D; ++d1)
      for (unsigned d2 = d1 + 1; d2 < D; ++d2) r1 += ogl_fix(g->retr[og_plaquette(Lt, i, d1, d2)]);
    for (unsigned d1 = 0; d1 < D; ++d1)
      for (unsigned d2 = 0; d2 < D; ++d2)
        if (d1 != d2) r2 += ogl_fix(g->retr[og_wilson(Lt, i, d1, d2, 2, 1)]);
  }
  if (L->resync == 1) {
    const double d1 = fabs(ogl_unfix(L->S1) - ogl_unfix(r1)), d2 = fabs(ogl_unfix(L->S2) - ogl_unfix(r2));
    const double drift = L->axis == 1 ? d1 : L->axis == 2 ? d2 : fmax(d1, d2);
    if (drift > L->maxdrift) L->maxdrift = drift;
    if (d1 > L->drift1) L->drift1 = d1;
    if (d2 > L->drift2) L->drift2 = d2;
  }
  L->S1 = r1;
  L->S2 = r2;
  L->resync = 0;
}

static long long llabs_(long long v) { return v < 0 ? -v : v; }

static int dev_inside(const ogl_dev *L, long long E1, long long E2) {
  if (L->axis == 3) return E1 >= L->lo && E1 <= L->hi && E2 >= L->lo2 && E2 <= L->hi2;
  const long long Ew = L->axis == 1 ? E1 : E2;
  return Ew >= L->lo && Ew <= L->hi;
}

/* the pass's acceptance rule on the running energies before (b) / after (a) */
typedef struct {
  int axis, drive;
  long long lo, hi, lo2, hi2, mid, mid2;
  double hw, hw2;
} dev_rule_t;

static dev_rule_t dev_rule_of(const ogl_dev *L, int drive) {
  dev_rule_t R;
  R.axis = L->axis;
  R.drive = drive;
  R.lo = L->lo;
  R.hi = L->hi;
  R.lo2 = L->lo2;
  R.hi2 = L->hi2;
  R.mid = L->lo + (L->hi - L->lo) / 2;
  R.mid2 = L->lo2 + (L->hi2 - L->lo2) / 2;
  R.hw = 0.5 * L->W;
  R.hw2 = 0.5 * L->W2;
  return R;
}

static double dev_dist2(const dev_rule_t *R, long long E1, long long E2) {
  const double x = ogl_unfix(E1 - R->mid) / R->hw, y = ogl_unfix(E2 - R->mid2) / R->hw2;
  return x * x + y * y;
}

static int dev_rule(const dev_rule_t *R, long long E1b, long long E2b, long long E1a, long long E2a) {
  if (R->axis == 3) {
    if (R->drive) return dev_dist2(R, E1a, E2a) < dev_dist2(R, E1b, E2b);
    return E1a >= R->lo && E1a <= R->hi && E2a >= R->lo2 && E2a <= R->hi2;
  }
  const long long Eb = R->axis == 1 ? E1b : E2b, Ea = R->axis == 1 ? E1a : E2a;
  if (R->drive) return llabs_(Ea - R->mid) < llabs_(Eb - R->mid);
  return Ea >= R->lo && Ea <= R->hi;
}

/* One colour pass: every link proposes (NHIT tilted or greedy hits), then the
 * proposals are accepted in ascending site order while the running windowed
 * energy obeys the rule; refused links are restored. */
static void dev_pass(ogl_dev *L, unsigned mu, unsigned colour) {
  og_lat *Lt = &L->c.L;
  const og_group *g = Lt->g;
  const unsigned D = Lt->D, V = Lt->V;
  const int ns = 2 * (int)(D - 1), nr = 6 * (int)(D - 1);
  const long long Ew = L->axis == 1 ? L->S1 : L->S2;
  const int drive = L->phase == OGL_SEED && !dev_inside(L, L->S1, L->S2);
  const dev_rule_t R = dev_rule_of(L, drive);
  const double sign = ogl_unfix(Ew) < L->Ewin ? 1.0 : -1.0;
  const double gx = ogl_unfix(L->S1 - R.mid) / R.hw / R.hw, gy = ogl_unfix(L->S2 - R.mid2) / R.hw2 / R.hw2;
  const uint32_t key[2] = {(uint32_t)L->c.seed, OG_KEY_LLR};
  const uint32_t swl = (uint32_t)L->sweep, swh = (uint32_t)(L->sweep >> 32);
  const uint8_t *col = L->c.col + (size_t)mu * V;
  unsigned np = 0;
  int32_t st[2 * OG_DMAX], re[6 * OG_DMAX];
  for (unsigned i = 0; i < V; ++i) {
    if (col[i] != colour) continue;
    og_staples(Lt, i, mu, st);
    og_rect(Lt, i, mu, re);
    const size_t l = (size_t)i * D + mu;
    int32_t b = Lt->a[l];
    const int32_t b0 = b;
    double r1 = 0, m1 = 0;
    long long q1 = 0, q2 = 0;
    for (int j = 0; j < ns; ++j) {
      double v = g->retr[MUL(b, st[j])];
      r1 += v;
      q1 += ogl_fix(v);
    }
    for (int k = 0; k < nr; ++k) {
      double v = g->retr[MUL(b, re[k])];
      m1 += v;
      q2 += ogl_fix(v);
    }
    unsigned nacc = 0;
    for (int h = 0; h < OG_NHIT; ++h) {
      uint32_t ctr[4] = {(uint32_t)l, (uint32_t)h, swl, swh}, w[4];
      og_philox4x32_10(ctr, key, w);
      const double u1 = og_u53(w[0], w[1]), u2 = og_u53(w[2], w[3]);
      const int32_t bnew = MUL(b, g->pool[(size_t)((double)g->npool * u1)]);
      double r1n = 0, m1n = 0;
      for (int j = 0; j < ns; ++j) r1n += g->retr[MUL(bnew, st[j])];
      for (int k = 0; k < nr; ++k) m1n += g->retr[MUL(bnew, re[k])];
      const double dS1 = r1n - r1, dS2 = m1n - m1;
      int accept;
      if (L->axis == 3) {
        if (drive)
          accept = gx * dS1 + gy * dS2 < 0.0;
        else
          accept = og_exp(-L->a * dS1 - L->a2 * dS2) > u2;
      } else {
        const double dEw = L->axis == 1 ? dS1 : dS2, dEbg = L->axis == 1 ? dS2 : dS1;
        if (drive)
          accept = sign * dEw > 0.0;
        else
          accept = og_exp(-L->a * dEw - L->bg * dEbg) > u2;
      }
      if (accept) {
        b = bnew;
        r1 = r1n;
        m1 = m1n;
        ++nacc;
      }
    }
    long long d1 = 0, d2 = 0;
    if (b != b0) {
      Lt->a[l] = b;
      long long p1 = 0, p2 = 0;
      for (int j = 0; j < ns; ++j) p1 += ogl_fix(g->retr[MUL(b, st[j])]);
      for (int k = 0; k < nr; ++k) p2 += ogl_fix(g->retr[MUL(b, re[k])]);
      d1 = p1 - q1;
      d2 = p2 - q2;
    }
    L->old[np] = b0;
    L->d1[np] = d1;
    L->d2[np] = d2;
    L->na[np] = nacc;
    np++;
  }
  /* decisions in ascending site order */
  unsigned q = 0;
  for (unsigned i = 0; i < V; ++i) {
    if (col[i] != colour) continue;
    if (dev_rule(&R, L->S1, L->S2, L->S1 + L->d1[q], L->S2 + L->d2[q])) {
      L->S1 += L->d1[q];
      L->S2 += L->d2[q];
      L->acc += L->na[q];
    } else {
      Lt->a[(size_t)i * D + mu] = L->old[q];
    }
    ++q;
  }
}

static void dev_sweep_end(ogl_dev *L) {
  L->sweep++;
  L->hit += (uint64_t)L->c.L.V * L->c.L.D * OG_NHIT;
  const long long Ew = L->axis == 2 ? L->S2 : L->S1; /* the S1 window for axes 1 and 3 */
  if (L->phase == OGL_SEED) {
    L->seed_sweeps++;
    if (dev_inside(L, L->S1, L->S2)) {
      L->phase = L->NRM ? OGL_RM : OGL_DONE; /* NRM = 0: drive-in only */
      L->reached = 1;
      L->k = 0;
      L->m = 0;
      L->sumE = 0;
      L->sumE2 = 0;
      L->resync = 1;
    } else if (L->seed_sweeps >= L->cap) {
      L->phase = OGL_DONE;
      L->reached = 0;
    }
    return;
  }
  L->sumE += ogl_unfix(Ew);
  if (L->axis == 3) L->sumE2 += ogl_unfix(L->S2);
  L->k++;
  if (L->k == L->K) {
    const double a_old = L->a, a2_old = L->a2;
    const double gr = L->sumE / L->K - L->Ewin;
    L->a += (12.0 / (L->W * L->W * (L->m + 1))) * gr;
    L->last_da = fabs(L->a - a_old);
    if (L->axis == 3) { /* llr2d.cpp:247-248 */
      const double g2 = L->sumE2 / L->K - L->Ewin2;
      L->a2 += (12.0 / (L->W2 * L->W2 * (L->m + 1))) * g2;
      L->last_da = fmax(L->last_da, fabs(L->a2 - a2_old));
      L->sumE2 = 0;
    }
    L->m++;
    L->k = 0;
    L->sumE = 0;
    L->resync = 1;
    if ((L->tol > 0.0 && L->last_da < L->tol) || L->m >= L->NRM) L->phase = OGL_DONE;
  }
}

void ogl_dev_step(ogl_dev *L, unsigned n) {
  dev_resync(L);
  for (unsigned it = 0; it < n; ++it) {
    if (L->phase == OGL_SEED || L->phase == OGL_RM) {
      for (unsigned mu = 0; mu < L->c.L.D; ++mu)
        for (unsigned c = 0; c < L->c.ncol[mu]; ++c) {
          int any = 0;
          const uint8_t *col = L->c.col + (size_t)mu * L->c.L.V;
          for (unsigned i = 0; i < L->c.L.V && !any; ++i) any = col[i] == c;
          if (any && (L->phase == OGL_SEED || L->phase == OGL_RM)) dev_pass(L, mu, c);
        }
      dev_sweep_end(L);
    }
    dev_resync(L);
  }
}
