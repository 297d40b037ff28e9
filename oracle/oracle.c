/*
 * oracle.c -- CPU restatement of the reference's Metropolis hot path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker, never the product.
 * Each function cites the reference file:line it follows.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

/* ------------------------------------------------------------------ group */

static char *slurp(const char *path, size_t *len) {
  gzFile f = gzopen(path, "rb");
  if (!f) return NULL;
  size_t cap = 1 << 20, n = 0;
  char *buf = (char *)malloc(cap + 1);
  for (;;) {
    if (n == cap) {
      cap *= 2;
      buf = (char *)realloc(buf, cap + 1);
    }
    int r = gzread(f, buf + n, (unsigned)(cap - n));
    if (r < 0) { free(buf); gzclose(f); return NULL; }
    if (r == 0) break;
    n += (size_t)r;
  }
  gzclose(f);
  buf[n] = 0;
  *len = n;
  return buf;
}

/* group.c:13-81: P, P ReTr, P ImTr, P*P table; trailing tokens ignored;
 * identity = first n whose row and column are the identity map; inverse of n
 * = first m with n*m = m*n = id.  Proposal pool: dym-mod-metro.cpp:239-245. */
int og_group_load(const char *path, og_group *g) {
  memset(g, 0, sizeof *g);
  size_t len;
  char *buf = slurp(path, &len);
  if (!buf) return -1;
  char *p = buf, *e;
  long Pl = strtol(p, &e, 10);
  if (e == p || Pl <= 0 || Pl > 5040) { free(buf); return -2; }
  p = e;
  unsigned P = (unsigned)Pl;
  g->P = P;
  g->retr = (double *)malloc(sizeof(double) * P);
  g->imtr = (double *)malloc(sizeof(double) * P);
  g->mult = (int32_t *)malloc(sizeof(int32_t) * (size_t)P * P);
  g->inv = (int32_t *)malloc(sizeof(int32_t) * P);
  for (unsigned n = 0; n < P; n++) {
    g->retr[n] = strtod(p, &e);
    if (e == p) goto bad;
    p = e;
  }
  for (unsigned n = 0; n < P; n++) {
    g->imtr[n] = strtod(p, &e);
    if (e == p) goto bad;
    p = e;
  }
  for (size_t n = 0; n < (size_t)P * P; n++) {
    long v = strtol(p, &e, 10);
    if (e == p || v < 0 || v >= (long)P) goto bad;
    g->mult[n] = (int32_t)v;
    p = e;
  }
  free(buf);
  buf = NULL;
  {
    int found = 0;
    for (unsigned n = 0; n < P && !found; n++) {
      int is_id = 1;
      for (unsigned m = 0; m < P; m++)
        if (g->mult[(size_t)n * P + m] != (int32_t)m || g->mult[(size_t)m * P + n] != (int32_t)m) {
          is_id = 0;
          break;
        }
      if (is_id) { g->id = (int32_t)n; found = 1; }
    }
    if (!found) { og_group_free(g); return -3; }
    for (unsigned n = 0; n < P; n++) {
      int f = 0;
      for (unsigned m = 0; m < P; m++)
        if (g->mult[(size_t)n * P + m] == g->id && g->mult[(size_t)m * P + n] == g->id) {
          g->inv[n] = (int32_t)m;
          f = 1;
          break;
        }
      if (!f) { og_group_free(g); return -4; }
    }
  }
  {
    double min_retr = 0, nn_retr = 10;
    for (unsigned i = 0; i < P; ++i) if (g->retr[i] < min_retr) min_retr = g->retr[i];
    for (unsigned i = 0; i < P; ++i)
      if (g->retr[i] > min_retr && g->retr[i] < nn_retr) nn_retr = g->retr[i];
    g->pool = (int32_t *)malloc(sizeof(int32_t) * P);
    g->npool = 0;
    for (unsigned i = 0; i < P; ++i)
      if (g->retr[i] < nn_retr + 1e-6 && g->retr[i] > min_retr) g->pool[g->npool++] = (int32_t)i;
    g->min_retr = min_retr;
    g->nn_retr = nn_retr;
  }
  return 0;
bad:
  free(buf);
  og_group_free(g);
  return -2;
}

void og_group_free(og_group *g) {
  free(g->retr); free(g->imtr); free(g->mult); free(g->inv); free(g->pool);
  memset(g, 0, sizeof *g);
}

/* --------------------------------------------------------------- lattice */

/* V = Nt*Nx^(D-1) (dym-mod-metro.cpp:222-223); site index
 * i = x0 + Nt*(x1 + Nx*(x2 + ...)) (lattice.c:39-51). */
int og_lat_init(og_lat *L, const og_group *g, unsigned D, unsigned Nt, unsigned Nx) {
  memset(L, 0, sizeof *L);
  if (D < 2 || D > OG_DMAX || Nt == 0 || Nx == 0) return -1;
  L->D = D; L->Nt = Nt; L->Nx = Nx; L->g = g;
  unsigned V = 1;
  for (unsigned d = 0; d < D; ++d) {
    L->ext[d] = d == 0 ? Nt : Nx;
    L->stride[d] = V;
    V *= L->ext[d];
  }
  L->V = V;
  L->a = (int32_t *)malloc(sizeof(int32_t) * (size_t)V * D);
  return 0;
}

void og_lat_free(og_lat *L) { free(L->a); L->a = NULL; }

void og_getpos(const og_lat *L, unsigned i, unsigned *pos) {
  for (unsigned d = 0; d < L->D; ++d) { pos[d] = i % L->ext[d]; i /= L->ext[d]; }
}

/* lattice.c:22-26 with the table of init_nn (55-72) computed on the fly. */
unsigned og_step(const og_lat *L, unsigned i, unsigned d, int s) {
  unsigned n = L->ext[d], st = L->stride[d];
  unsigned x = (i / st) % n;
  unsigned xn = s > 0 ? (x + 1 == n ? 0 : x + 1) : (x == 0 ? n - 1 : x - 1);
  return i + (xn - x) * st; /* unsigned wrap is exact */
}

#define MUL(x, y) (g->mult[(size_t)(x) * P + (unsigned)(y)])
#define INV(x) (g->inv[(x)])

/* Reads link (s,dir); records its position when rs != NULL. */
static inline int32_t rd(const og_lat *L, unsigned s, unsigned dir, uint32_t *rs, unsigned *nrs) {
  if (rs) rs[(*nrs)++] = s * L->D + dir;
  return L->a[(size_t)s * L->D + dir];
}

/* dym-mod-metro.cpp:63-78 (st != NULL) and :83-142 (re != NULL). */
static void staples_impl(const og_lat *L, unsigned i, unsigned d, int32_t *st, int32_t *re,
                         uint32_t *rs, unsigned *nrs) {
  const og_group *g = L->g;
  const unsigned P = g->P, D = L->D;
#define STEP(x, dd, s) og_step(L, (x), (dd), (s))
#define A(x, dd) rd(L, (x), (dd), rs, nrs)
  unsigned i1 = STEP(i, d, 1);
  int k = 0;
  if (st) {
    for (unsigned d1 = 0; d1 < D; ++d1) if (d1 != d) {
      int32_t gg = A(i1, d1);
      gg = MUL(gg, INV(A(STEP(i, d1, 1), d)));
      gg = MUL(gg, INV(A(i, d1)));
      st[k++] = gg;
      gg = INV(A(STEP(i1, d1, -1), d1));
      unsigned i2 = STEP(i, d1, -1);
      gg = MUL(gg, INV(A(i2, d)));
      gg = MUL(gg, A(i2, d1));
      st[k++] = gg;
    }
  }
  if (re) {
    const int32_t id = g->id;
    k = 0;
    for (unsigned d1 = 0; d1 < D; ++d1) if (d1 != d) {
      int32_t gg;
      unsigned i2;
      gg = id;
      gg = MUL(gg, A(STEP(i, d, 1), d1));
      i2 = STEP(STEP(i, d, 1), d1, 1);
      gg = MUL(gg, A(i2, d1));
      gg = MUL(gg, INV(A(STEP(STEP(i2, d1, 1), d, -1), d)));
      gg = MUL(gg, INV(A(STEP(i, d1, 1), d1)));
      gg = MUL(gg, INV(A(i, d1)));
      re[k++] = gg;

      gg = id;
      gg = MUL(gg, A(STEP(i, d, 1), d));
      i2 = STEP(STEP(i, d, 1), d, 1);
      gg = MUL(gg, A(i2, d1));
      gg = MUL(gg, INV(A(STEP(STEP(i2, d1, 1), d, -1), d)));
      gg = MUL(gg, INV(A(STEP(i, d1, 1), d)));
      gg = MUL(gg, INV(A(i, d1)));
      re[k++] = gg;

      gg = id;
      gg = MUL(gg, A(STEP(i, d, 1), d1));
      i2 = STEP(STEP(i, d, 1), d1, 1);
      gg = MUL(gg, INV(A(STEP(i2, d, -1), d)));
      gg = MUL(gg, INV(A(STEP(STEP(i2, d, -1), d, -1), d)));
      gg = MUL(gg, INV(A(STEP(i, d, -1), d1)));
      gg = MUL(gg, A(STEP(i, d, -1), d));
      re[k++] = gg;

      gg = id;
      gg = MUL(gg, INV(A(STEP(STEP(i, d, 1), d1, -1), d1)));
      i2 = STEP(STEP(i, d, 1), d1, -1);
      gg = MUL(gg, INV(A(STEP(i2, d1, -1), d1)));
      gg = MUL(gg, INV(A(STEP(STEP(i2, d1, -1), d, -1), d)));
      gg = MUL(gg, A(STEP(STEP(i, d1, -1), d1, -1), d1));
      gg = MUL(gg, A(STEP(i, d1, -1), d1));
      re[k++] = gg;

      gg = id;
      gg = MUL(gg, A(STEP(i, d, 1), d));
      i2 = STEP(STEP(i, d, 1), d, 1);
      gg = MUL(gg, INV(A(STEP(i2, d1, -1), d1)));
      gg = MUL(gg, INV(A(STEP(STEP(i2, d1, -1), d, -1), d)));
      gg = MUL(gg, INV(A(STEP(i, d1, -1), d)));
      gg = MUL(gg, A(STEP(i, d1, -1), d1));
      re[k++] = gg;

      gg = id;
      gg = MUL(gg, INV(A(STEP(STEP(i, d, 1), d1, -1), d1)));
      i2 = STEP(STEP(i, d, 1), d1, -1);
      gg = MUL(gg, INV(A(STEP(i2, d, -1), d)));
      gg = MUL(gg, INV(A(STEP(STEP(i2, d, -1), d, -1), d)));
      gg = MUL(gg, A(STEP(STEP(i, d, -1), d1, -1), d1));
      gg = MUL(gg, A(STEP(i, d, -1), d));
      re[k++] = gg;
    }
  }
#undef STEP
#undef A
  (void)i1;
}

void og_staples(const og_lat *L, unsigned i, unsigned d, int32_t *st) {
  staples_impl(L, i, d, st, NULL, NULL, NULL);
}
void og_rect(const og_lat *L, unsigned i, unsigned d, int32_t *re) {
  staples_impl(L, i, d, NULL, re, NULL, NULL);
}
unsigned og_readset(const og_lat *L, unsigned i, unsigned d, uint32_t *links) {
  int32_t st[2 * OG_DMAX], re[6 * OG_DMAX];
  unsigned n = 0;
  staples_impl(L, i, d, st, re, links, &n);
  return n;
}

/* lattice.c:76-83 */
int32_t og_plaquette(const og_lat *L, unsigned i, unsigned d1, unsigned d2) {
  const og_group *g = L->g;
  const unsigned P = g->P, D = L->D;
  int32_t r = g->id;
  r = MUL(r, L->a[(size_t)i * D + d1]);
  r = MUL(r, L->a[(size_t)og_step(L, i, d1, 1) * D + d2]);
  r = MUL(r, INV(L->a[(size_t)og_step(L, i, d2, 1) * D + d1]));
  r = MUL(r, INV(L->a[(size_t)i * D + d2]));
  return r;
}

/* lattice.c:86-94 */
int32_t og_polyakov(const og_lat *L, unsigned i) {
  const og_group *g = L->g;
  const unsigned P = g->P, D = L->D;
  int32_t r = g->id;
  unsigned i0 = i;
  do {
    r = MUL(r, L->a[(size_t)i * D + 0]);
    i = og_step(L, i, 0, 1);
  } while (i != i0);
  return r;
}

/* lattice.c:116-124 */
int32_t og_wilson(const og_lat *L, unsigned i, unsigned d1, unsigned d2, unsigned n1, unsigned n2) {
  const og_group *g = L->g;
  const unsigned P = g->P, D = L->D;
  int32_t r = g->id;
  for (unsigned j = 0; j < n1; ++j) { r = MUL(r, L->a[(size_t)i * D + d1]); i = og_step(L, i, d1, 1); }
  for (unsigned j = 0; j < n2; ++j) { r = MUL(r, L->a[(size_t)i * D + d2]); i = og_step(L, i, d2, 1); }
  for (unsigned j = 0; j < n1; ++j) { i = og_step(L, i, d1, -1); r = MUL(r, INV(L->a[(size_t)i * D + d1])); }
  for (unsigned j = 0; j < n2; ++j) { i = og_step(L, i, d2, -1); r = MUL(r, INV(L->a[(size_t)i * D + d2])); }
  return r;
}

/* ------------------------------------------------------------ random streams */

/* std::minstd_rand0 = linear_congruential_engine<uint_fast32_t,16807,0,2^31-1>;
 * seed(s): x = s mod m, 1 if that is 0 (libstdc++ 11 random.h).  The
 * reference seeds site i with the int iseed+i (dym-mod-metro.cpp:228), which
 * the 64-bit result_type receives modulo 2^64. */
uint64_t og_minstd_seed(long long s) {
  uint64_t u = (uint64_t)s;
  uint64_t x = u % 2147483647ull;
  return x == 0 ? 1 : x;
}

uint64_t og_minstd_next(uint64_t *x) {
  *x = (*x * 16807ull) % 2147483647ull;
  return *x;
}

/* generate_canonical<double,53>(minstd_rand0), libstdc++ 11 random.tcc:3345-3379:
 * R = max-min+1 = 2147483646, log2r = 30, m = 2 draws. */
double og_canonical(uint64_t *x) {
  const long double R = 2147483646.0L;
  double sum = 0.0, tmp = 1.0;
  for (int k = 2; k != 0; --k) {
    sum += (double)(og_minstd_next(x) - 1) * tmp;
    tmp = (double)((long double)tmp * R);
  }
  double ret = sum / tmp;
  if (ret >= 1.0) ret = nextafter(1.0, 0.0);
  return ret;
}

/* uniform_int_distribution<int>(lo,hi)(minstd_rand0): urngrange = 2^31-3 is
 * neither 2^32-1 nor 2^64-1, so the two-division downscaling path of
 * uniform_int_dist.h:296-322 (libstdc++ 11) is taken. */
int og_randint(uint64_t *x, int lo, int hi) {
  const uint64_t urngrange = 2147483646ull - 1ull;
  const uint64_t urange = (uint64_t)((unsigned)hi - (unsigned)lo);
  const uint64_t uerange = urange + 1;
  const uint64_t scaling = urngrange / uerange;
  const uint64_t past = uerange * scaling;
  uint64_t ret;
  do ret = og_minstd_next(x) - 1;
  while (ret >= past);
  ret /= scaling;
  return (int)(ret + (unsigned)lo);
}

/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11: multipliers 0xD2511F53, 0xCD9E8D57; key Weyl increments
 * 0x9E3779B9, 0xBB67AE85; ten rounds. */
void og_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in [0,1) from two 32-bit words. */
double og_u53(uint32_t hi, uint32_t lo) {
  return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6)) * 0x1.0p-53;
}

/* exp(x): Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, ln2 split into a
 * 32-bit head (exact products k*hi for |k| < 2^21) and a tail; e^r from its
 * Taylor series through r^13 (truncation < 5e-18 relative), Horner order;
 * scaling by 2^k through the exponent field.  Basic IEEE operations only, so
 * any compiler without FMA contraction gives the same bits. */
double og_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return INFINITY;
  if (x < -745.1332191019412) return 0.0;
  const double log2e = 0x1.71547652b82fep+0;
  const double ln2hi = 0x1.62e42fee00000p-1;
  const double ln2lo = 0x1.a39ef35793c76p-33;
  double t = x * log2e;
  int k = (int)(t < 0.0 ? t - 0.5 : t + 0.5);
  double kd = (double)k;
  double r = (x - kd * ln2hi) - kd * ln2lo;
  double p = 0x1.6124613a86d09p-33;
  p = p * r + 0x1.1eed8eff8d898p-29;
  p = p * r + 0x1.ae64567f544e4p-26;
  p = p * r + 0x1.27e4fb7789f5cp-22;
  p = p * r + 0x1.71de3a556c734p-19;
  p = p * r + 0x1.a01a01a01a01ap-16;
  p = p * r + 0x1.a01a01a01a01ap-13;
  p = p * r + 0x1.6c16c16c16c17p-10;
  p = p * r + 0x1.1111111111111p-7;
  p = p * r + 0x1.5555555555555p-5;
  p = p * r + 0x1.5555555555555p-3;
  p = p * r + 0x1.0000000000000p-1;
  double er = 1.0 + (r + (r * r) * p);
  /* er * 2^k; k in [-1075, 1024] */
  union { double d; uint64_t u; } s1, s2;
  if (k > 1000) {
    s1.u = (uint64_t)(k - 1000 + 1023) << 52;
    s2.u = (uint64_t)(1000 + 1023) << 52;
    return (er * s1.d) * s2.d;
  }
  if (k < -1000) {
    s1.u = (uint64_t)(k + 1000 + 1023) << 52;
    s2.u = (uint64_t)(-1000 + 1023) << 52;
    return (er * s1.d) * s2.d;
  }
  s1.u = (uint64_t)(k + 1023) << 52;
  return er * s1.d;
}

/* ------------------------------------------------------------- colouring */

/* Colour of site x for direction mu: (x_mu mod 2) + 2*((sum_{nu!=mu} x_nu) mod m)
 * with m = 3 when every transverse extent is a multiple of 3 and the mu extent
 * is even, else m = 4 when every transverse extent is a multiple of 4, else a
 * greedy colouring of the conflict graph in ascending site order.  Two
 * same-direction links conflict when one is read by the other's staples. */
int og_colouring(const og_lat *L, uint8_t *col, unsigned *ncol) {
  const unsigned D = L->D, V = L->V;
  uint32_t rs[512];
  for (unsigned mu = 0; mu < D; ++mu) {
    int even = L->ext[mu] % 2 == 0, t3 = 1, t4 = 1;
    for (unsigned nu = 0; nu < D; ++nu) if (nu != mu) {
      if (L->ext[nu] % 3) t3 = 0;
      if (L->ext[nu] % 4) t4 = 0;
    }
    unsigned m = (even && t3) ? 3 : (even && t4) ? 4 : 0;
    uint8_t *c = col + (size_t)mu * V;
    if (m) {
      unsigned pos[OG_DMAX];
      for (unsigned i = 0; i < V; ++i) {
        og_getpos(L, i, pos);
        unsigned s = 0;
        for (unsigned nu = 0; nu < D; ++nu) if (nu != mu) s += pos[nu];
        c[i] = (uint8_t)((pos[mu] & 1) + 2 * (s % m));
      }
      ncol[mu] = 2 * m;
    } else {
      unsigned maxc = 0;
      for (unsigned i = 0; i < V; ++i) {
        uint64_t used = 0;
        unsigned n = og_readset(L, i, mu, rs);
        for (unsigned q = 0; q < n; ++q) {
          unsigned s = rs[q] / D, dir = rs[q] % D;
          if (dir == mu && s < i) used |= 1ull << c[s];
        }
        unsigned cc = 0;
        while (used >> cc & 1) ++cc;
        if (cc >= 64) return -1;
        c[i] = (uint8_t)cc;
        if (cc + 1 > maxc) maxc = cc + 1;
      }
      ncol[mu] = maxc;
    }
  }
  return 0;
}

unsigned long og_colour_conflicts(const og_lat *L, const uint8_t *col) {
  const unsigned D = L->D, V = L->V;
  uint32_t rs[512];
  unsigned long bad = 0;
  for (unsigned mu = 0; mu < D; ++mu)
    for (unsigned i = 0; i < V; ++i) {
      unsigned n = og_readset(L, i, mu, rs);
      for (unsigned q = 0; q < n; ++q) {
        unsigned s = rs[q] / D, dir = rs[q] % D;
        if (dir == mu && s != i && col[(size_t)mu * V + s] == col[(size_t)mu * V + i]) ++bad;
      }
    }
  return bad;
}

/* ----------------------------------------------------------------- chain */

/* Initialisation as dym-mod-metro.cpp:227-268: minstd engines seeded iseed+i;
 * hot links drawn in link order from site 0's engine (uniform_int 0..P-1);
 * philox: link l drawn as floor(P*u) from Philox(key=(seed,"DSAI"), ctr=(l,0,0,0)). */
int og_chain_init(og_chain *c, const og_group *g, unsigned D, unsigned Nt, unsigned Nx,
                  double beta0, double beta1, double beta2, int32_t seed, int cold,
                  int rng, int order) {
  memset(c, 0, sizeof *c);
  if (og_lat_init(&c->L, g, D, Nt, Nx)) return -1;
  c->beta0 = beta0; c->beta1 = beta1; c->beta2 = beta2;
  c->rng = rng; c->order = order; c->seed = seed;
  const unsigned V = c->L.V;
  if (rng == OG_RNG_MINSTD) {
    c->ms = (uint64_t *)malloc(sizeof(uint64_t) * V);
    for (unsigned i = 0; i < V; ++i) c->ms[i] = og_minstd_seed((long long)(int)(seed + (int)i));
  }
  for (size_t l = 0; l < (size_t)V * D; ++l) {
    if (cold) { c->L.a[l] = g->id; continue; }
    if (rng == OG_RNG_MINSTD) {
      c->L.a[l] = og_randint(&c->ms[0], 0, (int)g->P - 1);
    } else {
      uint32_t ctr[4] = {(uint32_t)l, 0, 0, 0}, key[2] = {(uint32_t)seed, OG_KEY_INIT}, o[4];
      og_philox4x32_10(ctr, key, o);
      c->L.a[l] = (int32_t)((double)g->P * og_u53(o[0], o[1]));
    }
  }
  if (order == OG_ORDER_COLOUR) {
    c->col = (uint8_t *)malloc((size_t)D * V);
    if (og_colouring(&c->L, c->col, c->ncol)) return -2;
  }
  return 0;
}

void og_chain_free(og_chain *c) {
  og_lat_free(&c->L);
  free(c->ms);
  free(c->col);
  memset(c, 0, sizeof *c);
}

/* dym-mod-metro.cpp:161-192 for link (i,d).  philox: hits h and h+1 (h even)
 * of sweep s share Philox(key=(seed,"DSA1"), ctr=(i*D+d, h/2, s_lo, s_hi));
 * hit h takes words 2(h&1), 2(h&1)+1: proposal pool[floor(npool w/2^32)] and
 * acceptance uniform w'/2^32. */
void og_update_link(og_chain *c, unsigned i, unsigned d) {
  const og_lat *L = &c->L;
  const og_group *g = L->g;
  const unsigned P = g->P, D = L->D;
  const int ns = 2 * (int)(D - 1), nr = 6 * (int)(D - 1);
  int32_t st[2 * OG_DMAX], re[6 * OG_DMAX];
  og_staples(L, i, d, st);
  og_rect(L, i, d, re);
  const size_t l = (size_t)i * D + d;
  int32_t b = L->a[l];
  double r1 = 0, m1 = 0;
  for (int j = 0; j < ns; ++j) r1 += g->retr[MUL(b, st[j])];
  for (int k = 0; k < nr; ++k) m1 += g->retr[MUL(b, re[k])];
  uint32_t o[4] = {0, 0, 0, 0};
  for (int h = 0; h < OG_NHIT; ++h) {
    double u2;
    size_t jp;
    if (c->rng == OG_RNG_PHILOX) {
      if ((h & 1) == 0) {
        uint32_t ctr[4] = {(uint32_t)l, (uint32_t)(h >> 1), (uint32_t)c->sweep, (uint32_t)(c->sweep >> 32)};
        uint32_t key[2] = {(uint32_t)c->seed, OG_KEY_SWEEP};
        og_philox4x32_10(ctr, key, o);
      }
      jp = (size_t)(((uint64_t)o[2 * (h & 1)] * g->npool) >> 32);
      u2 = (double)o[2 * (h & 1) + 1] * 0x1.0p-32;
    } else {
      jp = (size_t)((double)g->npool * og_canonical(&c->ms[i]));
      u2 = 0; /* drawn below, after the proposal's action (same stream order) */
    }
    int32_t bnew = MUL(b, g->pool[jp]);
    double r1n = 0, m1n = 0;
    for (int j = 0; j < ns; ++j) r1n += g->retr[MUL(bnew, st[j])];
    for (int k = 0; k < nr; ++k) m1n += g->retr[MUL(bnew, re[k])];
    double oldact = c->beta0 + c->beta1 * r1 + c->beta2 * m1;
    double newact = c->beta0 + c->beta1 * r1n + c->beta2 * m1n;
    double probrat = (newact <= oldact) ? 1.0 : og_exp(-(newact - oldact));
    if (c->rng == OG_RNG_MINSTD) u2 = og_canonical(&c->ms[i]);
    ++c->hit;
    if (probrat > u2) {
      c->L.a[l] = bnew;
      b = bnew; r1 = r1n; m1 = m1n;
      ++c->acc;
    }
  }
}

/* dym-mod-metro.cpp:146-197 at one thread (reference order), or by colours. */
void og_chain_sweep(og_chain *c, unsigned n) {
  const unsigned D = c->L.D, V = c->L.V;
  unsigned pos[OG_DMAX];
  for (unsigned it = 0; it < n; ++it) {
    for (unsigned d = 0; d < D; d++) {
      if (c->order == OG_ORDER_REFERENCE) {
        for (int parity = 0; parity < 2; ++parity)
          for (unsigned i = 0; i < V; i++) {
            og_getpos(&c->L, i, pos);
            unsigned s = 0;
            for (unsigned q = 0; q < D; ++q) s += pos[q];
            if ((int)(s & 1) != parity) continue;
            og_update_link(c, i, d);
          }
      } else {
        const uint8_t *col = c->col + (size_t)d * V;
        for (unsigned cc = 0; cc < c->ncol[d]; ++cc)
          for (unsigned i = 0; i < V; i++)
            if (col[i] == cc) og_update_link(c, i, d);
      }
    }
    c->sweep++;
  }
}

/* ----------------------------------------------------------- measurement */

static double hist_dot(const uint64_t *cnt, const double *tab, unsigned P) {
  double s = 0.0;
  for (unsigned e = 0; e < P; ++e)
    if (cnt[e]) s += (double)cnt[e] * tab[e];
  return s;
}

/* dym-mod-metro.cpp:282-318 (getpoly: lattice.c:98-113). */
void og_measure(const og_lat *L, unsigned wk, unsigned wl, int mode, double *out) {
  const og_group *g = L->g;
  const unsigned D = L->D, V = L->V, Nt = L->Nt, P = g->P;
  unsigned pos[OG_DMAX];
  if (mode == 0) {
    double lre = 0, lim = 0;
    for (unsigned i = 0; i < V; ++i) {
      og_getpos(L, i, pos);
      if (pos[0] != 0) continue;
      int32_t p = og_polyakov(L, i);
      lre += g->retr[p];
      lim += g->imtr[p];
    }
    out[0] = lre / (V / Nt);
    out[1] = lim / (V / Nt);
    double sp = 0;
    for (unsigned i = 0; i < V; ++i)
      for (unsigned d1 = 0; d1 < D; ++d1)
        for (unsigned d2 = d1 + 1; d2 < D; ++d2) sp += g->retr[og_plaquette(L, i, d1, d2)];
    out[2] = sp / (V * D * (D - 1) / 2);
    double sr = 0;
    for (unsigned i = 0; i < V; ++i)
      for (unsigned d1 = 0; d1 < D; ++d1)
        for (unsigned d2 = 0; d2 < D; ++d2)
          if (d1 != d2) sr += g->retr[og_wilson(L, i, d1, d2, 2, 1)];
    out[3] = sr / (V * D * (D - 1));
    for (unsigned k = 0; k < wk; ++k)
      for (unsigned l = 0; l < wl; ++l) {
        double lw = 0;
        for (unsigned i = 0; i < V; ++i)
          for (unsigned d1 = 1; d1 < D; ++d1) lw += g->retr[og_wilson(L, i, d1, 0, k + 1, l + 1)];
        out[4 + k * wl + l] = lw / (V * (D - 1));
      }
    return;
  }
  uint64_t *cnt = (uint64_t *)calloc(P, sizeof(uint64_t));
  for (unsigned i = 0; i < V; ++i) {
    og_getpos(L, i, pos);
    if (pos[0] == 0) cnt[og_polyakov(L, i)]++;
  }
  out[0] = hist_dot(cnt, g->retr, P) / (V / Nt);
  out[1] = hist_dot(cnt, g->imtr, P) / (V / Nt);
  memset(cnt, 0, sizeof(uint64_t) * P);
  for (unsigned i = 0; i < V; ++i)
    for (unsigned d1 = 0; d1 < D; ++d1)
      for (unsigned d2 = d1 + 1; d2 < D; ++d2) cnt[og_plaquette(L, i, d1, d2)]++;
  out[2] = hist_dot(cnt, g->retr, P) / (V * D * (D - 1) / 2);
  memset(cnt, 0, sizeof(uint64_t) * P);
  for (unsigned i = 0; i < V; ++i)
    for (unsigned d1 = 0; d1 < D; ++d1)
      for (unsigned d2 = 0; d2 < D; ++d2)
        if (d1 != d2) cnt[og_wilson(L, i, d1, d2, 2, 1)]++;
  out[3] = hist_dot(cnt, g->retr, P) / (V * D * (D - 1));
  for (unsigned k = 0; k < wk; ++k)
    for (unsigned l = 0; l < wl; ++l) {
      memset(cnt, 0, sizeof(uint64_t) * P);
      for (unsigned i = 0; i < V; ++i)
        for (unsigned d1 = 1; d1 < D; ++d1) cnt[og_wilson(L, i, d1, 0, k + 1, l + 1)]++;
      out[4 + k * wl + l] = hist_dot(cnt, g->retr, P) / (V * (D - 1));
    }
  free(cnt);
}

void og_measure_abs(const og_lat *L, unsigned wk, unsigned wl, double *out) {
  const og_group *g = L->g;
  const unsigned D = L->D, V = L->V, Nt = L->Nt;
  unsigned pos[OG_DMAX];
  double lre = 0, lim = 0;
  for (unsigned i = 0; i < V; ++i) {
    og_getpos(L, i, pos);
    if (pos[0] != 0) continue;
    int32_t p = og_polyakov(L, i);
    lre += fabs(g->retr[p]);
    lim += fabs(g->imtr[p]);
  }
  out[0] = lre / (V / Nt);
  out[1] = lim / (V / Nt);
  double sp = 0;
  for (unsigned i = 0; i < V; ++i)
    for (unsigned d1 = 0; d1 < D; ++d1)
      for (unsigned d2 = d1 + 1; d2 < D; ++d2) sp += fabs(g->retr[og_plaquette(L, i, d1, d2)]);
  out[2] = sp / (V * D * (D - 1) / 2);
  double sr = 0;
  for (unsigned i = 0; i < V; ++i)
    for (unsigned d1 = 0; d1 < D; ++d1)
      for (unsigned d2 = 0; d2 < D; ++d2)
        if (d1 != d2) sr += fabs(g->retr[og_wilson(L, i, d1, d2, 2, 1)]);
  out[3] = sr / (V * D * (D - 1));
  for (unsigned k = 0; k < wk; ++k)
    for (unsigned l = 0; l < wl; ++l) {
      double lw = 0;
      for (unsigned i = 0; i < V; ++i)
        for (unsigned d1 = 1; d1 < D; ++d1) lw += fabs(g->retr[og_wilson(L, i, d1, 0, k + 1, l + 1)]);
      out[4 + k * wl + l] = lw / (V * (D - 1));
    }
}

/* ------------------------------------------------------------------ LLR */

/* llr.cpp:46-55 in sequential order */
double og_S1(const og_lat *L) {
  double s = 0;
  for (unsigned i = 0; i < L->V; ++i)
    for (unsigned d1 = 0; d1 < L->D; ++d1)
      for (unsigned d2 = d1 + 1; d2 < L->D; ++d2) s += L->g->retr[og_plaquette(L, i, d1, d2)];
  return s;
}

/* llr.cpp:139-149 */
double og_sum_m1(const og_lat *L) {
  const og_group *g = L->g;
  const unsigned P = g->P, D = L->D;
  const int nr = 6 * (int)(D - 1);
  int32_t re[6 * OG_DMAX];
  double s = 0;
  for (unsigned i = 0; i < L->V; ++i)
    for (unsigned d = 0; d < D; ++d) {
      og_rect(L, i, d, re);
      int32_t b = L->a[(size_t)i * D + d];
      for (int k = 0; k < nr; ++k) s += g->retr[MUL(b, re[k])];
    }
  return s;
}
