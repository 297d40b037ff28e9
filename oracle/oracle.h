/*
 * oracle.h -- CPU restatement of the reference's Metropolis hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * check and the CPU-baseline leg of the benchmark may load it; the product
 * (dym_symanzik_action_amd/) never does.  It restates, in plain C:
 *   group.c:13-81                 table loader (identity/inverse rule)
 *   dym-mod-metro.cpp:239-245     proposal pool
 *   lattice.c:22-124              geometry and closed loops
 *   dym-mod-metro.cpp:63-142      plaquette and rectangle staples
 *   dym-mod-metro.cpp:146-197     the sweep (hit loop, acceptance)
 *   dym-mod-metro.cpp:281-318     the measurement block
 *   llr.cpp:46-250, llr2d.cpp:178-259   LLR constrained update + Robbins-Monro
 * plus the two random streams the path can run on:
 *   minstd-compat: std::minstd_rand0 per site with libstdc++ 11
 *     generate_canonical<double,53> and uniform_int_distribution (downscaling),
 *     as the reference consumes them (dym-mod-metro.cpp:148,174,187,227-232,255);
 *   philox: Philox4x32-10 (Salmon et al., SC'11), one call per two hits.
 * and two visiting orders:
 *   reference: direction -> site parity -> ascending site (dym-mod-metro.cpp:152-159)
 *   colour:    direction -> colour -> site, with a colouring in which no two
 *              links of one pass share a plaquette or a rectangle.
 */
#ifndef DSA_ORACLE_H
#define DSA_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OG_NHIT 20
#define OG_DMAX 8

typedef struct {
  unsigned P;
  double *retr, *imtr;
  int32_t *mult; /* P*P row-major: mult[n*P + m] = n*m */
  int32_t *inv;
  int32_t id;
  int32_t *pool;
  unsigned npool;
  double min_retr, nn_retr;
} og_group;

/* 0 on success, <0 on a malformed table (the reference aborts). */
int og_group_load(const char *path, og_group *g);
void og_group_free(og_group *g);

typedef struct {
  unsigned D, Nt, Nx, V;
  unsigned ext[OG_DMAX];
  unsigned stride[OG_DMAX];
  const og_group *g;
  int32_t *a; /* V*D, a[i*D + mu] (reference order) */
} og_lat;

int og_lat_init(og_lat *L, const og_group *g, unsigned D, unsigned Nt, unsigned Nx);
void og_lat_free(og_lat *L);
unsigned og_step(const og_lat *L, unsigned i, unsigned d, int s);
void og_getpos(const og_lat *L, unsigned i, unsigned *pos);

void og_staples(const og_lat *L, unsigned i, unsigned d, int32_t *st);
void og_rect(const og_lat *L, unsigned i, unsigned d, int32_t *st);
/* the link positions (site*D+dir) the staples of (i,d) read, in read order;
 * returns the count (<= 6*(D-1)*5 + 2*(D-1)*3) */
unsigned og_readset(const og_lat *L, unsigned i, unsigned d, uint32_t *links);

int32_t og_plaquette(const og_lat *L, unsigned i, unsigned d1, unsigned d2);
int32_t og_polyakov(const og_lat *L, unsigned i);
int32_t og_wilson(const og_lat *L, unsigned i, unsigned d1, unsigned d2, unsigned n1, unsigned n2);

/* ---- random streams ---- */
uint64_t og_minstd_seed(long long s);
uint64_t og_minstd_next(uint64_t *x);
double og_canonical(uint64_t *x);
int og_randint(uint64_t *x, int lo, int hi);
void og_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double og_u53(uint32_t hi, uint32_t lo);
double og_exp(double x);

enum { OG_RNG_MINSTD = 0, OG_RNG_PHILOX = 1 };
enum { OG_ORDER_REFERENCE = 0, OG_ORDER_COLOUR = 1 };

#define OG_KEY_SWEEP 0x44534131u /* "DSA1" */
#define OG_KEY_INIT 0x44534149u  /* "DSAI" */
#define OG_KEY_LLR 0x4453414cu   /* "DSAL" */

/* colour of every site for direction mu: col[mu*V + i]; ncol[mu] colours */
int og_colouring(const og_lat *L, uint8_t *col, unsigned *ncol);
/* number of same-direction conflicts of a colouring (0 = race-free) */
unsigned long og_colour_conflicts(const og_lat *L, const uint8_t *col);

typedef struct {
  og_lat L;
  double beta0, beta1, beta2;
  int rng, order;
  int32_t seed;
  uint64_t *ms;     /* minstd: per-site state */
  uint8_t *col;     /* colour order: D*V colours */
  unsigned ncol[OG_DMAX];
  uint64_t sweep;   /* sweeps done */
  uint64_t hit, acc;
} og_chain;

int og_chain_init(og_chain *c, const og_group *g, unsigned D, unsigned Nt, unsigned Nx,
                  double beta0, double beta1, double beta2, int32_t seed, int cold,
                  int rng, int order);
void og_chain_free(og_chain *c);
void og_chain_sweep(og_chain *c, unsigned n);
/* one link's update (the body of dym-mod-metro.cpp:158-192) */
void og_update_link(og_chain *c, unsigned i, unsigned d);

/* measurement: out = [rep, imp, plaq, rect, W(0,0) .. W(wk-1, wl-1)] (k outer).
 * mode 0: the reference's one-thread summation order;
 * mode 1: per-element integer histograms, then sum_e count[e]*ReTr[e] in
 *         ascending element order. */
void og_measure(const og_lat *L, unsigned wk, unsigned wl, int mode, double *out);
/* sum of |terms| for each measured quantity (tolerance scale), same layout */
void og_measure_abs(const og_lat *L, unsigned wk, unsigned wl, double *out);

/* ---- LLR (llr.cpp / llr2d.cpp), serial reference order ---- */
double og_S1(const og_lat *L);            /* llr.cpp:46-55 (sequential order) */
double og_sum_m1(const og_lat *L);        /* llr.cpp:139-149 */

#ifdef __cplusplus
}
#endif
#endif
