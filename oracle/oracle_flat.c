/*
 * oracle_flat.c -- handle-based entry points of the oracle for ctypes
 * (TEST INFRASTRUCTURE ONLY; used by oracle/pyoracle.py).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

void *ogp_group_load(const char *path) {
  og_group *g = (og_group *)calloc(1, sizeof *g);
  if (og_group_load(path, g)) { free(g); return NULL; }
  return g;
}
void ogp_group_free(void *g) { if (g) { og_group_free((og_group *)g); free(g); } }
unsigned ogp_group_P(void *g) { return ((og_group *)g)->P; }
int ogp_group_id(void *g) { return ((og_group *)g)->id; }
unsigned ogp_group_npool(void *g) { return ((og_group *)g)->npool; }
void ogp_group_tables(void *gv, int32_t *mult, int32_t *inv, double *retr, double *imtr, int32_t *pool) {
  og_group *g = (og_group *)gv;
  size_t P = g->P;
  if (mult) memcpy(mult, g->mult, P * P * 4);
  if (inv) memcpy(inv, g->inv, P * 4);
  if (retr) memcpy(retr, g->retr, P * 8);
  if (imtr) memcpy(imtr, g->imtr, P * 8);
  if (pool) memcpy(pool, g->pool, g->npool * 4);
}

void *ogp_chain_new(void *g, unsigned D, unsigned Nt, unsigned Nx, double b0, double b1, double b2, int seed,
                    int cold, int rng, int order) {
  og_chain *c = (og_chain *)calloc(1, sizeof *c);
  if (og_chain_init(c, (og_group *)g, D, Nt, Nx, b0, b1, b2, seed, cold, rng, order)) {
    og_chain_free(c);
    free(c);
    return NULL;
  }
  return c;
}
void ogp_chain_free(void *c) { if (c) { og_chain_free((og_chain *)c); free(c); } }
void ogp_chain_sweep(void *c, unsigned n) { og_chain_sweep((og_chain *)c, n); }
void ogp_chain_get_links(void *cv, int32_t *out) {
  og_chain *c = (og_chain *)cv;
  memcpy(out, c->L.a, (size_t)c->L.V * c->L.D * 4);
}
void ogp_chain_set_links(void *cv, const int32_t *in) {
  og_chain *c = (og_chain *)cv;
  memcpy(c->L.a, in, (size_t)c->L.V * c->L.D * 4);
}
void ogp_chain_counters(void *cv, uint64_t *hit, uint64_t *acc) {
  og_chain *c = (og_chain *)cv;
  *hit = c->hit;
  *acc = c->acc;
}
uint64_t ogp_chain_sweeps(void *cv) { return ((og_chain *)cv)->sweep; }
void ogp_chain_set_sweeps(void *cv, uint64_t s) { ((og_chain *)cv)->sweep = s; }
void ogp_chain_get_mstate(void *cv, uint32_t *out) {
  og_chain *c = (og_chain *)cv;
  for (unsigned i = 0; i < c->L.V; ++i) out[i] = (uint32_t)c->ms[i];
}
void ogp_chain_measure(void *cv, unsigned wk, unsigned wl, int mode, double *out) {
  og_measure(&((og_chain *)cv)->L, wk, wl, mode, out);
}
void ogp_chain_measure_abs(void *cv, unsigned wk, unsigned wl, double *out) {
  og_measure_abs(&((og_chain *)cv)->L, wk, wl, out);
}
void ogp_chain_staples(void *cv, int32_t *st, int32_t *re) {
  og_chain *c = (og_chain *)cv;
  const unsigned D = c->L.D, ns = 2 * (D - 1), nr = 6 * (D - 1);
  for (unsigned i = 0; i < c->L.V; ++i)
    for (unsigned d = 0; d < D; ++d) {
      size_t l = (size_t)i * D + d;
      og_staples(&c->L, i, d, st + l * ns);
      og_rect(&c->L, i, d, re + l * nr);
    }
}
/* kind: 0 plaq, 1 rect, 2 poly, 3 wilson(k+1,l+1) -- the harness's enumeration */
void ogp_chain_loops(void *cv, int kind, unsigned k, unsigned l, int32_t *out) {
  og_chain *c = (og_chain *)cv;
  const og_lat *L = &c->L;
  const unsigned D = L->D, V = L->V;
  size_t n = 0;
  if (kind == 0) {
    for (unsigned i = 0; i < V; ++i)
      for (unsigned d1 = 0; d1 < D; ++d1)
        for (unsigned d2 = d1 + 1; d2 < D; ++d2) out[n++] = og_plaquette(L, i, d1, d2);
  } else if (kind == 1) {
    for (unsigned i = 0; i < V; ++i)
      for (unsigned d1 = 0; d1 < D; ++d1)
        for (unsigned d2 = 0; d2 < D; ++d2)
          if (d1 != d2) out[n++] = og_wilson(L, i, d1, d2, 2, 1);
  } else if (kind == 2) {
    for (unsigned i = 0; i < V; i += L->Nt) out[n++] = og_polyakov(L, i);
  } else {
    for (unsigned i = 0; i < V; ++i)
      for (unsigned d1 = 1; d1 < D; ++d1) out[n++] = og_wilson(L, i, d1, 0, k + 1, l + 1);
  }
}
/* colouring of the chain's lattice: col[mu*V+i], ncol[mu]; returns conflicts */
long ogp_chain_colouring(void *cv, uint8_t *col, unsigned *ncol) {
  og_chain *c = (og_chain *)cv;
  if (og_colouring(&c->L, col, ncol)) return -1;
  return (long)og_colour_conflicts(&c->L, col);
}
unsigned ogp_chain_readset(void *cv, unsigned i, unsigned d, uint32_t *out) {
  return og_readset(&((og_chain *)cv)->L, i, d, out);
}
double ogp_chain_S1(void *cv) { return og_S1(&((og_chain *)cv)->L); }
double ogp_chain_sum_m1(void *cv) { return og_sum_m1(&((og_chain *)cv)->L); }

void ogp_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out) { og_philox4x32_10(ctr, key, out); }
double ogp_exp(double x) { return og_exp(x); }
void ogp_exp_many(const double *x, double *y, size_t n) {
  for (size_t i = 0; i < n; ++i) y[i] = og_exp(x[i]);
}
uint64_t ogp_minstd_seed(long long s) { return og_minstd_seed(s); }
double ogp_canonical(uint64_t *x) { return og_canonical(x); }
int ogp_randint(uint64_t *x, int lo, int hi) { return og_randint(x, lo, hi); }

/* ---- LLR ---- */
#include "oracle_llr.h"
void *ogp_llr_new(void *g, unsigned D, unsigned Nt, unsigned Nx, int axis, double bg, double Ewin, double hw,
                  double a0, int seed, int cold, unsigned K, unsigned NRM, double tol, unsigned cap, double Ewin2,
                  double hw2, double a20) {
  ogl_dev *L = (ogl_dev *)calloc(1, sizeof *L);
  if (ogl_dev_init(L, (og_group *)g, D, Nt, Nx, axis, bg, Ewin, hw, a0, seed, cold, K, NRM, tol, cap, Ewin2, hw2,
                   a20)) {
    free(L);
    return NULL;
  }
  return L;
}
void ogp_llr_free(void *L) { if (L) { ogl_dev_free((ogl_dev *)L); free(L); } }
void ogp_llr_step(void *L, unsigned n) { ogl_dev_step((ogl_dev *)L, n); }
/* start from a given configuration (energies recomputed, not compared) */
void ogp_llr_set_links(void *Lv, const int32_t *in) {
  ogl_dev *L = (ogl_dev *)Lv;
  memcpy(L->c.L.a, in, (size_t)L->c.L.V * L->c.L.D * 4);
  L->resync = 2;
}
/* out: a, S1, S2, last_da, maxdrift, phase, m, seed_sweeps, reached, sweep, hit, acc, a2, drift1, drift2 */
void ogp_llr_status(void *Lv, double *out) {
  ogl_dev *L = (ogl_dev *)Lv;
  out[0] = L->a; out[1] = (double)L->S1 / 4294967296.0; out[2] = (double)L->S2 / 4294967296.0;
  out[3] = L->last_da; out[4] = L->maxdrift; out[5] = L->phase; out[6] = L->m; out[7] = L->seed_sweeps;
  out[8] = L->reached; out[9] = (double)L->sweep; out[10] = (double)L->hit; out[11] = (double)L->acc;
  out[12] = L->a2;
  out[13] = L->drift1;
  out[14] = L->drift2;
}
void ogp_llr_links(void *Lv, int32_t *out) {
  ogl_dev *L = (ogl_dev *)Lv;
  memcpy(out, L->c.L.a, (size_t)L->c.L.V * L->c.L.D * 4);
}

/* ---- the reference llr's serial chain (ogl_ref_*), step by step ---- */
void *ogp_lref_new(void *g, unsigned D, unsigned Nt, unsigned Nx, int axis, double bg, int seed) {
  ogl_ref *R = (ogl_ref *)calloc(1, sizeof *R);
  if (ogl_ref_init(R, (og_group *)g, D, Nt, Nx, axis, bg, seed)) {
    free(R);
    return NULL;
  }
  return R;
}
void ogp_lref_free(void *Rv) {
  if (Rv) {
    og_chain_free(&((ogl_ref *)Rv)->c);
    free(Rv);
  }
}
double ogp_lref_rect_mult(void *Rv) { return ((ogl_ref *)Rv)->rect_mult; }
/* llr.cpp:326: rnd[i].seed(iseed + r*1000003 + i) */
void ogp_lref_reseed(void *Rv, unsigned r) {
  ogl_ref *R = (ogl_ref *)Rv;
  for (unsigned i = 0; i < R->c.L.V; ++i)
    R->c.ms[i] = og_minstd_seed((long long)(uint32_t)((uint32_t)R->c.seed + r * 1000003u + i));
}
/* run_pass's start (llr.cpp:258-260): hot links from site 0's stream or cold */
void ogp_lref_start(void *Rv, int hot) {
  ogl_ref *R = (ogl_ref *)Rv;
  og_lat *L = &R->c.L;
  const size_t nl = (size_t)L->V * L->D;
  for (size_t i = 0; i < nl; ++i) L->a[i] = hot ? og_randint(&R->c.ms[0], 0, (int)L->g->P - 1) : L->g->id;
  R->S1 = og_S1(L);
  R->S2 = og_sum_m1(L) / R->rect_mult;
}
int ogp_lref_seed_window(void *Rv, double Elo, double Ehi, double a_guess) {
  return ogl_ref_seed_window((ogl_ref *)Rv, Elo, Ehi, a_guess);
}
double ogp_lref_rm_solve(void *Rv, double Ewin, double hw, double a0, unsigned K, unsigned NRM) {
  return ogl_ref_rm_solve((ogl_ref *)Rv, Ewin, hw, a0, K, NRM);
}
/* out: S1, S2, maxdrift, hit, acc */
void ogp_lref_state(void *Rv, double *out) {
  ogl_ref *R = (ogl_ref *)Rv;
  out[0] = R->S1; out[1] = R->S2; out[2] = R->maxdrift; out[3] = (double)R->c.hit; out[4] = (double)R->c.acc;
}
void ogp_lref_links(void *Rv, int32_t *out) {
  ogl_ref *R = (ogl_ref *)Rv;
  memcpy(out, R->c.L.a, (size_t)R->c.L.V * R->c.L.D * 4);
}
/* one update_constrained sweep (llr.cpp:178-219) at tilt a */
void ogp_lref_sweep(void *Rv, double a, double Elo, double Ehi, int drive) {
  ogl_ref_update_constrained((ogl_ref *)Rv, a, Elo, Ehi, drive);
}
