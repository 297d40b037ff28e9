/*
 * oracle_llr.h -- CPU restatement of the LLR drivers (TEST INFRASTRUCTURE ONLY).
 *
 * ogl_ref_*: llr.cpp:46-345 as the reference runs it at one thread (serial
 *            visiting order, per-hit window test on the running energies,
 *            minstd streams).
 * ogl_dev_*: the device LLR algorithm of this build (per-pass window test in
 *            fixed point, Philox streams), restated for bit-exact checks.
 *            It is not the reference's algorithm, so it is pinned to the
 *            reference statistically (a_n of the reference llr and both tilts
 *            of the reference llr2d on their own fixtures, tests/test_oracle.py)
 *            and by invariants (incremental energies equal a recomputation,
 *            the window holds once reached).
 */
#ifndef DSA_ORACLE_LLR_H
#define DSA_ORACLE_LLR_H
#include "oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  og_chain c;
  int axis;
  double bg;
  double S1, S2;
  double rect_mult;
  double maxdrift;
} ogl_ref;

int ogl_ref_init(ogl_ref *R, const og_group *g, unsigned D, unsigned Nt, unsigned Nx, int axis, double bg, int seed);
void ogl_ref_calibrate(ogl_ref *R);
void ogl_ref_update_constrained(ogl_ref *R, double a_tilt, double Elo, double Ehi, int driveIn);
int ogl_ref_seed_window(ogl_ref *R, double Elo, double Ehi, double a_guess);
double ogl_ref_rm_solve(ogl_ref *R, double Ewin, double hw, double a0, unsigned K, unsigned NRM);
int ogl_ref_main(int argc, char **argv);

/* ---- the device algorithm (per-pass window, fixed point, Philox "DSAL") ---- */
enum { OGL_SEED = 0, OGL_RM = 1, OGL_DONE = 2 };

typedef struct {
  og_chain c;        /* lattice, colouring; Philox key = c.seed */
  int axis;          /* 1, 2: one window; 3: box on (S1, S2) (llr2d) */
  double bg, Ewin, W;
  long long lo, hi;
  double Ewin2, W2;  /* axis 3: the S2 window */
  long long lo2, hi2;
  unsigned K, NRM, cap;
  double tol;
  long long S1, S2;  /* fixed point, units 2^-32 */
  double a, sumE, last_da, maxdrift;
  double drift1, drift2; /* the drift of S1 and of S2 separately */
  double a2, sumE2;  /* axis 3: tilt of S2 and its sum */
  unsigned phase, k, m, seed_sweeps, reached, resync;
  uint64_t sweep, acc, hit;
  int32_t *old;      /* scratch of one pass: previous values, */
  long long *d1, *d2; /* proposed energy changes */
  unsigned *na;      /* and accepted hits, in ascending site order */
} ogl_dev;

int ogl_dev_init(ogl_dev *L, const og_group *g, unsigned D, unsigned Nt, unsigned Nx, int axis, double bg,
                 double Ewin, double hw, double a0, int seed, int cold, unsigned K, unsigned NRM, double tol,
                 unsigned cap, double Ewin2, double hw2, double a20);
void ogl_dev_free(ogl_dev *L);
/* n sweeps, with the resync bookkeeping the device does around them */
void ogl_dev_step(ogl_dev *L, unsigned n);
long long ogl_fix(double x);

#ifdef __cplusplus
}
#endif
#endif
