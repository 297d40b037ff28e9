/*
 * dsa.h -- C ABI of libdsa.so, the MI355X (gfx950) discrete-group Symanzik
 * Monte Carlo.
 *
 * The reference exposes its path as global mutable state behind extern "C"
 * headers (group.h:8-16, lattice.h:7-16) driven by dym-mod-metro's main; this
 * ABI replaces that with opaque handles.  Every call returns 0 on success and
 * a negative DSA_E* code on failure (never aborts); dsa_last_error() gives the
 * calling thread's last message.  The caller owns handles and host buffers;
 * the library owns device memory.  Calls on one handle must be serialised by
 * the caller; distinct handles may be used from different threads.  A group
 * must outlive every chain created from it.
 *
 * Link arrays crossing this ABI are int32 in the reference order a[i*D + mu]
 * with site i = x0 + Nt*(x1 + Nx*(x2 + ...)) (lattice.c:39-51, x0 = time).
 */
#ifndef DSA_H
#define DSA_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DSA_OK = 0,
  DSA_E_ARG = -1,     /* invalid argument */
  DSA_E_IO = -2,      /* cannot read a file */
  DSA_E_FORMAT = -3,  /* malformed group table */
  DSA_E_DEVICE = -4,  /* HIP runtime error or no device */
  DSA_E_RANGE = -5,   /* size outside what the kernels support */
  DSA_E_STATE = -6    /* call not valid in the handle's state */
};

enum { DSA_RNG_PHILOX = 0, DSA_RNG_MINSTD = 1 };
enum { DSA_ORDER_COLOUR = 0, DSA_ORDER_REFERENCE = 1 };
enum { DSA_START_HOT = 0, DSA_START_COLD = 1 };
enum { DSA_SUM_TREE = 0, DSA_SUM_ORDERED = 1 };
enum { DSA_LOOP_PLAQ = 0, DSA_LOOP_RECT = 1, DSA_LOOP_POLY = 2, DSA_LOOP_WILSON = 3 };

typedef struct dsa_group dsa_group;
typedef struct dsa_chain dsa_chain;

typedef struct {
  uint32_t P;         /* group order (group.h:9) */
  int32_t id;         /* identity element (group.h:15) */
  uint32_t npool;     /* proposal pool size, "small group size" (dym-mod-metro.cpp:245) */
  double min_retr;    /* dym-mod-metro.cpp:239-241 */
  double nn_retr;     /* dym-mod-metro.cpp:240-242 */
} dsa_group_info;

typedef struct {
  uint32_t D, Nt, Nx;          /* lattice.h:7; V = Nt*Nx^(D-1) */
  double beta0, beta1, beta2;  /* lattice.h:8 */
  int32_t seed;                /* argv[8] of dym-mod-metro */
  int32_t start;               /* DSA_START_HOT | DSA_START_COLD (dym-mod-metro.cpp:250-255) */
  int32_t rng;                 /* DSA_RNG_PHILOX (default) | DSA_RNG_MINSTD (the reference's streams) */
  int32_t order;               /* DSA_ORDER_COLOUR (default) | DSA_ORDER_REFERENCE (one-thread order) */
} dsa_chain_params;

/* Replaces load_group (group.c:13-81) and the proposal pool
 * (dym-mod-metro.cpp:237-245).  Reads text or gzip. */
int dsa_group_load(const char *path, dsa_group **out);
/* Replaces reads of the globals P, id (group.h:9,15) and the pool statistics. */
int dsa_group_info_get(const dsa_group *g, dsa_group_info *info);
/* Copies the tables behind the globals ReTr, ImTr, mult, inv (group.h:10-14);
 * any pointer may be NULL.  mult is P*P int32 row-major, pool is npool int32. */
int dsa_group_tables(const dsa_group *g, int32_t *mult, int32_t *inv, double *retr, double *imtr, int32_t *pool);
/* The defining-representation (SU(3)) matrices appended to a group file after
 * its table, as dym-mod-metro-savecfg.cpp:249-262 reads them: out[18*e + part
 * + 2*(c + 3*r)], P*18 doubles.  DSA_E_STATE if the file has none. */
int dsa_group_matrices(const dsa_group *g, double *out);
/* No counterpart in the reference (the sweep kernels' product form): the
 * group's representation by 3x3 matrices over GF(4) = {0, 1, w, w^2}, found
 * at load time and checked against every product of the table.  mat[e] holds
 * element e's matrix with entry (r, c) in bits 2(3r + c) (2 e1 + e0 for
 * e1 w + e0), P uint32; either pointer may be NULL; key (4096 uint16) maps
 * the first two rows of a matrix (mat & 0xFFF) to its element, 0xFFFF where
 * none.  DSA_E_STATE if the group has no such form. */
int dsa_group_field(const dsa_group *g, uint32_t *mat, uint16_t *key);
void dsa_group_free(dsa_group *g);

/* Replaces main's set-up (dym-mod-metro.cpp:222-268): volume, per-site
 * streams, hot or cold start, neighbour geometry, visiting order. */
int dsa_chain_create(const dsa_group *g, const dsa_chain_params *p, dsa_chain **out);
void dsa_chain_destroy(dsa_chain *c);
/* n calls of update() (dym-mod-metro.cpp:146-197). */
int dsa_chain_sweep(dsa_chain *c, uint32_t n);
/* Cumulative Metropolis counters hit, acc (dym-mod-metro.cpp:60,194-195). */
int dsa_chain_counters(dsa_chain *c, uint64_t *hit, uint64_t *acc);
/* Sweeps done so far (the Philox counter); with set, resume from a checkpoint. */
int dsa_chain_get_sweep(dsa_chain *c, uint64_t *sweep);
int dsa_chain_set_sweep(dsa_chain *c, uint64_t sweep);
/* The measurement block (dym-mod-metro.cpp:282-318; getpoly lattice.c:98-113):
 * out[0..1] = mean Re/Im Polyakov loop, out[2] = mean plaquette ReTr,
 * out[3] = mean 2x1 rectangle ReTr, out[4 + k*wl + l] = mean (k+1)x(l+1)
 * spatial-temporal Wilson loop, k < wk <= Nx, l < wl <= Nt.  sum_mode
 * DSA_SUM_TREE: fixed-shape FP64 tree (deterministic); DSA_SUM_ORDERED: the
 * reference's one-thread summation order (one lane; small lattices). */
int dsa_chain_measure(dsa_chain *c, uint32_t wk, uint32_t wl, int32_t sum_mode, double *out);
/* dsa_chain_measure without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code between HIP graph
 * captures, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as dsa_chain_measure (bit-identical), but the
 * per-block partial sums, the Wilson-loop line products and the result
 * staging live in a library-owned static device array of
 * DSA_NOALLOC_SCRATCH_BYTES per GPU instead of memory the chain allocates
 * on first use; calls no hipMalloc/hipFree.  DSA_E_RANGE when the lattice
 * and grid need more: 8*(4+wk*wl)*(1 + blocks) bytes (blocks <= 1024, tree
 * mode) plus 2*V*(wl + (D-1)*wk) for a Wilson grid, each region rounded up
 * to 256.  Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU (chains measured concurrently from several threads) race and
 * must be serialised by the caller; each call has finished with the array
 * when it returns. */
#define DSA_NOALLOC_SCRATCH_BYTES (32u * 1024u * 1024u)
int dsa_chain_measure_noalloc(dsa_chain *c, uint32_t wk, uint32_t wl, int32_t sum_mode, double *out);
/* Read / write the gauge field `a` (lattice.h:9, lattice.c:7,15). *//* Read / write the gauge field `a` (lattice.h:9, lattice.c:7,15). */
int dsa_chain_get_links(dsa_chain *c, int32_t *a);
int dsa_chain_set_links(dsa_chain *c, const int32_t *a);
/* Minstd per-site stream states (V values) for checkpointing in that mode. */
int dsa_chain_get_rng_state(dsa_chain *c, uint32_t *state);
int dsa_chain_set_rng_state(dsa_chain *c, const uint32_t *state);
/* get_staples / get_rect_loops (dym-mod-metro.cpp:63-142) of every link,
 * computed by the sweep kernel's staple code: st[(i*D+mu)*2(D-1) + j],
 * re[(i*D+mu)*6(D-1) + k]. */
int dsa_chain_staples(dsa_chain *c, int32_t *st, int32_t *re);
/* Element index of every closed loop of one kind (lattice.c:76-124), in the
 * reference's enumeration: PLAQ [i][d1<d2]; RECT wilson(i,d1,d2,2,1) [i][d1!=d2];
 * POLY polyakov of the t=0 sites ascending; WILSON wilson(i,d1,0,k+1,l+1) [i][d1>=1]. */
int dsa_chain_loops(dsa_chain *c, int32_t kind, uint32_t k, uint32_t l, int32_t *out);
/* Number of passes per sweep and colours per direction of the visiting order. */
int dsa_chain_schedule_info(dsa_chain *c, uint32_t *npasses, uint32_t *ncolours);
/* The visiting order a chain of this shape would use, computed on the host
 * (no device needed).  Call with NULL arrays to get *npasses and *nsites;
 * then pass_mu[npasses], pass_count[npasses], sites[nsites] receive, pass by
 * pass, the direction and the ascending site list.  Replaces the reference's
 * site_parity loop (dym-mod-metro.cpp:152-159, 261-268). */
int dsa_schedule_query(uint32_t D, uint32_t Nt, uint32_t Nx, int32_t order, uint32_t *npasses, uint32_t *nsites,
                       uint32_t *pass_mu, uint32_t *pass_count, uint32_t *sites);
/* Milliseconds the last dsa_chain_sweep spent on the device (HIP events). */
int dsa_chain_last_sweep_ms(dsa_chain *c, double *ms);

/* ---- configuration files ----------------------------------------------------
 * Raw configuration: the int32 link array a[i*D+mu] in host byte order, V*D
 * entries, no header -- llr2d's save_cfg/load_cfg (llr2d.cpp:262-271).  Load
 * fails (DSA_E_RANGE) unless the file holds exactly n entries, and rejects
 * entries outside [0, P) on a chain. */
int dsa_cfg_save(const char *path, const int32_t *a, uint64_t n);
int dsa_cfg_load(const char *path, int32_t *a, uint64_t n);
int dsa_chain_save_cfg(dsa_chain *c, const char *path);
int dsa_chain_load_cfg(dsa_chain *c, const char *path);
/* Writes the chain's configuration as SU(3) matrices in the Kentucky byte
 * order that dym-mod-metro-savecfg.cpp:55-86 (savekyconfig) writes: V*4*18
 * big-endian doubles, link (x, d) element (r, c, part) at index x1 + Nx*(x2 +
 * Nx*(x3 + Nx*(x0 + Nt*(part + 2*(r + 3*(c + 3*((d+3)%4))))))), to
 * "<prefix>lat-beta1%.3f-beta2%.3f-nt%02d-nx%02d-num%04d".  D = 4 and a group
 * file with appended matrices only.  The file name is copied to name (cap
 * bytes) when name is not NULL.  The matrices are expanded on the device. */
int dsa_chain_save_kentucky(dsa_chain *c, const char *prefix, int32_t num, char *name, uint64_t cap);

/* Sweep kernel choice (no counterpart in the reference; all kernels give
 * bit-identical links and counters).  TABLE: every trial action from element
 * ReTr values through L2 gathers of the P x P table (the reference's
 * arithmetic, dym-mod-metro.cpp:164-192).  CLASS: trial actions from
 * quantised conjugacy-class values held in LDS (int16 values when they fit,
 * else byte class indices); a hit whose decision could depend on the
 * quantisation or the table's within-class ReTr rounding is decided again
 * from element values.  CLASS_NARROW forces the byte-index form.  AUTO
 * (default): the class kernel when the group's class tables fit the
 * device's LDS (it is faster at every lattice size measured, 8^4 to
 * 32^3 x 64), in the byte-index form where that fits, else as CLASS.
 * MATRIX (groups whose file appends SU(3) matrices, e.g. S(1080)): the
 * staple-times-link products by table gathers, their weighted sum as an FP32
 * 3x3 complex matrix from those matrices in LDS, and each trial action as one
 * trace of a product with it; a hit whose decision the rounding bound could
 * change is decided again from element values.
 * FIELD (groups with a GF(4) form: a faithful representation by 3x3
 * matrices over GF(4), found and checked against every product of the table
 * when the group is loaded, e.g. S(1080), S(216), S(108)): the byte-index
 * class update with its staple-times-link products formed as packed GF(4)
 * matrix products in registers instead of table gathers; AUTO takes it
 * wherever CLASS_NARROW would run and the form exists, except for a sweep
 * call after one whose acceptance was below 1e-3 (a frozen chain: its links
 * near one another make the table gathers cache hits, and the byte-index
 * form then also tests each hit on its plaquette terms first).
 * margin_scale (default 1) multiplies the bound
 * that selects those re-decisions: a huge value makes every hit re-decided
 * (a test of that path). */
enum { DSA_KERNEL_AUTO = 0, DSA_KERNEL_TABLE = 1, DSA_KERNEL_CLASS = 2, DSA_KERNEL_CLASS_NARROW = 3,
       DSA_KERNEL_MATRIX = 4, DSA_KERNEL_FIELD = 5 };
int dsa_chain_set_kernel(dsa_chain *c, int32_t kernel, double margin_scale);
/* The kernel the chain's next sweep call runs: TABLE, CLASS (the int16-value
 * form), CLASS_NARROW (the byte-index form), MATRIX or FIELD. */
int dsa_chain_kernel_form(dsa_chain *c, int32_t *form);
/* Hits the CLASS or MATRIX kernel re-decided from element values since creation. */
int dsa_chain_exact_decisions(dsa_chain *c, uint64_t *n);

/* ---- many chains at once -------------------------------------------------
 * Replaces the reference's process-per-point scans (evermore.sh:1-23,
 * jobs2run:1-55; each process runs dym-mod-metro.cpp:199-337): chains of
 * different shapes (one D per batch), groups, couplings and seeds are swept
 * together, one launch per pass index.  Each chain's links equal those of the
 * same dsa_chain run alone.  Usage: create, add chains, finalize, then sweep /
 * measure / run_cycles.  dsa_batch_add returns the chain's index. */
typedef struct dsa_batch dsa_batch;
int dsa_batch_create(dsa_batch **out);
int dsa_batch_add(dsa_batch *b, const dsa_group *g, const dsa_chain_params *p, uint32_t *index);
int dsa_batch_finalize(dsa_batch *b);
/* Before finalize: the sweep kernel of the batch's chains (no counterpart in
 * the reference; identical results either way).  TABLE: element values
 * (dym-mod-metro.cpp:164-192).  CLASS / CLASS_NARROW: the byte-index class
 * update of DSA_KERNEL_CLASS_NARROW for every chain (an error at finalize
 * for a group without class tables that fit).  FIELD: that update with the
 * GF(4) products of DSA_KERNEL_FIELD for every chain (an error at finalize
 * for a group without that form).  AUTO (default): the class update for the
 * chains where it fits and its decision bound is below 1e-3, with GF(4)
 * products where the group has the form, the table update for the others. */
int dsa_batch_set_kernel(dsa_batch *b, int32_t kernel, double margin_scale);
/* Hits of the batch's class updates re-decided from element values. */
int dsa_batch_exact_decisions(dsa_batch *b, uint64_t *n);
/* Before finalize: links per sweep block (one per lane): 256, 512 or 1024,
 * or 0 (default) to let finalize time each size over two sweeps and keep the
 * fastest, restoring the chains' state afterwards, so results do not depend
 * on the choice (no counterpart in the reference, which runs one process per
 * point). */
int dsa_batch_set_task_links(dsa_batch *b, uint32_t links);
/* The size in use and, when tuned, ms per sweep of each candidate (256, 512,
 * 1024; the first ncand written to ms_per_sweep, which may be NULL). */
int dsa_batch_task_links(dsa_batch *b, uint32_t *links, double *ms_per_sweep, uint32_t ncand);
void dsa_batch_destroy(dsa_batch *b);
/* n sweeps of every chain (update(), dym-mod-metro.cpp:146-197). */
int dsa_batch_sweep(dsa_batch *b, uint32_t n);
/* out[c*4 + 0..3]: mean Re P, Im P, plaquette, 2x1 rectangle of chain c
 * (dym-mod-metro.cpp:282-303), FP64 fixed-shape tree per chain. */
int dsa_batch_measure(dsa_batch *b, double *out);
/* ncycles x (sweeps_per_cycle sweeps + measure), captured once in a hipGraph
 * and replayed; out[k*nchains*4 + c*4 + q] is cycle k's measurement. */
int dsa_batch_run_cycles(dsa_batch *b, uint32_t ncycles, uint32_t sweeps_per_cycle, double *out);
int dsa_batch_info(dsa_batch *b, uint32_t *nchains, uint32_t *npasses, uint32_t *nblocks_per_sweep);
int dsa_batch_get_links(dsa_batch *b, uint32_t chain, int32_t *a);
/* per-chain cumulative hit and acc counters (arrays of nchains) */
int dsa_batch_counters(dsa_batch *b, uint64_t *hit, uint64_t *acc);
int dsa_batch_last_ms(dsa_batch *b, double *ms);

/* ---- LLR: energy-window constrained chains ---------------------------------
 * Replaces update_constrained / seed_window / rm_solve of llr.cpp:178-250
 * (llr2d.cpp:178-259 for two windows).  Each chain samples
 * exp(-a E_w - bg E_bg) restricted to E_w in [Ewin - hw, Ewin + hw], E_w = S1
 * (axis 1, plaquettes) or S2 (axis 2, rectangles); it first drives its
 * energy into the window (greedy moves, at most `cap` sweeps), then runs
 * Robbins-Monro iterations a += 12/(W^2 (m+1)) (<E_w> - Ewin), W = 2 hw, each
 * over K constrained sweeps, until |delta a| < tol or NRM iterations.  The
 * window is tested per colour pass (a pass leaving the window is undone) and
 * energies are kept in 2^-32 fixed point.  Chains live in the slots of one
 * engine (one D per engine) and are advanced together. */
typedef struct {
  uint32_t D, Nt, Nx;
  int32_t axis;       /* 1: window S1 (llr.cpp AXIS=1), 2: window S2,
                         3: box on (S1, S2) with tilts (a0, a20) (llr2d.cpp) */
  double bg;          /* coupling of the other action (llr.cpp BG) */
  double Ewin, hw;    /* window centre and half width */
  double a0;          /* initial tilt */
  int32_t seed;       /* Philox key of the chain */
  int32_t start;      /* DSA_START_HOT | DSA_START_COLD (ignored when links are given) */
  uint32_t K, NRM;    /* constrained sweeps per RM iteration; iteration cap
                         (NRM = 0: stop once the window is reached, the
                         drive-in of llr2d's seed pass) */
  double tol;         /* stop once |delta a| < tol; 0 runs NRM iterations */
  uint32_t cap;       /* drive-in sweep cap (0: 20000, llr.cpp:225) */
  /* axis 3 only: the S2 window and its initial tilt; Ewin, hw, a0 are the S1
   * ones (llr2d.cpp:180-259: drive-in greedy on the distance to the cell
   * centre in units of the half widths, RM on both tilts) */
  double Ewin2, hw2, a20;
} dsa_llr_params;

/* DSA_LLR_REJECTED: a served request refused at admission (footprint above
 * the engine's whole budget, or parameters the engine refuses). */
enum { DSA_LLR_SEED = 0, DSA_LLR_RM = 1, DSA_LLR_DONE = 2, DSA_LLR_IDLE = 3, DSA_LLR_REJECTED = 4 };

typedef struct {
  double a;           /* current tilt a_n */
  double S1, S2;      /* current plaquette / rectangle actions */
  double last_da;     /* |delta a| of the last iteration */
  double maxdrift;    /* max |incremental - recomputed| windowed energy */
  uint32_t phase;     /* DSA_LLR_* */
  uint32_t iterations, seed_sweeps, reached;
  uint64_t sweeps, hit, acc;
  double a2;          /* axis 3: the S2 tilt */
  double drift1, drift2;  /* max |incremental - recomputed| of S1 and of S2
                             (llr2d.cpp:254-255 tracks both) */
} dsa_llr_status;

typedef struct dsa_llr dsa_llr;
/* An engine of at most max_chains chains of dimension D; hbm_budget bounds the
 * summed chain footprints (0: no bound). */
int dsa_llr_create(uint32_t D, uint32_t max_chains, uint64_t hbm_budget, dsa_llr **out);
void dsa_llr_destroy(dsa_llr *L);
int dsa_llr_footprint(const dsa_llr_params *p, uint64_t *bytes);
/* Admits a chain; init_links (reference order, may be NULL) seeds it from a
 * given configuration, as run_pass does from the previous window (llr.cpp:256-274). */
int dsa_llr_add(dsa_llr *L, const dsa_group *g, const dsa_llr_params *p, const int32_t *init_links, uint32_t *slot);
int dsa_llr_remove(dsa_llr *L, uint32_t slot);
/* Pass kernel for chains added from now on (no counterpart in the reference;
 * results are identical either way).  TABLE: element ReTr values through L2
 * gathers, as update_constrained sums them (llr.cpp:178-219).  CLASS or
 * CLASS_NARROW: the byte class tables in LDS with certified decisions, each
 * undecided hit decided again from element values; an error when the group
 * has no class tables that fit.  FIELD: the class path with the
 * staple-times-link products from the group's GF(4) form (DSA_KERNEL_FIELD);
 * an error when the group has no such form.  AUTO (default): the class path
 * where it fits, with GF(4) products where the group has the form.
 * margin_scale >= 1 widens the bound (a huge value decides every hit from
 * element values). */
int dsa_llr_set_kernel(dsa_llr *L, int32_t kernel, double margin_scale);
/* nsweeps sweeps of every active chain (each in its own phase). */
int dsa_llr_step(dsa_llr *L, uint32_t nsweeps);
int dsa_llr_status_get(dsa_llr *L, uint32_t slot, dsa_llr_status *out);
int dsa_llr_get_links(dsa_llr *L, uint32_t slot, int32_t *a);
/* Continuous-batching service of nreq window requests arriving at arrival_s[r]
 * seconds: admission in arrival order by free slot and footprint
 * (dsa_llr_footprint; the sum over admitted chains never exceeds the
 * engine's hbm_budget), `chunk` sweeps per round for every active chain,
 * retirement on convergence.  Outputs per request: admit and done times
 * (seconds from the call) and the final status; a request that can never
 * fit is retired on arrival with phase DSA_LLR_REJECTED and the others are
 * served.  Returns an error only for a device failure.  (The reference
 * serves windows one process at a time: scripts/llr2d_parallel.sh:28-37.) */
int dsa_llr_serve(dsa_llr *L, const dsa_group *const *groups, const dsa_llr_params *req, const double *arrival_s,
                  uint32_t nreq, uint32_t chunk, double *admit_s, double *done_s, dsa_llr_status *status);

/* ---- the reference llr's serial chain, replayed on the device ---------------
 * Replaces llr.cpp's own chain (llr.cpp:178-274, 296-326) exactly as the
 * reference runs it at one thread, for bit-exact replay: per-site
 * std::minstd_rand0 streams seeded iseed + i (unsigned), hot links from site
 * 0's stream through uniform_int_distribution(0, P-1), RECT_MULT calibrated
 * from one link flip (calibrate_rect_mult), links visited serially (d, then
 * i) with the window tested after every hit on running double energies,
 * full-lattice energies recomputed in the reference's serial order.  Links,
 * energies, RECT_MULT and a_n equal the reference's bit for bit (the one
 * deviation possible is glibc's exp against this library's, within 1 ulp).
 * One lane walks the lattice, so this is a parity and replay path for small
 * lattices; dsa_llr_* is the throughput path.  axis 1 or 2 (llr.cpp AXIS),
 * bg its BG.  Create performs main's set-up (llr.cpp:314-320). */
typedef struct dsa_llr_ref dsa_llr_ref;
int dsa_llr_ref_create(const dsa_group *g, uint32_t D, uint32_t Nt, uint32_t Nx, int32_t axis, double bg,
                       int32_t seed, dsa_llr_ref **out);
void dsa_llr_ref_destroy(dsa_llr_ref *R);
/* RECT_MULT as calibrate_rect_mult found it (llr.cpp:155-171). */
int dsa_llr_ref_rect_mult(dsa_llr_ref *R, double *rect_mult);
/* rnd[i].seed(iseed + repeat*1000003 + i) (llr.cpp:326). */
int dsa_llr_ref_reseed(dsa_llr_ref *R, uint32_t repeat);
/* run_pass's start (llr.cpp:258-260): DSA_START_HOT draws every link from
 * site 0's stream, DSA_START_COLD sets the identity; then resync_E. */
int dsa_llr_ref_start(dsa_llr_ref *R, int32_t start);
/* seed_window (llr.cpp:223-231): greedy drive-in sweeps until the running
 * windowed energy is in [Elo, Ehi] or 20000 sweeps, then resync_E. */
int dsa_llr_ref_seed_window(dsa_llr_ref *R, double Elo, double Ehi, double a_guess, int32_t *reached);
/* rm_solve (llr.cpp:234-250): NRM iterations of K constrained sweeps. */
int dsa_llr_ref_rm_solve(dsa_llr_ref *R, double Ewin, double hw, double a0, uint32_t K, uint32_t NRM, double *a);
/* Running energies S1, S2 (llr.cpp:38-39), g_maxdrift and the hit/acc totals;
 * any pointer may be NULL. */
int dsa_llr_ref_state(dsa_llr_ref *R, double *S1, double *S2, double *maxdrift, uint64_t *hit, uint64_t *acc);
int dsa_llr_ref_get_links(dsa_llr_ref *R, int32_t *a);

/* Measurement helpers (no counterpart in the reference, which times with
 * gettimeofday around update(), timer.cpp:11-27).  Mode 1: every pass launch
 * is bracketed by HIP events on the chain's stream (each event record also
 * orders a cache release between passes, which costs the sweep ~9 %); mode
 * 2: one event pair around each dsa_chain_sweep call, so the time per launch
 * includes the gaps between launches.  dsa_chain_kernel_time returns the
 * summed milliseconds and the pass launches they cover since the mode was
 * set; mode 0 switches it off. */
int dsa_chain_set_profiling(dsa_chain *c, int32_t mode);
int dsa_chain_kernel_time(dsa_chain *c, double *ms, uint64_t *launches);
/* Random uint16 gathers from the group's P x P table, `blocks` x 256 lanes,
 * 24 gathers in flight per lane, `rounds` rounds: the ceiling the sweep's
 * table gathers are bound by. */
int dsa_probe_gather(const dsa_group *g, uint32_t blocks, uint32_t rounds, double *gathers_per_s);
/* A plain device-to-device copy kernel of `bytes`, averaged over `reps`. */
int dsa_probe_copy(uint64_t bytes, uint32_t reps, double *ms_per_copy);

const char *dsa_last_error(void);
const char *dsa_version(void);

#ifdef __cplusplus
}
#endif
#endif
