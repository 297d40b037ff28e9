// kernels.hpp -- launch interface of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dsa {

constexpr int kDMax = 5;

// Lattice geometry on the device.  Site i = x0 + Nt*(x1 + Nx*(x2 + ...))
// (lattice.c:39-51).  Links live per direction (mu*V + pos), and when Nt is
// even pos = parity(x)*V/2 + i/2 (sites with equal i/2 differ in x0 only, so
// exactly one of each pair is even): a per-direction, per-parity layout.
struct Geom {
  int D;
  int eo;
  uint32_t V, halfV;
  uint32_t ext[kDMax];
  uint32_t stride[kDMax];
};

struct Tables {
  const uint16_t *mult;  // P*P
  const uint16_t *inv;   // P
  const double *retr;    // P
  const double *imtr;    // P
  const uint16_t *pool;  // npool
  uint32_t P, npool, id;
  // conjugacy classes (null when the group has none usable): class of each
  // element, cprod[j*P + t] = class(pool[j] * t) padded to 16 bytes, and one
  // ReTr per class in units of 2^-24
  const uint8_t *cls, *cprod;
  const int32_t *rcq;
  const int32_t *mdq;  // per class, the bound of Group::mdq (null if none)
  uint32_t ncls;
  // wide form: class values in units of 2^-13 (null if absent)
  const int16_t *cval, *cval0;
  // the group's SU(3) matrices in FP32, 20 floats per element (re[9], im[9],
  // 2 unused), row-major (null where the group file has none)
  const float *mat;
  // the group's GF(4) form (gf4.hpp; null where it has none): packed matrix
  // of each element and of its inverse, and the element of each key
  const uint32_t *fmat, *finv;
  const uint16_t *fkey;
};

enum RngKind : int { kRngMinstd = 0, kRngPhilox = 1 };

struct SweepArgs {
  uint16_t *links;
  const uint32_t *sites;  // the pass's sites
  uint32_t *mstate;       // minstd per-site state (V) or null
  unsigned long long *acc;
  Tables T;
  Geom G;
  uint32_t count;
  int mu;
  int rng;
  double beta0, beta1, beta2;
  uint32_t seed;
  uint32_t sweep_lo, sweep_hi;
  // class-product kernel (kClsNone, kClsNarrow, kClsWide) or the matrix
  // kernel (kClsMatrix), the bound on |exact - estimated| trial-action
  // differences below which a decision is recomputed exactly (the matrix
  // kernel's grows by cmargin_acc per accepted hit), and the count of such
  // recomputations
  int use_cls;
  double cmargin, cmargin_acc;
  unsigned long long *fall;
  // narrow class form: test each hit on its plaquette terms first, with the
  // rectangle terms bounded (rejects without the rectangle lookups; pays
  // where nearly every hit of a wavefront is rejected that way, i.e. at very
  // low acceptance -- a wavefront runs the rectangle terms if any lane needs
  // them); identical results either way
  int bound;
  // the device's LDS bytes per block: the small-pass class forms run only
  // where their draw arrays fit beside the tables
  size_t lds_max;
  // kClsField: every pass through the field kernel (else passes below the
  // small-pass sizes take the narrow small-pass kernel)
  int field_all;
  // compute units of the device (host side: the field kernel's grid)
  uint32_t ncu;
};

size_t sweep_lds_bytes(const Tables &T);
// kClsField: the narrow class update with the staple-times-link products
// formed from the group's GF(4) matrices (no table gathers for them)
enum ClsMode : int { kClsNone = 0, kClsNarrow = 1, kClsWide = 2, kClsMatrix = 3, kClsField = 4 };
size_t sweep_cls_lds_bytes(const Tables &T, int mode);
__host__ __device__ inline size_t pad16(size_t n) { return (n + 15) & ~(size_t)15; }
hipError_t launch_sweep(const SweepArgs &A, int max_blocks, hipStream_t s);

// ---- many independent chains in one launch per pass ----
struct BatchChain {
  uint16_t *links;
  const uint32_t *sites;  // the chain's schedule (all passes concatenated)
  uint32_t *mstate;       // minstd state or null
  Tables T;
  Geom G;
  double beta0, beta1, beta2;
  uint32_t seed;
  int rng;
  uint64_t sweep_base;    // sweeps done before joining the batch counter
  // class-product update (byte tables in LDS): on or off, its decision
  // bound, and the batch's count of hits re-decided from element values
  int cls;
  double cmargin;
  unsigned long long *fall;
  int bound;  // as SweepArgs::bound
  // with cls: the staple-times-link products from the group's GF(4) form
  // (its tables staged after the class tables)
  int field;
};

// The plaquette-first test is chosen for a sweep call when the chain's
// acceptance over its previous call was below this.
constexpr double kBoundAcceptance = 1e-3;

// A block's work: `count` (<= 256) consecutive sites of one chain's pass.
struct BatchTask {
  uint32_t chain;
  uint32_t offset;  // into the chain's sites
  uint32_t count;
  int32_t mu;
};

// A block's share of one chain's measurement: sites [site0, site0 + nsites).
struct ObsTask {
  uint32_t chain;
  uint32_t site0;
  uint32_t nsites;
};

// ---- LLR: energy-window constrained chains (llr.cpp, llr2d.cpp) ----
enum LlrPhase : uint32_t { kLlrSeed = 0, kLlrRM = 1, kLlrDone = 2, kLlrIdle = 3 };

// One pass of a chain's schedule: sites [offset, offset + count), direction mu.
struct PassRange {
  uint32_t offset, count;
  int32_t mu;
  uint32_t pad;
};

struct LlrChain {
  BatchChain base;   // links, sites, tables, geometry, Philox seed
  uint16_t *old;     // previous link values of the current pass, indexed like sites
  long long *dE1;    // per link: proposed change of S1 (fixed point), indexed like sites
  long long *dE2;    // per link: proposed change of S2
  uint8_t *nacc;     // per link: accepted hits of its proposal, + 0x80 if it moved the link
  const uint32_t *pos;  // per link: its index in the links array, indexed like sites
  const PassRange *passes;
  int axis;          // 1: window S1 (plaquettes), 2: window S2 (rectangles), 3: box on both
  double bg;         // coupling of the other action (axes 1, 2)
  long long lo, hi;  // window [Ewin - hw, Ewin + hw] in fixed point (axis 3: on S1)
  double Ewin, W;    // centre and width 2*hw (Robbins-Monro)
  long long lo2, hi2;  // axis 3: the S2 window
  double Ewin2, W2;
  uint32_t K, NRM;   // constrained sweeps per RM iteration; iteration cap
  double tol;        // stop once |delta a| < tol (0: run NRM iterations)
  uint32_t cap;      // drive-in sweep cap
  uint32_t npass;    // passes per sweep of this chain
  // class path (byte class tables in LDS, certified decisions): on or off,
  // and the terms of the bound on |element - class| sums: value spread and
  // quantisation (ce), element-sum rounding (cf), magnitude (cr) of the
  // plaquette (1) and rectangle (2) parts; cmscale multiplies the bound
  int cls;
  double cmscale, ce1, ce2, cf1, cf2, cr1, cr2;
  int field;  // with cls: the staple-times-link products from the GF(4) form
};

struct LlrState {
  long long S1, S2;  // running plaquette / rectangle actions (fixed point)
  long long r1, r2;  // resync accumulators
  double a;          // tilt a_n (axis 3: of S1)
  double sumE;       // sum of the windowed energy over the iteration's sweeps (axis 3: S1)
  double a2, sumE2;  // axis 3: tilt of S2 and its sum
  double last_da;    // |delta a| of the last RM step
  double maxdrift;   // max |incremental - recomputed| energy
  double drift1, drift2;  // the same for S1 and for S2 separately (llr2d.cpp:254-255)
  uint32_t phase, k, m, seed_sweeps, drive, resync, reached, pad;
  uint64_t sweep;    // Philox sweep counter
  unsigned long long acc, hit;
};

// LDS of the LLR pass kernel for one group: ReTr, inverses and pool, plus
// (class path) the class of each element and the npool x P class table,
// plus (field) the GF(4) matrices, their inverses' and the key table.
__host__ __device__ inline size_t llr_lds_bytes(const Tables &T, bool cls, bool field = false) {
  const size_t e = (size_t)T.P * sizeof(double) + (size_t)T.P * sizeof(uint16_t) + (size_t)T.npool * sizeof(uint16_t);
  if (!cls) return e;
  const size_t c = pad16(e) + pad16(T.P) + pad16((size_t)T.npool * T.P);
  return field ? c + 2 * pad16(4 * (size_t)T.P) + 4096 * sizeof(uint16_t) : c;
}

// Static LDS of the LLR pass kernel (class values, the class path's uniforms).
size_t llr_pass_static_lds();

// Proposals: every link of pass `pass` of every active chain.
hipError_t launch_llr_pass(int D, const LlrChain *chains, LlrState *st, const BatchTask *tasks, uint32_t ntasks,
                           size_t lds_bytes, hipStream_t s);
// Per chain (one wave each): accept the pass's link proposals in site order
// while the running energy obeys the window (or, in drive-in, approaches it),
// restoring the links whose proposal is refused.
// one block per chain idx[0 .. n) (chains 0 .. n when idx is null)
hipError_t launch_llr_decide(const LlrChain *chains, LlrState *st, const uint32_t *idx, uint32_t n, uint32_t pass,
                             hipStream_t s);
// idx: the chains idx[0..n), or chains 0..n-1 when null; the resync's
// tasks cover the same chains
hipError_t launch_llr_sweep_end(const LlrChain *chains, LlrState *st, const uint32_t *idx, uint32_t n,
                                hipStream_t s);
hipError_t launch_llr_resync(const LlrChain *chains, LlrState *st, const uint32_t *idx, uint32_t n,
                             const ObsTask *tasks, uint32_t ntasks, hipStream_t s);

// ---- the reference llr's serial chain, replayed (llr_ref.hip) ----
struct LrefState {
  double S1, S2;       // running energies (llr.cpp:38-39)
  double rect_mult;    // RECT_MULT (llr.cpp:135, calibrated :155-171)
  double maxdrift;     // g_maxdrift
  double sum;          // sum of Ecur over the current RM iteration's sweeps
  double a;            // the tilt of rm_solve
  double cal_m1_old, cal_m1_new;  // calibration scratch
  unsigned long long hit, acc;
  uint32_t drive;      // drive-in sweeps of the current seed_window
  uint32_t cal_b;
};
struct LrefArgs {
  uint16_t *links;
  uint32_t *ms;        // per-site minstd_rand0 states
  Tables T;
  Geom G;
  int axis;
  double bg;
  LrefState *st;
  double *t1, *t2;     // terms of recompute_S1 (V D(D-1)/2) and sum_m1 (V D 6(D-1))
};
// update_constrained (driveIn sweeps skipped once inside [Elo, Ehi] or after
// cap of them; a NaN tilt takes st->a)
hipError_t launch_lref_update(const LrefArgs &A, double a_tilt, double Elo, double Ehi, int driveIn, uint32_t cap,
                              hipStream_t s);
// sums[0] = recompute_S1, sums[1] = sum_m1, each in the reference's serial order
hipError_t launch_lref_sums(const LrefArgs &A, double *sums, hipStream_t s);
// mode 0: resync_E; mode 1: the end of Robbins-Monro iteration m (a, drift, resync)
hipError_t launch_lref_resync(const LrefArgs &A, const double *sums, int mode, double Ewin, double W, uint32_t m,
                              uint32_t K, hipStream_t s);
hipError_t launch_lref_hot(const LrefArgs &A, hipStream_t s);
hipError_t launch_lref_seed(const LrefArgs &A, uint32_t base, hipStream_t s);
hipError_t launch_lref_calibrate(const LrefArgs &A, const double *sums_old, const double *sums_new, int step,
                                 hipStream_t s);
hipError_t launch_lref_set(const LrefArgs &A, double a, uint32_t drive, hipStream_t s);

// `block` lanes per block (a multiple of 64, at most kBatchBlockMax); every
// task holds at most `block` links.
constexpr int kBatchBlockMax = 1024;
hipError_t launch_sweep_batch(int D, const BatchChain *chains, const BatchTask *tasks, uint32_t ntasks,
                              uint32_t block, size_t lds_bytes, const unsigned long long *sweep_ctr,
                              unsigned long long *acc,
                              hipStream_t s);
hipError_t launch_count_sweep(unsigned long long *sweep_ctr, hipStream_t s);
// out[c*4 + q]: mean Re P, Im P, plaquette, 2x1 rectangle of chain c; partials
// holds 4*ntasks doubles; chain c owns tasks [first[c], first[c] + n[c]).
hipError_t launch_measure_batch(double *out, double *partials, const BatchChain *chains, const ObsTask *tasks,
                                uint32_t ntasks, const uint32_t *first, const uint32_t *n, uint32_t nchains,
                                hipStream_t s);

// Philox hot start: link l (reference order l = i*D+mu) <- floor(P*u).
hipError_t launch_init_philox(uint16_t *links, const Geom &G, uint32_t P, uint32_t seed, hipStream_t s);
hipError_t launch_fill(uint16_t *links, const Geom &G, uint16_t v, hipStream_t s);
// reference-order int32 a[i*D+mu] <-> device layout
hipError_t launch_pack(uint16_t *links, const int32_t *aos, const Geom &G, hipStream_t s);
hipError_t launch_unpack(int32_t *aos, const uint16_t *links, const Geom &G, hipStream_t s);

// SU(3) matrices of every link in the Kentucky layout (D = 4): out[o + V*(part
// + 2*(r + 3*(c + 3*((d+3)%4))))] for output site o = x1 + Nx*(x2 + Nx*(x3 +
// Nx*x0)), each double byte-swapped to big-endian; mats[18*e + part + 2*(c+3r)].
hipError_t launch_kentucky(double *out, const uint16_t *links, const double *mats, const Geom &G, hipStream_t s);

// Staples of every link (parity checks): st[(i*D+mu)*2(D-1)+j], re[...*6(D-1)+k]
hipError_t launch_staples(int32_t *st, int32_t *re, const uint16_t *links, const Tables &T, const Geom &G,
                          hipStream_t s);

// ---- observables ----
enum LoopKind : int { kLoopPlaq = 0, kLoopRect = 1, kLoopPoly = 2, kLoopWilson = 3 };
// element index of every loop, in the reference's enumeration order
hipError_t launch_loops(int32_t *out, int kind, int k, int l, const uint16_t *links, const uint16_t *lines,
                        const Tables &T, const Geom &G, uint32_t nmax_t, uint32_t nmax_x, hipStream_t s);

// line products L_d(x,n) = U_d(x)U_d(x+d)...U_d(x+(n-1)d), n = 1..nmax_d;
// lines[(d == 0 ? n-1 : nmax_t + (d-1)*nmax_x + n-1) * V + site]
hipError_t launch_lines(uint16_t *lines, const uint16_t *links, const Tables &T, const Geom &G, uint32_t nmax_t,
                        uint32_t nmax_x, hipStream_t s);

// Deterministic FP64 tree sums.  partials: per block results; out: final values
//   out[0..1] = Re/Im Polyakov sums, out[2] = plaquette sum, out[3] = rectangle
//   sum, out[4 + k*wl + l] = Wilson (k+1)x(l+1) sums (unnormalised).
uint32_t obs_blocks(const Geom &G);
hipError_t launch_measure_tree(double *out, double *partials, const uint16_t *links, const uint16_t *lines,
                               const Tables &T, const Geom &G, uint32_t wk, uint32_t wl, uint32_t nmax_t,
                               uint32_t nmax_x, hipStream_t s);
// Bytes of the per-GPU static work space of the entry points that allocate
// nothing, and its address on the current device (looked up once per device).
constexpr size_t kScratchBytes = size_t(32) << 20;
hipError_t scratch_region(void **ptr);
// The reference's one-thread order (dym-mod-metro.cpp:282-318): one lane.
hipError_t launch_measure_ordered(double *out, const uint16_t *links, const uint16_t *lines, const Tables &T,
                                  const Geom &G, uint32_t wk, uint32_t wl, uint32_t nmax_t, uint32_t nmax_x,
                                  hipStream_t s);

}  // namespace dsa
