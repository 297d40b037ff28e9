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
};

size_t sweep_lds_bytes(const Tables &T);
hipError_t launch_sweep(const SweepArgs &A, int max_blocks, hipStream_t s);

// Philox hot start: link l (reference order l = i*D+mu) <- floor(P*u).
hipError_t launch_init_philox(uint16_t *links, const Geom &G, uint32_t P, uint32_t seed, hipStream_t s);
hipError_t launch_fill(uint16_t *links, const Geom &G, uint16_t v, hipStream_t s);
// reference-order int32 a[i*D+mu] <-> device layout
hipError_t launch_pack(uint16_t *links, const int32_t *aos, const Geom &G, hipStream_t s);
hipError_t launch_unpack(int32_t *aos, const uint16_t *links, const Geom &G, hipStream_t s);

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
// The reference's one-thread order (dym-mod-metro.cpp:282-318): one lane.
hipError_t launch_measure_ordered(double *out, const uint16_t *links, const uint16_t *lines, const Tables &T,
                                  const Geom &G, uint32_t wk, uint32_t wl, uint32_t nmax_t, uint32_t nmax_x,
                                  hipStream_t s);

}  // namespace dsa
