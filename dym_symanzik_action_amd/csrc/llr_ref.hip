// llr_ref.hip -- the reference llr's serial constrained chain on the device:
// the replay of llr.cpp (update_constrained llr.cpp:178-219, the energy
// recomputations :46-55 and :139-151, calibrate_rect_mult :155-171, the
// Robbins-Monro step :234-250 and run_pass's starts :256-274) with the
// reference's own streams, visiting order and floating-point operations, so
// links, energies and a_n equal the reference's bit for bit.
//
// The reference's update is serial: every hit tests the window on the
// running energy, which every earlier hit of the sweep has changed, so one
// lane walks the links in the reference's order (d, then i).  The full-lattice
// sums it recomputes are summed in its serial order too: their terms are
// formed in parallel into a buffer and one lane adds them in sequence.  This
// is a parity and replay path for small lattices; the device LLR engine
// (llr.hip) is the throughput path.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "dsa_math.hpp"
#include "kernels.hpp"

namespace dsa {

namespace {

// update_constrained: one sweep.  driveIn sweeps are skipped once the running
// energy is inside [Elo, Ehi] or `cap` drive-in sweeps have run (the loop
// condition of seed_window, llr.cpp:225-228), so the host can enqueue several
// at once; a tilt of NaN takes the state's a (the rm_solve iterations).
template <int D>
__global__ void k_lref_update(uint16_t *links, uint32_t *ms, Tables Tb, Geom G, int axis, double bg, LrefState *st,
                              double a_tilt, double Elo, double Ehi, int driveIn, uint32_t cap) {
  extern __shared__ __align__(16) unsigned char smem[];
  double *s_retr = reinterpret_cast<double *>(smem);
  for (uint32_t e = threadIdx.x; e < Tb.P; e += blockDim.x) s_retr[e] = Tb.retr[e];
  __syncthreads();
  if (threadIdx.x != 0) return;
  constexpr int NS = 2 * (D - 1), NR = 6 * (D - 1);
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  double S1 = st->S1, S2 = st->S2;
  if (driveIn) {
    const double Ec = axis == 1 ? S1 : S2;
    if ((Ec >= Elo && Ec <= Ehi) || st->drive >= cap) return;
    st->drive += 1;
  }
  if (a_tilt != a_tilt) a_tilt = st->a;
  const double Emid = 0.5 * (Elo + Ehi);
  const double npool = (double)Tb.npool;
  unsigned long long hit = 0, acc = 0;
  for (int d = 0; d < D; ++d)
    for (uint32_t i = 0; i < G.V; ++i) {
      const Site<D> s = decode<D>(G, i);
      uint32_t stp[NS], re[NR];
      build_staples<D>(links, G, T, s, d, stp, re);
      const size_t at = (size_t)d * G.V + lpos(G, i, s.par);
      uint32_t b = links[at];
      double r1 = 0, m1 = 0;
      for (int j = 0; j < NS; ++j) r1 += s_retr[T.M(b, stp[j])];
      for (int k = 0; k < NR; ++k) m1 += s_retr[T.M(b, re[k])];
      uint32_t x = ms[i];
      for (int h = 0; h < kNHit; ++h) {
        const uint32_t bnew = T.M(b, Tb.pool[(uint32_t)(npool * minstd_canonical(x))]);
        double r1n = 0, m1n = 0;
        for (int j = 0; j < NS; ++j) r1n += s_retr[T.M(bnew, stp[j])];
        for (int k = 0; k < NR; ++k) m1n += s_retr[T.M(bnew, re[k])];
        const double dS1 = r1n - r1, dS2 = m1n - m1;
        const double dEw = axis == 1 ? dS1 : dS2;
        const double dEbg = axis == 1 ? dS2 : dS1;
        const double Ec = axis == 1 ? S1 : S2;
        const double Enew = Ec + dEw;
        bool accept = false;
        if (driveIn && (Ec < Elo || Ec > Ehi)) {
          accept = fabs(Enew - Emid) < fabs(Ec - Emid);
        } else if (Enew >= Elo && Enew <= Ehi) {
          const double w = dsa_exp(-a_tilt * dEw - bg * dEbg);
          accept = w > minstd_canonical(x);
        }
        ++hit;
        if (accept) {
          links[at] = (uint16_t)bnew;
          S1 += dS1;
          S2 += dS2;
          b = bnew;
          r1 = r1n;
          m1 = m1n;
          ++acc;
        }
      }
      ms[i] = x;
    }
  st->S1 = S1;
  st->S2 = S2;
  st->hit += hit;
  st->acc += acc;
  if (!driveIn) st->sum += axis == 1 ? S1 : S2;
}

// The terms of recompute_S1 (llr.cpp:46-55: per site, ReTr of each plaquette
// d1 < d2) and of sum_m1 (llr.cpp:139-149: per link (i, d), ReTr[b r_k] of its
// rectangle staples), at their positions in the reference's serial order.
template <int D>
__global__ void k_lref_terms(const uint16_t *links, Tables Tb, Geom G, double *t1, double *t2) {
  constexpr int NS = 2 * (D - 1), NR = 6 * (D - 1), NP = D * (D - 1) / 2;
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < G.V; i += gridDim.x * blockDim.x) {
    int p = 0;
    for (int d1 = 0; d1 < D; ++d1)
      for (int d2 = d1 + 1; d2 < D; ++d2) t1[(size_t)i * NP + p++] = Tb.retr[plaq_elem(links, G, T, i, d1, d2)];
    const Site<D> s = decode<D>(G, i);
    for (int d = 0; d < D; ++d) {
      uint32_t stp[NS], re[NR];
      build_staples<D>(links, G, T, s, d, stp, re);
      const uint32_t b = links[(size_t)d * G.V + lpos(G, i, s.par)];
      for (int k = 0; k < NR; ++k) t2[((size_t)i * D + d) * NR + k] = Tb.retr[T.M(b, re[k])];
    }
  }
}

// One lane adds n terms in sequence (the reference's one-thread sum).
__global__ void k_lref_serial_sum(const double *t, size_t n, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0;
  size_t k = 0;
  for (; k + 8 <= n; k += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = t[k + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; k < n; ++k) s += t[k];
  *out = s;
}

// After the sums: sums[0] = recompute_S1, sums[1] = sum_m1.  mode 0: resync_E
// (S1, S2 <- the recomputed energies); mode 1: the end of a Robbins-Monro
// iteration (llr.cpp:242-247): a from <E> over the K sweeps, the drift of the
// running energy, then resync_E.
__global__ void k_lref_resync(LrefState *st, const double *sums, int axis, int mode, double Ewin, double W,
                              uint32_t m, uint32_t K) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double S1r = sums[0], S2r = sums[1] / st->rect_mult;
  if (mode == 1) {
    const double g = st->sum / K - Ewin;
    st->a += (12.0 / (W * W * (m + 1))) * g;
    const double drift = fabs((axis == 1 ? st->S1 : st->S2) - (axis == 1 ? S1r : S2r));
    if (drift > st->maxdrift) st->maxdrift = drift;
  }
  st->S1 = S1r;
  st->S2 = S2r;
  st->sum = 0;
}

// run_pass's hot start (llr.cpp:258): every link, in the order a[i*D+d], from
// site 0's stream through uniform_int_distribution(0, P-1) (libstdc++'s
// downscaling: draws at or above P * floor((2^31-3)/P) are redrawn).
__global__ void k_lref_hot(uint16_t *links, uint32_t *ms, Geom G, uint32_t P) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t urngrange = 2147483645u, scaling = urngrange / P, past = P * scaling;
  uint32_t x = ms[0];
  const uint32_t nl = G.V * (uint32_t)G.D;
  for (uint32_t l = 0; l < nl; ++l) {
    uint32_t r;
    do r = minstd_next(x) - 1u;
    while (r >= past);
    const uint32_t i = l / (uint32_t)G.D, d = l - i * (uint32_t)G.D;
    links[(size_t)d * G.V + lpos(G, i, site_parity(G, i))] = (uint16_t)(r / scaling);
  }
  ms[0] = x;
}

// rnd[i].seed(base + i) with unsigned 32-bit arithmetic (llr.cpp:314, 326):
// the engine state is the seed mod 2^31 - 1, or 1 where that is 0.
__global__ void k_lref_seed(uint32_t *ms, uint32_t V, uint32_t base) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
    const uint32_t s = (base + i) % 2147483647u;
    ms[i] = s ? s : 1u;
  }
}

// calibrate_rect_mult (llr.cpp:155-171) around the two full sums: step 0
// forms m1 of link (V/2, 0) before and after b -> b pool[npool/2] and flips
// it; step 1 restores it and sets RECT_MULT = (T_new - T_old) / (m1_new -
// m1_old), T_old = sum_m1 before the flip (sums_old[1]), T_new after it
// (sums_new[1]).
template <int D>
__global__ void k_lref_calibrate(uint16_t *links, Tables Tb, Geom G, LrefState *st, const double *sums_old,
                                 const double *sums_new, int step) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  constexpr int NS = 2 * (D - 1), NR = 6 * (D - 1);
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  const uint32_t i = G.V / 2;
  const Site<D> s = decode<D>(G, i);
  const size_t at = lpos(G, i, s.par);  // direction 0
  if (step == 0) {
    uint32_t stp[NS], re[NR];
    build_staples<D>(links, G, T, s, 0, stp, re);
    const uint32_t b = links[at];
    double m1_old = 0;
    for (int k = 0; k < NR; ++k) m1_old += Tb.retr[T.M(b, re[k])];
    const uint32_t bnew = T.M(b, Tb.pool[Tb.npool / 2]);
    double m1_new = 0;
    for (int k = 0; k < NR; ++k) m1_new += Tb.retr[T.M(bnew, re[k])];
    st->cal_m1_old = m1_old;
    st->cal_m1_new = m1_new;
    st->cal_b = b;
    links[at] = (uint16_t)bnew;
  } else {
    links[at] = (uint16_t)st->cal_b;
    st->rect_mult = (sums_new[1] - sums_old[1]) / (st->cal_m1_new - st->cal_m1_old);
  }
}

__global__ void k_lref_set(LrefState *st, double a, uint32_t drive) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  st->a = a;
  st->drive = drive;
  st->sum = 0;
}

constexpr int kLrefBlock = 64;

inline uint32_t lref_grid(uint64_t n) {
  const uint64_t b = (n + 255) / 256;
  return (uint32_t)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

}  // namespace

hipError_t launch_lref_update(const LrefArgs &A, double a_tilt, double Elo, double Ehi, int driveIn, uint32_t cap,
                              hipStream_t s) {
  const size_t lds = (size_t)A.T.P * sizeof(double);
  const dim3 g(1), b(kLrefBlock);
  switch (A.G.D) {
    case 2: hipLaunchKernelGGL(k_lref_update<2>, g, b, lds, s, A.links, A.ms, A.T, A.G, A.axis, A.bg, A.st, a_tilt,
                               Elo, Ehi, driveIn, cap); break;
    case 3: hipLaunchKernelGGL(k_lref_update<3>, g, b, lds, s, A.links, A.ms, A.T, A.G, A.axis, A.bg, A.st, a_tilt,
                               Elo, Ehi, driveIn, cap); break;
    case 4: hipLaunchKernelGGL(k_lref_update<4>, g, b, lds, s, A.links, A.ms, A.T, A.G, A.axis, A.bg, A.st, a_tilt,
                               Elo, Ehi, driveIn, cap); break;
    case 5: hipLaunchKernelGGL(k_lref_update<5>, g, b, lds, s, A.links, A.ms, A.T, A.G, A.axis, A.bg, A.st, a_tilt,
                               Elo, Ehi, driveIn, cap); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_lref_sums(const LrefArgs &A, double *sums, hipStream_t s) {
  const dim3 g(lref_grid(A.G.V)), b(256);
  switch (A.G.D) {
    case 2: hipLaunchKernelGGL(k_lref_terms<2>, g, b, 0, s, A.links, A.T, A.G, A.t1, A.t2); break;
    case 3: hipLaunchKernelGGL(k_lref_terms<3>, g, b, 0, s, A.links, A.T, A.G, A.t1, A.t2); break;
    case 4: hipLaunchKernelGGL(k_lref_terms<4>, g, b, 0, s, A.links, A.T, A.G, A.t1, A.t2); break;
    case 5: hipLaunchKernelGGL(k_lref_terms<5>, g, b, 0, s, A.links, A.T, A.G, A.t1, A.t2); break;
    default: return hipErrorInvalidValue;
  }
  const size_t D = (size_t)A.G.D, V = A.G.V;
  hipLaunchKernelGGL(k_lref_serial_sum, dim3(1), dim3(64), 0, s, A.t1, V * D * (D - 1) / 2, sums);
  hipLaunchKernelGGL(k_lref_serial_sum, dim3(1), dim3(64), 0, s, A.t2, V * D * 6 * (D - 1), sums + 1);
  return hipGetLastError();
}

hipError_t launch_lref_resync(const LrefArgs &A, const double *sums, int mode, double Ewin, double W, uint32_t m,
                              uint32_t K, hipStream_t s) {
  hipLaunchKernelGGL(k_lref_resync, dim3(1), dim3(64), 0, s, A.st, sums, A.axis, mode, Ewin, W, m, K);
  return hipGetLastError();
}

hipError_t launch_lref_hot(const LrefArgs &A, hipStream_t s) {
  hipLaunchKernelGGL(k_lref_hot, dim3(1), dim3(64), 0, s, A.links, A.ms, A.G, A.T.P);
  return hipGetLastError();
}

hipError_t launch_lref_seed(const LrefArgs &A, uint32_t base, hipStream_t s) {
  hipLaunchKernelGGL(k_lref_seed, dim3(lref_grid(A.G.V)), dim3(256), 0, s, A.ms, A.G.V, base);
  return hipGetLastError();
}

hipError_t launch_lref_calibrate(const LrefArgs &A, const double *sums_old, const double *sums_new, int step,
                                 hipStream_t s) {
  const dim3 g(1), b(64);
  switch (A.G.D) {
    case 2: hipLaunchKernelGGL(k_lref_calibrate<2>, g, b, 0, s, A.links, A.T, A.G, A.st, sums_old, sums_new, step);
      break;
    case 3: hipLaunchKernelGGL(k_lref_calibrate<3>, g, b, 0, s, A.links, A.T, A.G, A.st, sums_old, sums_new, step);
      break;
    case 4: hipLaunchKernelGGL(k_lref_calibrate<4>, g, b, 0, s, A.links, A.T, A.G, A.st, sums_old, sums_new, step);
      break;
    case 5: hipLaunchKernelGGL(k_lref_calibrate<5>, g, b, 0, s, A.links, A.T, A.G, A.st, sums_old, sums_new, step);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_lref_set(const LrefArgs &A, double a, uint32_t drive, hipStream_t s) {
  hipLaunchKernelGGL(k_lref_set, dim3(1), dim3(64), 0, s, A.st, a, drive);
  return hipGetLastError();
}

}  // namespace dsa
