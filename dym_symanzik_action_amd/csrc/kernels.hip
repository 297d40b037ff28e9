// kernels.hip -- gfx950 kernels for the discrete-group Symanzik Metropolis
// sweep and its observables.
//
// The sweep is integer gather work: every staple product and every trial
// action is a lookup in the P x P multiplication table (uint16, 2.3 MB for
// P = 1080, resident in each XCD's 4 MB L2), with ReTr, the inverse table and
// the proposal pool in LDS.  No MFMA: there is no dense contraction here.
#include <hip/hip_runtime.h>

#include "dsa_math.hpp"
#include "kernels.hpp"

namespace dsa {

namespace {

constexpr int kBlock = 256;

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD one
// contiguous range of the pass so neighbouring links share an L2.
__device__ inline uint32_t xcd_block() {
  const uint32_t b = blockIdx.x, n = gridDim.x;
  if ((n & 7u) == 0) return (b & 7u) * (n >> 3) + (b >> 3);
  return b;
}

__device__ inline int wrapc(int v, int n) {
  if (v < 0) v += n;
  if (v < 0) v += n;
  if (v >= n) v -= n;
  if (v >= n) v -= n;
  return v;
}

template <int D>
struct Site {
  int x[D];
  uint32_t i;
  uint32_t par;
};

template <int D>
__device__ inline Site<D> decode(const Geom &G, uint32_t i) {
  Site<D> s;
  s.i = i;
  uint32_t par = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const uint32_t q = i / G.ext[d];
    s.x[d] = (int)(i - q * G.ext[d]);
    par += (uint32_t)s.x[d];
    i = q;
  }
  s.par = par & 1u;
  return s;
}

// s.x[d] for a run-time d without indexing a register array (no scratch).
template <int D>
__device__ inline int coord(const Site<D> &s, int d) {
  int r = s.x[0];
#pragma unroll
  for (int q = 1; q < D; ++q) r = (q == d) ? s.x[q] : r;
  return r;
}

// Site at x + a e_mu + b e_nu (periodic); its parity in *par.
template <int D>
__device__ inline uint32_t shifted(const Geom &G, const Site<D> &s, int mu, int a, int nu, int b, uint32_t *par) {
  uint32_t i = s.i;
  uint32_t p = s.par;
  if (a) {
    const int x = coord<D>(s, mu), n = (int)G.ext[mu];
    const int y = wrapc(x + a, n);
    i += (uint32_t)(y - x) * G.stride[mu];
    p ^= (uint32_t)(y - x) & 1u;
  }
  if (b) {
    const int x = coord<D>(s, nu), n = (int)G.ext[nu];
    const int y = wrapc(x + b, n);
    i += (uint32_t)(y - x) * G.stride[nu];
    p ^= (uint32_t)(y - x) & 1u;
  }
  *par = p;
  return i;
}

__device__ inline uint32_t lpos(const Geom &G, uint32_t site, uint32_t par) {
  return G.eo ? (par ? G.halfV : 0u) + (site >> 1) : site;
}

// Link U_dir(x + a e_mu + b e_nu).
template <int D>
__device__ inline uint32_t U(const uint16_t *links, const Geom &G, const Site<D> &s, int dir, int mu,
                             int a, int nu, int b) {
  uint32_t par;
  const uint32_t site = shifted<D>(G, s, mu, a, nu, b, &par);
  return links[(size_t)dir * G.V + lpos(G, site, par)];
}

struct TabView {
  const uint16_t *__restrict__ mult;
  const uint16_t *inv;  // LDS or global
  uint32_t P;
  __device__ inline uint32_t M(uint32_t a, uint32_t b) const { return mult[a * P + b]; }
  __device__ inline uint32_t I(uint32_t a) const { return inv[a]; }
};

// The 2(D-1) plaquette and 6(D-1) rectangle staples of link (x, mu), in the
// reference's order and with its products: dym-mod-metro.cpp:63-78 and
// 83-142.  For each transverse nu (ascending): up, down plaquette staples;
// rectangles k0 (1x2 up), k1 (2x1 forward-up), k2 (2x1 backward-up), k3 (1x2
// down), k4 (2x1 forward-down), k5 (2x1 backward-down).  The reference's
// leading mult[id][.] is the identity map and is skipped.
template <int D>
__device__ inline void build_staples(const uint16_t *links, const Geom &G, const TabView &T,
                                     const Site<D> &s, int mu, uint32_t *st, uint32_t *re) {
  const uint32_t umu_p = U<D>(links, G, s, mu, mu, 1, mu, 0);   // U_mu(x+mu)
  const uint32_t umu_m = U<D>(links, G, s, mu, mu, -1, mu, 0);  // U_mu(x-mu)
#pragma unroll
  for (int j = 0; j < D - 1; ++j) {
    const int nu = j + (j >= mu ? 1 : 0);
    // U_nu at x + a mu + b nu
    const uint32_t n_10 = U<D>(links, G, s, nu, mu, 1, nu, 0);
    const uint32_t n_00 = U<D>(links, G, s, nu, mu, 0, nu, 0);
    const uint32_t n_1m = U<D>(links, G, s, nu, mu, 1, nu, -1);
    const uint32_t n_0m = U<D>(links, G, s, nu, mu, 0, nu, -1);
    const uint32_t n_11 = U<D>(links, G, s, nu, mu, 1, nu, 1);
    const uint32_t n_01 = U<D>(links, G, s, nu, mu, 0, nu, 1);
    const uint32_t n_20 = U<D>(links, G, s, nu, mu, 2, nu, 0);
    const uint32_t n_m0 = U<D>(links, G, s, nu, mu, -1, nu, 0);
    const uint32_t n_1M = U<D>(links, G, s, nu, mu, 1, nu, -2);
    const uint32_t n_0M = U<D>(links, G, s, nu, mu, 0, nu, -2);
    const uint32_t n_2m = U<D>(links, G, s, nu, mu, 2, nu, -1);
    const uint32_t n_mm = U<D>(links, G, s, nu, mu, -1, nu, -1);
    // U_mu at x + a mu + b nu
    const uint32_t m_01 = U<D>(links, G, s, mu, mu, 0, nu, 1);
    const uint32_t m_0m = U<D>(links, G, s, mu, mu, 0, nu, -1);
    const uint32_t m_02 = U<D>(links, G, s, mu, mu, 0, nu, 2);
    const uint32_t m_11 = U<D>(links, G, s, mu, mu, 1, nu, 1);
    const uint32_t m_m1 = U<D>(links, G, s, mu, mu, -1, nu, 1);
    const uint32_t m_0M = U<D>(links, G, s, mu, mu, 0, nu, -2);
    const uint32_t m_1m = U<D>(links, G, s, mu, mu, 1, nu, -1);
    const uint32_t m_mm = U<D>(links, G, s, mu, mu, -1, nu, -1);

    // plaquette up: U_nu(x+mu) U_mu(x+nu)^-1 U_nu(x)^-1
    st[2 * j] = T.M(T.M(n_10, T.I(m_01)), T.I(n_00));
    // plaquette down: U_nu(x+mu-nu)^-1 U_mu(x-nu)^-1 U_nu(x-nu)
    st[2 * j + 1] = T.M(T.M(T.I(n_1m), T.I(m_0m)), n_0m);
    // k0
    re[6 * j + 0] = T.M(T.M(T.M(T.M(n_10, n_11), T.I(m_02)), T.I(n_01)), T.I(n_00));
    // k1
    re[6 * j + 1] = T.M(T.M(T.M(T.M(umu_p, n_20), T.I(m_11)), T.I(m_01)), T.I(n_00));
    // k2
    re[6 * j + 2] = T.M(T.M(T.M(T.M(n_10, T.I(m_01)), T.I(m_m1)), T.I(n_m0)), umu_m);
    // k3
    re[6 * j + 3] = T.M(T.M(T.M(T.M(T.I(n_1m), T.I(n_1M)), T.I(m_0M)), n_0M), n_0m);
    // k4
    re[6 * j + 4] = T.M(T.M(T.M(T.M(umu_p, T.I(n_2m)), T.I(m_1m)), T.I(m_0m)), n_0m);
    // k5
    re[6 * j + 5] = T.M(T.M(T.M(T.M(T.I(n_1m), T.I(m_0m)), T.I(m_mm)), n_mm), umu_m);
  }
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// One pass: every listed link of direction mu gets NHIT Metropolis hits
// (dym-mod-metro.cpp:158-192).  Links of one pass share no plaquette and no
// rectangle, so their updates are independent.
template <int D>
__global__ __launch_bounds__(kBlock) void k_sweep(SweepArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t P = A.T.P, npool = A.T.npool;
  double *s_retr = reinterpret_cast<double *>(smem);
  uint16_t *s_inv = reinterpret_cast<uint16_t *>(s_retr + P);
  uint16_t *s_pool = s_inv + P;
  for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) {
    s_retr[t] = A.T.retr[t];
    s_inv[t] = A.T.inv[t];
  }
  for (uint32_t t = threadIdx.x; t < npool; t += blockDim.x) s_pool[t] = A.T.pool[t];
  __syncthreads();

  const TabView T{A.T.mult, s_inv, P};
  const Geom &G = A.G;
  const int mu = A.mu;
  const double npool_d = (double)npool;
  constexpr int NS = 2 * (D - 1), NR = 6 * (D - 1);
  unsigned long long acc = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = xcd_block() * blockDim.x + threadIdx.x; k < A.count; k += stride) {
    const uint32_t i = A.sites[k];
    const Site<D> s = decode<D>(G, i);
    uint32_t st[NS], re[NR];
    build_staples<D>(A.links, G, T, s, mu, st, re);
    const size_t lp = (size_t)mu * G.V + lpos(G, i, s.par);
    uint32_t b = A.links[lp];
    const uint32_t b0 = b;
    double r1 = 0, m1 = 0;
#pragma unroll
    for (int j = 0; j < NS; ++j) r1 += s_retr[T.M(b, st[j])];
#pragma unroll
    for (int q = 0; q < NR; ++q) m1 += s_retr[T.M(b, re[q])];
    uint32_t ms = 0;
    if (A.rng == kRngMinstd) ms = A.mstate[i];
    const uint32_t link = i * (uint32_t)D + (uint32_t)mu;
    for (int h = 0; h < kNHit; ++h) {
      double u1, u2;
      if (A.rng == kRngPhilox) {
        const U4 w = philox4x32_10(U4{link, (uint32_t)h, A.sweep_lo, A.sweep_hi}, A.seed, kKeySweep);
        u1 = u53(w.x, w.y);
        u2 = u53(w.z, w.w);
      } else {
        u1 = minstd_canonical(ms);
        u2 = minstd_canonical(ms);  // same draw order as the reference: proposal, then acceptance
      }
      const uint32_t bnew = T.M(b, s_pool[(uint32_t)(npool_d * u1)]);
      double r1n = 0, m1n = 0;
#pragma unroll
      for (int j = 0; j < NS; ++j) r1n += s_retr[T.M(bnew, st[j])];
#pragma unroll
      for (int q = 0; q < NR; ++q) m1n += s_retr[T.M(bnew, re[q])];
      const double oldact = A.beta0 + A.beta1 * r1 + A.beta2 * m1;
      const double newact = A.beta0 + A.beta1 * r1n + A.beta2 * m1n;
      const double probrat = (newact <= oldact) ? 1.0 : dsa_exp(-(newact - oldact));
      if (probrat > u2) {
        b = bnew;
        r1 = r1n;
        m1 = m1n;
        ++acc;
      }
    }
    if (b != b0) A.links[lp] = (uint16_t)b;
    if (A.rng == kRngMinstd) A.mstate[i] = ms;
  }
  acc = wave_sum_u64(acc);
  if ((threadIdx.x & 63u) == 0 && acc) atomicAdd(A.acc, acc);
}

__global__ void k_init_philox(uint16_t *links, Geom G, uint32_t P, uint32_t seed) {
  const uint32_t nl = G.V * (uint32_t)G.D;
  for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < nl; l += gridDim.x * blockDim.x) {
    const U4 w = philox4x32_10(U4{l, 0u, 0u, 0u}, seed, kKeyInit);
    const uint32_t e = (uint32_t)((double)P * u53(w.x, w.y));
    const uint32_t i = l / (uint32_t)G.D, mu = l - i * (uint32_t)G.D;
    uint32_t par = 0, r = i;
    for (int d = 0; d < G.D; ++d) {
      const uint32_t q = r / G.ext[d];
      par += r - q * G.ext[d];
      r = q;
    }
    links[(size_t)mu * G.V + lpos(G, i, par & 1u)] = (uint16_t)e;
  }
}

__global__ void k_fill(uint16_t *links, uint32_t n, uint16_t v) {
  for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < n; l += gridDim.x * blockDim.x) links[l] = v;
}

__device__ inline uint32_t site_parity(const Geom &G, uint32_t i) {
  uint32_t par = 0;
  for (int d = 0; d < G.D; ++d) {
    const uint32_t q = i / G.ext[d];
    par += i - q * G.ext[d];
    i = q;
  }
  return par & 1u;
}

__global__ void k_pack(uint16_t *links, const int32_t *aos, Geom G) {
  const uint32_t nl = G.V * (uint32_t)G.D;
  for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < nl; l += gridDim.x * blockDim.x) {
    const uint32_t i = l / (uint32_t)G.D, mu = l - i * (uint32_t)G.D;
    links[(size_t)mu * G.V + lpos(G, i, site_parity(G, i))] = (uint16_t)aos[l];
  }
}

__global__ void k_unpack(int32_t *aos, const uint16_t *links, Geom G) {
  const uint32_t nl = G.V * (uint32_t)G.D;
  for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < nl; l += gridDim.x * blockDim.x) {
    const uint32_t i = l / (uint32_t)G.D, mu = l - i * (uint32_t)G.D;
    aos[l] = links[(size_t)mu * G.V + lpos(G, i, site_parity(G, i))];
  }
}

template <int D>
__global__ void k_staples(int32_t *st_out, int32_t *re_out, const uint16_t *links, Tables Tb, Geom G) {
  constexpr int NS = 2 * (D - 1), NR = 6 * (D - 1);
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  const uint32_t nl = G.V * (uint32_t)D;
  for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < nl; l += gridDim.x * blockDim.x) {
    const uint32_t i = l / D;
    const int mu = (int)(l - i * D);
    const Site<D> s = decode<D>(G, i);
    uint32_t st[NS], re[NR];
    build_staples<D>(links, G, T, s, mu, st, re);
    for (int j = 0; j < NS; ++j) st_out[(size_t)l * NS + j] = (int32_t)st[j];
    for (int q = 0; q < NR; ++q) re_out[(size_t)l * NR + q] = (int32_t)re[q];
  }
}

// ---------------------------------------------------------------- loops

__device__ inline uint32_t link_at(const uint16_t *links, const Geom &G, uint32_t site, int dir) {
  return links[(size_t)dir * G.V + lpos(G, site, site_parity(G, site))];
}

__device__ inline uint32_t step_site(const Geom &G, uint32_t i, int d, int sgn) {
  const uint32_t n = G.ext[d], st = G.stride[d];
  const uint32_t x = (i / st) % n;
  const uint32_t y = sgn > 0 ? (x + 1 == n ? 0 : x + 1) : (x == 0 ? n - 1 : x - 1);
  return i + (y - x) * st;
}

// lattice.c:76-83
__device__ inline uint32_t plaq_elem(const uint16_t *links, const Geom &G, const TabView &T, uint32_t i, int d1,
                                     int d2) {
  uint32_t g = link_at(links, G, i, d1);
  g = T.M(g, link_at(links, G, step_site(G, i, d1, 1), d2));
  g = T.M(g, T.I(link_at(links, G, step_site(G, i, d2, 1), d1)));
  g = T.M(g, T.I(link_at(links, G, i, d2)));
  return g;
}

// wilson(i, d1, d2, 2, 1), lattice.c:116-124
__device__ inline uint32_t rect_elem(const uint16_t *links, const Geom &G, const TabView &T, uint32_t i, int d1,
                                     int d2) {
  const uint32_t x1 = step_site(G, i, d1, 1), x2 = step_site(G, x1, d1, 1);
  const uint32_t x2u = step_site(G, x2, d2, 1);
  const uint32_t x1u = step_site(G, x2u, d1, -1), x0u = step_site(G, x1u, d1, -1);
  uint32_t g = link_at(links, G, i, d1);
  g = T.M(g, link_at(links, G, x1, d1));
  g = T.M(g, link_at(links, G, x2, d2));
  g = T.M(g, T.I(link_at(links, G, x1u, d1)));
  g = T.M(g, T.I(link_at(links, G, x0u, d1)));
  g = T.M(g, T.I(link_at(links, G, i, d2)));
  return g;
}

// lattice.c:86-94 (site i has x0 = 0)
__device__ inline uint32_t poly_elem(const uint16_t *links, const Geom &G, const TabView &T, uint32_t i) {
  uint32_t g = link_at(links, G, i, 0);
  uint32_t j = i;
  for (uint32_t t = 1; t < G.ext[0]; ++t) {
    j = step_site(G, j, 0, 1);
    g = T.M(g, link_at(links, G, j, 0));
  }
  return g;
}

__device__ inline size_t line_row(int d, uint32_t n, uint32_t nmax_t, uint32_t nmax_x) {
  return d == 0 ? (size_t)(n - 1) : (size_t)nmax_t + (size_t)(d - 1) * nmax_x + (n - 1);
}

// wilson(i, d1, 0, K, T) = L_d1(x,K) L_0(x+K d1,T) L_d1(x+T 0,K)^-1 L_0(x,T)^-1
// (the same element as the reference's link-by-link product: the table is a group).
__device__ inline uint32_t wilson_elem(const uint16_t *lines, const Geom &G, const TabView &T, uint32_t i, int d1,
                                       uint32_t K, uint32_t Tt, uint32_t nmax_t, uint32_t nmax_x) {
  const size_t V = G.V;
  const uint32_t xd1 = (i / G.stride[d1]) % G.ext[d1];
  const uint32_t x0 = i % G.ext[0];
  const uint32_t yd1 = (xd1 + K) % G.ext[d1];
  const uint32_t y0 = (x0 + Tt) % G.ext[0];
  const uint32_t iK = i + (yd1 - xd1) * G.stride[d1];
  const uint32_t iT = i + (y0 - x0);
  const uint32_t a = lines[line_row(d1, K, nmax_t, nmax_x) * V + i];
  const uint32_t b = lines[line_row(0, Tt, nmax_t, nmax_x) * V + iK];
  const uint32_t c = lines[line_row(d1, K, nmax_t, nmax_x) * V + iT];
  const uint32_t e = lines[line_row(0, Tt, nmax_t, nmax_x) * V + i];
  return T.M(T.M(T.M(a, b), T.I(c)), T.I(e));
}

__global__ void k_lines(uint16_t *lines, const uint16_t *links, Tables Tb, Geom G, uint32_t nmax_t,
                        uint32_t nmax_x) {
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < G.V; i += gridDim.x * blockDim.x) {
    for (int d = 0; d < G.D; ++d) {
      const uint32_t nmax = d == 0 ? nmax_t : nmax_x;
      if (!nmax) continue;
      uint32_t g = link_at(links, G, i, d);
      uint32_t j = i;
      lines[line_row(d, 1, nmax_t, nmax_x) * G.V + i] = (uint16_t)g;
      for (uint32_t n = 2; n <= nmax; ++n) {
        j = step_site(G, j, d, 1);
        g = T.M(g, link_at(links, G, j, d));
        lines[line_row(d, n, nmax_t, nmax_x) * G.V + i] = (uint16_t)g;
      }
    }
  }
}

__global__ void k_loops(int32_t *out, int kind, int k, int l, const uint16_t *links, const uint16_t *lines,
                        Tables Tb, Geom G, uint32_t nmax_t, uint32_t nmax_x) {
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  const int D = G.D;
  const uint32_t nsite = kind == kLoopPoly ? G.V / G.ext[0] : G.V;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nsite; q += gridDim.x * blockDim.x) {
    if (kind == kLoopPlaq) {
      int p = 0;
      const int np = D * (D - 1) / 2;
      for (int d1 = 0; d1 < D; ++d1)
        for (int d2 = d1 + 1; d2 < D; ++d2) out[(size_t)q * np + p++] = (int32_t)plaq_elem(links, G, T, q, d1, d2);
    } else if (kind == kLoopRect) {
      int p = 0;
      const int np = D * (D - 1);
      for (int d1 = 0; d1 < D; ++d1)
        for (int d2 = 0; d2 < D; ++d2)
          if (d1 != d2) out[(size_t)q * np + p++] = (int32_t)rect_elem(links, G, T, q, d1, d2);
    } else if (kind == kLoopPoly) {
      out[q] = (int32_t)poly_elem(links, G, T, q * G.ext[0]);
    } else {
      for (int d1 = 1; d1 < D; ++d1)
        out[(size_t)q * (D - 1) + (d1 - 1)] =
            (int32_t)wilson_elem(lines, G, T, q, d1, (uint32_t)k + 1, (uint32_t)l + 1, nmax_t, nmax_x);
    }
  }
}

// ------------------------------------------------------------ reductions

// Fixed-shape block sum: wave64 shuffle tree, then the waves in order.
__device__ inline double block_sum(double v, double *s_part) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) s_part[w] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) r += s_part[q];
  __syncthreads();
  return r;
}

constexpr uint32_t kObsMaxBlocks = 1024;

__global__ __launch_bounds__(kBlock) void k_obs_local(double *partials, const uint16_t *links, Tables Tb, Geom G) {
  __shared__ double s_part[kBlock / 64];
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  const int D = G.D;
  double pre = 0, pim = 0, pl = 0, rc = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < G.V; i += gridDim.x * blockDim.x) {
    if (i % G.ext[0] == 0) {
      const uint32_t p = poly_elem(links, G, T, i);
      pre += Tb.retr[p];
      pim += Tb.imtr[p];
    }
    for (int d1 = 0; d1 < D; ++d1)
      for (int d2 = d1 + 1; d2 < D; ++d2) pl += Tb.retr[plaq_elem(links, G, T, i, d1, d2)];
    for (int d1 = 0; d1 < D; ++d1)
      for (int d2 = 0; d2 < D; ++d2)
        if (d1 != d2) rc += Tb.retr[rect_elem(links, G, T, i, d1, d2)];
  }
  const double a = block_sum(pre, s_part);
  const double b = block_sum(pim, s_part);
  const double c = block_sum(pl, s_part);
  const double d = block_sum(rc, s_part);
  if (threadIdx.x == 0) {
    partials[(size_t)0 * gridDim.x + blockIdx.x] = a;
    partials[(size_t)1 * gridDim.x + blockIdx.x] = b;
    partials[(size_t)2 * gridDim.x + blockIdx.x] = c;
    partials[(size_t)3 * gridDim.x + blockIdx.x] = d;
  }
}

__global__ __launch_bounds__(kBlock) void k_obs_wilson(double *partials, const uint16_t *lines, Tables Tb, Geom G,
                                                       uint32_t wl, uint32_t nmax_t, uint32_t nmax_x) {
  __shared__ double s_part[kBlock / 64];
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  const uint32_t kl = blockIdx.y;
  const uint32_t K = kl / wl + 1, Tt = kl % wl + 1;
  double w = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < G.V; i += gridDim.x * blockDim.x)
    for (int d1 = 1; d1 < G.D; ++d1) w += Tb.retr[wilson_elem(lines, G, T, i, d1, K, Tt, nmax_t, nmax_x)];
  const double r = block_sum(w, s_part);
  if (threadIdx.x == 0) partials[(size_t)(4 + kl) * gridDim.x + blockIdx.x] = r;
}

// Normalisations of dym-mod-metro.cpp:291,303,317 and lattice.c:110-111, in
// the reference's unsigned integer arithmetic.
__device__ inline double obs_norm(const Geom &G, uint32_t q) {
  const uint32_t V = G.V, D = (uint32_t)G.D;
  if (q < 2) return (double)(V / G.ext[0]);
  if (q == 2) return (double)(V * D * (D - 1) / 2);
  if (q == 3) return (double)(V * D * (D - 1));
  return (double)(V * (D - 1));
}

__global__ void k_obs_final(double *out, const double *partials, Geom G, uint32_t nobs, uint32_t nb) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nobs) return;
  double s = 0;
  for (uint32_t b = 0; b < nb; ++b) s += partials[(size_t)q * nb + b];
  out[q] = s / obs_norm(G, q);
}

// One lane, the reference's summation order at one thread.
__global__ void k_obs_ordered(double *out, const uint16_t *links, const uint16_t *lines, Tables Tb, Geom G,
                              uint32_t wk, uint32_t wl, uint32_t nmax_t, uint32_t nmax_x) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  const int D = G.D;
  double lre = 0, lim = 0;
  for (uint32_t i = 0; i < G.V; i += G.ext[0]) {
    const uint32_t p = poly_elem(links, G, T, i);
    lre += Tb.retr[p];
    lim += Tb.imtr[p];
  }
  out[0] = lre / obs_norm(G, 0);
  out[1] = lim / obs_norm(G, 1);
  double sp = 0;
  for (uint32_t i = 0; i < G.V; ++i)
    for (int d1 = 0; d1 < D; ++d1)
      for (int d2 = d1 + 1; d2 < D; ++d2) sp += Tb.retr[plaq_elem(links, G, T, i, d1, d2)];
  out[2] = sp / obs_norm(G, 2);
  double sr = 0;
  for (uint32_t i = 0; i < G.V; ++i)
    for (int d1 = 0; d1 < D; ++d1)
      for (int d2 = 0; d2 < D; ++d2)
        if (d1 != d2) sr += Tb.retr[rect_elem(links, G, T, i, d1, d2)];
  out[3] = sr / obs_norm(G, 3);
  for (uint32_t k = 0; k < wk; ++k)
    for (uint32_t l = 0; l < wl; ++l) {
      double w = 0;
      for (uint32_t i = 0; i < G.V; ++i)
        for (int d1 = 1; d1 < D; ++d1) w += Tb.retr[wilson_elem(lines, G, T, i, d1, k + 1, l + 1, nmax_t, nmax_x)];
      out[4 + k * wl + l] = w / obs_norm(G, 4);
    }
}

inline uint32_t grid_for(uint64_t n, uint32_t cap = 8192) {
  uint64_t b = (n + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (uint32_t)b;
}

}  // namespace

size_t sweep_lds_bytes(const Tables &T) {
  return (size_t)T.P * sizeof(double) + (size_t)T.P * sizeof(uint16_t) + (size_t)T.npool * sizeof(uint16_t);
}

hipError_t launch_sweep(const SweepArgs &A, int max_blocks, hipStream_t s) {
  if (A.count == 0) return hipSuccess;
  uint32_t nb = (A.count + kBlock - 1) / kBlock;
  if (max_blocks > 0 && nb > (uint32_t)max_blocks) nb = (uint32_t)max_blocks;
  const size_t lds = sweep_lds_bytes(A.T);
  switch (A.G.D) {
    case 2: hipLaunchKernelGGL(k_sweep<2>, dim3(nb), dim3(kBlock), lds, s, A); break;
    case 3: hipLaunchKernelGGL(k_sweep<3>, dim3(nb), dim3(kBlock), lds, s, A); break;
    case 4: hipLaunchKernelGGL(k_sweep<4>, dim3(nb), dim3(kBlock), lds, s, A); break;
    case 5: hipLaunchKernelGGL(k_sweep<5>, dim3(nb), dim3(kBlock), lds, s, A); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_init_philox(uint16_t *links, const Geom &G, uint32_t P, uint32_t seed, hipStream_t s) {
  hipLaunchKernelGGL(k_init_philox, dim3(grid_for((uint64_t)G.V * G.D)), dim3(kBlock), 0, s, links, G, P, seed);
  return hipGetLastError();
}

hipError_t launch_fill(uint16_t *links, const Geom &G, uint16_t v, hipStream_t s) {
  const uint32_t n = G.V * (uint32_t)G.D;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, s, links, n, v);
  return hipGetLastError();
}

hipError_t launch_pack(uint16_t *links, const int32_t *aos, const Geom &G, hipStream_t s) {
  hipLaunchKernelGGL(k_pack, dim3(grid_for((uint64_t)G.V * G.D)), dim3(kBlock), 0, s, links, aos, G);
  return hipGetLastError();
}

hipError_t launch_unpack(int32_t *aos, const uint16_t *links, const Geom &G, hipStream_t s) {
  hipLaunchKernelGGL(k_unpack, dim3(grid_for((uint64_t)G.V * G.D)), dim3(kBlock), 0, s, aos, links, G);
  return hipGetLastError();
}

hipError_t launch_staples(int32_t *st, int32_t *re, const uint16_t *links, const Tables &T, const Geom &G,
                          hipStream_t s) {
  const dim3 g(grid_for((uint64_t)G.V * G.D)), b(kBlock);
  switch (G.D) {
    case 2: hipLaunchKernelGGL(k_staples<2>, g, b, 0, s, st, re, links, T, G); break;
    case 3: hipLaunchKernelGGL(k_staples<3>, g, b, 0, s, st, re, links, T, G); break;
    case 4: hipLaunchKernelGGL(k_staples<4>, g, b, 0, s, st, re, links, T, G); break;
    case 5: hipLaunchKernelGGL(k_staples<5>, g, b, 0, s, st, re, links, T, G); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_loops(int32_t *out, int kind, int k, int l, const uint16_t *links, const uint16_t *lines,
                        const Tables &T, const Geom &G, uint32_t nmax_t, uint32_t nmax_x, hipStream_t s) {
  hipLaunchKernelGGL(k_loops, dim3(grid_for(G.V)), dim3(kBlock), 0, s, out, kind, k, l, links, lines, T, G,
                     nmax_t, nmax_x);
  return hipGetLastError();
}

hipError_t launch_lines(uint16_t *lines, const uint16_t *links, const Tables &T, const Geom &G, uint32_t nmax_t,
                        uint32_t nmax_x, hipStream_t s) {
  hipLaunchKernelGGL(k_lines, dim3(grid_for(G.V)), dim3(kBlock), 0, s, lines, links, T, G, nmax_t, nmax_x);
  return hipGetLastError();
}

uint32_t obs_blocks(const Geom &G) { return grid_for(G.V, kObsMaxBlocks); }

hipError_t launch_measure_tree(double *out, double *partials, const uint16_t *links, const uint16_t *lines,
                               const Tables &T, const Geom &G, uint32_t wk, uint32_t wl, uint32_t nmax_t,
                               uint32_t nmax_x, hipStream_t s) {
  const uint32_t nb = obs_blocks(G);
  hipLaunchKernelGGL(k_obs_local, dim3(nb), dim3(kBlock), 0, s, partials, links, T, G);
  if (wk && wl) hipLaunchKernelGGL(k_obs_wilson, dim3(nb, wk * wl), dim3(kBlock), 0, s, partials, lines, T, G, wl,
                                   nmax_t, nmax_x);
  const uint32_t nobs = 4 + wk * wl;
  hipLaunchKernelGGL(k_obs_final, dim3((nobs + 63) / 64), dim3(64), 0, s, out, partials, G, nobs, nb);
  return hipGetLastError();
}

hipError_t launch_measure_ordered(double *out, const uint16_t *links, const uint16_t *lines, const Tables &T,
                                  const Geom &G, uint32_t wk, uint32_t wl, uint32_t nmax_t, uint32_t nmax_x,
                                  hipStream_t s) {
  hipLaunchKernelGGL(k_obs_ordered, dim3(1), dim3(64), 0, s, out, links, lines, T, G, wk, wl, nmax_t, nmax_x);
  return hipGetLastError();
}

}  // namespace dsa
