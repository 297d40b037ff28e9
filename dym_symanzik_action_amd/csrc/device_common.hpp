// device_common.hpp -- device helpers shared by the gfx950 kernel files:
// lattice geometry, link addressing, staple products, closed loops and the
// fixed-shape reductions.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "dsa_math.hpp"
#include "gf4.hpp"
#include "kernels.hpp"

namespace dsa {
namespace {

constexpr int kBlock = 256;

// The pointer as one into global memory, so that accesses through it are
// global (not flat) instructions.  Pointers read from device-side structs
// (the batch and LLR chains) are generic to the compiler, and a flat access
// counts against the LDS wait counter too: every wait for an LDS lookup
// would also wait for all the global loads in flight.  Only for pointers
// into global memory (flat and global addresses coincide there).
template <class T>
__device__ inline __attribute__((address_space(1))) T *gmem(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}

// Four dwords as one native vector: 16-byte LDS and global accesses.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// The uint16 at byte offset `off` of a global array: the load takes the
// base from scalar registers and one vector register of offset, with no
// 64-bit address arithmetic per load.
__device__ inline uint32_t u16_at(const uint16_t *base, uint32_t off) {
  using gchar = const __attribute__((address_space(1))) unsigned char;
  using gu16 = const __attribute__((address_space(1))) uint16_t;
  return *reinterpret_cast<gu16 *>(reinterpret_cast<gchar *>(gmem(base)) + off);
}

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

// A site: its index, parity and coordinates packed 12 bits each (extents
// below 4096), so that a coordinate along a run-time direction is a shift,
// not an index into a register array (which the compiler would place in
// scratch memory).
template <int D>
struct Site {
  uint64_t xs;
  uint32_t i;
  uint32_t par;
};

template <int D>
__device__ inline Site<D> decode(const Geom &G, uint32_t i) {
  Site<D> s;
  s.i = i;
  s.xs = 0;
  uint32_t par = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const uint32_t q = i / G.ext[d];
    const uint32_t x = i - q * G.ext[d];
    s.xs |= (uint64_t)x << (12 * d);
    par += x;
    i = q;
  }
  s.par = par & 1u;
  return s;
}

template <int D>
__device__ inline int coord(const Site<D> &s, int d) {
  return (int)((s.xs >> (12 * d)) & 0xfffu);
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

// Link U_dir(x + a e_mu + b e_nu).  A chain has fewer than 2^31 links, so
// the byte offset fits 32 bits.
template <int D>
__device__ inline uint32_t U(const uint16_t *links, const Geom &G, const Site<D> &s, int dir, int mu,
                             int a, int nu, int b) {
  uint32_t par;
  const uint32_t site = shifted<D>(G, s, mu, a, nu, b, &par);
  return u16_at(links, 2u * ((uint32_t)dir * G.V + lpos(G, site, par)));
}

// mult[a * P + b]: the order is at most 46340, so the byte offset a 2P + 2b
// fits 32 bits (and a < 2^24, 2P < 2^24 for the 24-bit multiply).
__device__ inline uint32_t table_at(const uint16_t *mult, uint32_t P, uint32_t a, uint32_t b) {
  return u16_at(mult, __umul24(a, 2u * P) + 2u * b);
}

struct TabView {
  const uint16_t *__restrict__ mult;
  const uint16_t *inv;  // LDS or global
  uint32_t P;
  __device__ inline uint32_t M(uint32_t a, uint32_t b) const { return table_at(mult, P, a, b); }
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

// t_k = s_k b for the staples s_k of build_staples (same order: for each nu,
// the up and down plaquette staples, then k0..k5), built directly with the
// products re-associated (the table is a group, so every association gives
// the same element): the suffixes U_nu(x)^-1 b, U_nu(x-nu) b and
// U_mu(x-mu) b and the prefixes U_nu(x+mu) U_mu(x+nu)^-1 and
// U_nu(x+mu-nu)^-1 U_mu(x-nu)^-1 are shared, and every element is at most
// three dependent table gathers after the link loads (85 gathers instead of
// 84 + 24 at D = 4).
template <int D>
__device__ inline void build_t(const uint16_t *links, const Geom &G, const TabView &T, const Site<D> &s, int mu,
                               uint32_t b, uint32_t *t) {
  constexpr int NS = 2 * (D - 1);
  const uint32_t umu_p = U<D>(links, G, s, mu, mu, 1, mu, 0);   // U_mu(x+mu)
  const uint32_t umu_m = U<D>(links, G, s, mu, mu, -1, mu, 0);  // U_mu(x-mu)
  const uint32_t C = T.M(umu_m, b);
#pragma unroll
  for (int j = 0; j < D - 1; ++j) {
    const int nu = j + (j >= mu ? 1 : 0);
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
    const uint32_t m_01 = U<D>(links, G, s, mu, mu, 0, nu, 1);
    const uint32_t m_0m = U<D>(links, G, s, mu, mu, 0, nu, -1);
    const uint32_t m_02 = U<D>(links, G, s, mu, mu, 0, nu, 2);
    const uint32_t m_11 = U<D>(links, G, s, mu, mu, 1, nu, 1);
    const uint32_t m_m1 = U<D>(links, G, s, mu, mu, -1, nu, 1);
    const uint32_t m_0M = U<D>(links, G, s, mu, mu, 0, nu, -2);
    const uint32_t m_1m = U<D>(links, G, s, mu, mu, 1, nu, -1);
    const uint32_t m_mm = U<D>(links, G, s, mu, mu, -1, nu, -1);
    // level 1: suffixes and pairs
    const uint32_t A = T.M(T.I(n_00), b);                // U_nu(x)^-1 b
    const uint32_t Bd = T.M(n_0m, b);                    // U_nu(x-nu) b
    const uint32_t X = T.M(n_10, T.I(m_01));             // plaquette-up prefix
    const uint32_t Y = T.M(T.I(n_1m), T.I(m_0m));        // plaquette-down prefix
    const uint32_t p0 = T.M(n_10, n_11), q0 = T.M(T.I(m_02), T.I(n_01));
    const uint32_t p1 = T.M(umu_p, n_20), q1 = T.M(T.I(m_11), T.I(m_01));
    const uint32_t q2 = T.M(T.I(m_m1), T.I(n_m0));
    const uint32_t p3 = T.M(T.I(n_1m), T.I(n_1M)), q3 = T.M(T.I(m_0M), n_0M);
    const uint32_t p4 = T.M(umu_p, T.I(n_2m)), q4 = T.M(T.I(m_1m), T.I(m_0m));
    const uint32_t q5 = T.M(T.I(m_mm), n_mm);
    // levels 2 and 3
    t[2 * j] = T.M(X, A);
    t[2 * j + 1] = T.M(Y, Bd);
    t[NS + 6 * j + 0] = T.M(T.M(p0, q0), A);
    t[NS + 6 * j + 1] = T.M(T.M(p1, q1), A);
    t[NS + 6 * j + 2] = T.M(T.M(X, q2), C);
    t[NS + 6 * j + 3] = T.M(T.M(p3, q3), Bd);
    t[NS + 6 * j + 4] = T.M(T.M(p4, q4), Bd);
    t[NS + 6 * j + 5] = T.M(T.M(Y, q5), C);
  }
}

// Group products from the GF(4) form (gf4.hpp) instead of the table: the
// packed matrix of each element and of its inverse in LDS (8 bytes per
// element), split into product factors in registers where used; an element
// is found from its matrix through the key table.
struct FieldView {
  const uint32_t *m;    // P
  const uint32_t *mi;   // P: matrix of the inverse
  const uint16_t *key;  // kGf4Keys
  __device__ inline Gf4L L(uint32_t e) const { return gf4_left(m[e]); }
  __device__ inline Gf4L Li(uint32_t e) const { return gf4_left(mi[e]); }
  __device__ inline Gf4R R(uint32_t e) const { return gf4_right(m[e]); }
  __device__ inline Gf4R Ri(uint32_t e) const { return gf4_right(mi[e]); }
  __device__ inline uint32_t E(uint32_t w) const { return key[gf4_key(w)]; }
};

// One transverse direction's products of build_t_field: j-th nu, the
// shared right factor Rb of b, U_mu(x+mu)'s left factor and C = U_mu(x-mu) b
// given; t8 = {t[2j], t[2j+1], t[NS+6j] .. t[NS+6j+5]}.
template <int D>
__device__ inline void build_t_field_dir(const uint16_t *links, const Geom &G, const FieldView &F, const Site<D> &s,
                                         int mu, int j, const Gf4R &Rb, const Gf4L &umu_p, const Gf4R &C,
                                         uint32_t *t8) {
  const int nu = j + (j >= mu ? 1 : 0);
  const Gf4L n_10 = F.L(U<D>(links, G, s, nu, mu, 1, nu, 0));
  const Gf4L n_00i = F.Li(U<D>(links, G, s, nu, mu, 0, nu, 0));
  const Gf4L n_1mi = F.Li(U<D>(links, G, s, nu, mu, 1, nu, -1));
  const Gf4L n_0m = F.L(U<D>(links, G, s, nu, mu, 0, nu, -1));
  const Gf4R n_11 = F.R(U<D>(links, G, s, nu, mu, 1, nu, 1));
  const Gf4R n_01i = F.Ri(U<D>(links, G, s, nu, mu, 0, nu, 1));
  const Gf4R n_20 = F.R(U<D>(links, G, s, nu, mu, 2, nu, 0));
  const Gf4R n_m0i = F.Ri(U<D>(links, G, s, nu, mu, -1, nu, 0));
  const Gf4R n_1Mi = F.Ri(U<D>(links, G, s, nu, mu, 1, nu, -2));
  const Gf4R n_0M = F.R(U<D>(links, G, s, nu, mu, 0, nu, -2));
  const Gf4R n_2mi = F.Ri(U<D>(links, G, s, nu, mu, 2, nu, -1));
  const Gf4R n_mm = F.R(U<D>(links, G, s, nu, mu, -1, nu, -1));
  const Gf4R m_01i = F.Ri(U<D>(links, G, s, mu, mu, 0, nu, 1));
  const Gf4R m_0mi = F.Ri(U<D>(links, G, s, mu, mu, 0, nu, -1));
  const Gf4L m_02i = F.Li(U<D>(links, G, s, mu, mu, 0, nu, 2));
  const Gf4L m_11i = F.Li(U<D>(links, G, s, mu, mu, 1, nu, 1));
  const Gf4L m_m1i = F.Li(U<D>(links, G, s, mu, mu, -1, nu, 1));
  const Gf4L m_0Mi = F.Li(U<D>(links, G, s, mu, mu, 0, nu, -2));
  const Gf4L m_1mi = F.Li(U<D>(links, G, s, mu, mu, 1, nu, -1));
  const Gf4L m_mmi = F.Li(U<D>(links, G, s, mu, mu, -1, nu, -1));
  // level 1: products of two links (A, Bd with b), right or left forms
  // as their later products need them
  const Gf4R A = gf4_right(gf4_mul(n_00i, Rb));         // U_nu(x)^-1 b
  const Gf4R Bd = gf4_right(gf4_mul(n_0m, Rb));         // U_nu(x-nu) b
  const uint32_t X = gf4_mul(n_10, m_01i);              // plaquette-up prefix
  const uint32_t Y = gf4_mul(n_1mi, m_0mi);             // plaquette-down prefix
  const Gf4L XL = gf4_left(X), YL = gf4_left(Y);
  const Gf4L p0 = gf4_left(gf4_mul(n_10, n_11)), p1 = gf4_left(gf4_mul(umu_p, n_20));
  const Gf4L p3 = gf4_left(gf4_mul(n_1mi, n_1Mi)), p4 = gf4_left(gf4_mul(umu_p, n_2mi));
  const Gf4R q0 = gf4_right(gf4_mul(m_02i, n_01i)), q1 = gf4_right(gf4_mul(m_11i, m_01i));
  const Gf4R q2 = gf4_right(gf4_mul(m_m1i, n_m0i)), q3 = gf4_right(gf4_mul(m_0Mi, n_0M));
  const Gf4R q4 = gf4_right(gf4_mul(m_1mi, m_0mi)), q5 = gf4_right(gf4_mul(m_mmi, n_mm));
  // levels 2 and 3
  t8[0] = F.E(gf4_mul(XL, A));
  t8[1] = F.E(gf4_mul(YL, Bd));
  t8[2] = F.E(gf4_mul(gf4_left(gf4_mul(p0, q0)), A));
  t8[3] = F.E(gf4_mul(gf4_left(gf4_mul(p1, q1)), A));
  t8[4] = F.E(gf4_mul(gf4_left(gf4_mul(XL, q2)), C));
  t8[5] = F.E(gf4_mul(gf4_left(gf4_mul(p3, q3)), Bd));
  t8[6] = F.E(gf4_mul(gf4_left(gf4_mul(p4, q4)), Bd));
  t8[7] = F.E(gf4_mul(gf4_left(gf4_mul(YL, q5)), C));
}

// build_t with every product a packed GF(4) matrix product (no memory
// access; the same 85 products).  Each link's matrix (or its inverse's) is
// read from LDS and split into the factor its one role needs (left or
// right); shared products are split once.  The loads stay where the
// compiler places them (near their uses): issuing all 62 first needs more
// than the 128 registers a lane has at four waves per SIMD, and the spills
// cost more than the waits they save.
template <int D>
__device__ inline void build_t_field(const uint16_t *links, const Geom &G, const FieldView &F, const Site<D> &s,
                                     int mu, uint32_t b, uint32_t *t) {
  constexpr int NS = 2 * (D - 1);
  const Gf4R Rb = F.R(b);
  const Gf4L umu_p = F.L(U<D>(links, G, s, mu, mu, 1, mu, 0));                       // U_mu(x+mu)
  const Gf4R C = gf4_right(gf4_mul(F.L(U<D>(links, G, s, mu, mu, -1, mu, 0)), Rb));   // U_mu(x-mu) b
#pragma unroll
  for (int j = 0; j < D - 1; ++j) {
    uint32_t t8[8];
    build_t_field_dir<D>(links, G, F, s, mu, j, Rb, umu_p, C, t8);
    t[2 * j] = t8[0];
    t[2 * j + 1] = t8[1];
#pragma unroll
    for (int k = 0; k < 6; ++k) t[NS + 6 * j + k] = t8[2 + k];
  }
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Per-chain constants of one link update.
struct LinkCtx {
  uint16_t *links;
  uint32_t *mstate;
  const double *s_retr;   // LDS
  const uint16_t *s_pool; // LDS
  double npool_d;
  uint32_t npool;
  double beta0, beta1, beta2;
  uint32_t seed, sweep_lo, sweep_hi;
  int rng;
};

// NHIT Metropolis hits on link (i, mu) (dym-mod-metro.cpp:158-192): staples
// built once from the current field, action contributions r1 (plaquettes) and
// m1 (rectangles) summed in the reference's order, proposals b -> b*g with g
// from the pool, acceptance min(1, exp(-dS)) against a fresh uniform.  Returns
// the number of accepted hits.
template <int D>
__device__ inline uint32_t update_link(const LinkCtx &C, const Geom &G, const TabView &T, int mu, uint32_t i) {
  constexpr int NS = 2 * (D - 1), NR = 6 * (D - 1);
  const Site<D> s = decode<D>(G, i);
  uint32_t st[NS], re[NR];
  build_staples<D>(C.links, G, T, s, mu, st, re);
  const size_t lp = (size_t)mu * G.V + lpos(G, i, s.par);
  uint32_t b = gmem(C.links)[lp];
  const uint32_t b0 = b;
  double r1 = 0, m1 = 0;
#pragma unroll
  for (int j = 0; j < NS; ++j) r1 += C.s_retr[T.M(b, st[j])];
#pragma unroll
  for (int q = 0; q < NR; ++q) m1 += C.s_retr[T.M(b, re[q])];
  uint32_t ms = 0;
  if (C.rng == kRngMinstd) ms = gmem(C.mstate)[i];
  const uint32_t link = i * (uint32_t)D + (uint32_t)mu;
  uint32_t acc = 0;
  U4 w{0, 0, 0, 0};
  for (int h = 0; h < kNHit; ++h) {
    double u2;
    uint32_t jp;
    if (C.rng == kRngPhilox) {
      if ((h & 1) == 0) w = philox4x32_10(U4{link, (uint32_t)h >> 1, C.sweep_lo, C.sweep_hi}, C.seed, kKeySweep);
      jp = hit_pool_index((h & 1) ? w.z : w.x, C.npool);
      u2 = hit_uniform((h & 1) ? w.w : w.y);
    } else {
      jp = (uint32_t)(C.npool_d * minstd_canonical(ms));
      u2 = minstd_canonical(ms);  // same draw order as the reference: proposal, then acceptance
    }
    const uint32_t bnew = T.M(b, C.s_pool[jp]);
    double r1n = 0, m1n = 0;
#pragma unroll
    for (int j = 0; j < NS; ++j) r1n += C.s_retr[T.M(bnew, st[j])];
#pragma unroll
    for (int q = 0; q < NR; ++q) m1n += C.s_retr[T.M(bnew, re[q])];
    const double oldact = C.beta0 + C.beta1 * r1 + C.beta2 * m1;
    const double newact = C.beta0 + C.beta1 * r1n + C.beta2 * m1n;
    const double probrat = (newact <= oldact) ? 1.0 : dsa_exp(-(newact - oldact));
    if (probrat > u2) {
      b = bnew;
      r1 = r1n;
      m1 = m1n;
      ++acc;
    }
  }
  if (b != b0) gmem(C.links)[lp] = (uint16_t)b;
  if (C.rng == kRngMinstd) gmem(C.mstate)[i] = ms;
  return acc;
}

// The Metropolis decision on an estimate dc of the action change dS with
// |dS - dc| <= M, certified against the exact one, without branches: 1
// accept, 0 reject, -1 undecided.  Above the band (dc > M): v_exp_f32 (2^x,
// ~1 ulp) on the float-rounded argument is within 5e-6 relative of exp(-dc)
// for dc < 60, so decide with 1.5e-5 of slack; and exp(-dS) < exp(-37) <
// 1e-15 <= u2 rejects.  Inside the band (|dc| <= M) exp(-dS) >= 1 - 2M
// accepts any smaller u2.  Rarely, the double exp settles what the float
// one leaves open.
__device__ inline int metro_certified(double dc, double M, double u2) {
  const double e = (double)__builtin_amdgcn_exp2f((float)(-fmin(fmax(dc, 0.0), 60.0) * 1.4426950408889634));
  const double w = 1.02 * M + 0x1p-16;
  const int above = u2 < e * (1.0 - w) ? 1 : (u2 >= e * (1.0 + w) ? 0 : -1);
  int dec = dc + M <= 0.0  ? 1
            : dc - M > 0.0 ? ((dc - M > 37.0 && u2 >= 1e-15) ? 0 : (dc < 60.0 ? above : -1))
                           : (u2 < 1.0 - (2.02 * M + 0x1p-46) ? 1 : -1);
  if (dec < 0 && dc - M > 0.0) {
    const double ed = dsa_exp(-dc);
    const double wd = 1.02 * M + 0x1p-46;
    dec = u2 < ed * (1.0 - wd) ? 1 : (u2 >= ed * (1.0 + wd) ? 0 : -1);
  }
  return dec;
}

// The same test in single precision, ahead of the double one: dc and M are
// floats whose band [dc - M, dc + M] holds the exact change; every bound
// keeps the double test's slack (2^-16 relative on the exponential covers
// its float argument, v_exp_f32 and u's rounding to float).  -1 where it
// cannot decide for sure.
__device__ inline int metro_certified_f(float dc, float M, float u) {
  if (dc + M < 0.f) return 1;
  const float lo = dc - M;
  if (lo > 0.f) {
    if (lo > 37.f) return u >= 2e-15f ? 0 : -1;  // exp(-37) < 1e-15
    if (dc >= 60.f) return -1;
    const float e = __builtin_amdgcn_exp2f(-dc * 1.44269504f);
    const float w = 1.02f * M + 0x1p-16f;
    return u < e * (1.f - w) ? 1 : (u >= e * (1.f + w) ? 0 : -1);
  }
  return u < 1.f - (2.02f * M + 0x1p-22f) ? 1 : -1;
}

// A certain rejection from a lower bound: the exact action change is at least
// dl - M.  True when exp(-(dl - M)) <= u2 for sure, by the float tier of
// metro_certified (the same slack); false leaves the hit to the full test.
__device__ inline bool metro_rejects(double dl, double M, double u2) {
  if (dl - M <= 0.0) return false;
  if (dl - M > 37.0 && u2 >= 1e-15) return true;
  if (dl >= 60.0) return false;
  const double e = (double)__builtin_amdgcn_exp2f((float)(-dl * 1.4426950408889634));
  return u2 >= e * (1.0 + (1.02 * M + 0x1p-16));
}

// The reference's decision of one hit (dym-mod-metro.cpp:176-188) from
// element values, for proposal b -> b p given t_k = s_k b: s_k = t_k b^-1,
// so ReTr(b s_k) = retr[b t_k b^-1] and ReTr(b' s_k) = retr[b' t_k b^-1],
// summed in the reference's order.  1 accept, 0 reject.
template <int D>
__device__ inline int exact_hit(const LinkCtx &C, const double *retr, const TabView &T, uint32_t b, uint32_t p,
                                const uint32_t *t, double u2) {
  constexpr int NS = 2 * (D - 1), NL = 8 * (D - 1);
  const uint32_t bn = T.M(b, p), bi = T.I(b);
  double e1 = 0, e2 = 0, e1n = 0, e2n = 0;
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const uint32_t tk = t[k];
    const double v = gmem(retr)[T.M(T.M(b, tk), bi)], vn = gmem(retr)[T.M(T.M(bn, tk), bi)];
    if (k < NS) {
      e1 += v;
      e1n += vn;
    } else {
      e2 += v;
      e2n += vn;
    }
  }
  const double oldact = C.beta0 + C.beta1 * e1 + C.beta2 * e2;
  const double newact = C.beta0 + C.beta1 * e1n + C.beta2 * e2n;
  const double probrat = (newact <= oldact) ? 1.0 : dsa_exp(-(newact - oldact));
  return probrat > u2 ? 1 : 0;
}

// The hit's products t_k by value, for the out-of-line decision below.
template <int D>
struct TermSet {
  uint32_t v[8 * (D - 1)];
};

// A hit the float tier left open: the double test on the class-valued
// change dc (band M), else the element values (exact_hit).  Out of line: it
// runs for a few hits in 10^7, and inline its registers and code weigh on
// every hit.  Returns the decision (bit 0) and whether the element values
// made it (bit 1).
template <int D>
__device__ __attribute__((noinline)) int decide_open(const double *retr, const uint16_t *mult, const uint16_t *inv,
                                                     uint32_t P, double beta0, double beta1, double beta2, double dc,
                                                     double M, uint32_t b, uint32_t p, TermSet<D> ts, double u2) {
  const int dec = metro_certified(dc, M, u2);
  if (dec >= 0) return dec;
  const TabView T{mult, inv, P};
  const LinkCtx C{nullptr, nullptr, nullptr, nullptr, 0.0, 0u, beta0, beta1, beta2, 0u, 0u, 0u, 0};
  return exact_hit<D>(C, retr, T, b, p, ts.v, u2) | 2;
}

// Class-product form of update_link, with identical results.
//
// ReTr is a class function, so the trial term ReTr(b' s_k) with b' = b p_j
// equals ReTr(p_j t_k) for t_k = s_k b (conjugate by b): the class of p_j t_k
// is looked up in an LDS table instead of an L2 gather, and t_k changes only
// when a hit is accepted (t_k <- t_k p_j).  Class sums are exact integers of
// quantised class values (ClsNarrow: 2^-24 units through a class index;
// ClsWide: 2^-13 units held in the table itself, one lookup per term).  They
// differ from the reference's element-value sums by at most the table's
// within-class spread plus quantisation and rounding, bounded by `margin` for
// the change of action; a hit whose decision could differ within that bound
// is decided again from element values exactly as update_link does, so the
// links and counters are those of update_link.
struct ClsNarrow {
  const uint8_t *s_cprod;  // LDS, npool x P: 4 x class of pool[j] * t
  const uint8_t *s_cls;    // LDS, P: 4 x class
  const int32_t *s_rcq;    // LDS (static array), class ReTr in units of 2^-24
  // LDS: per class, the most one hit can change a term whose product is in
  // that class (Group::mdq, units of 2^-24); read by the BOUND update only
  const int32_t *s_md = nullptr;
  static constexpr double kScale = 0x1p-24;
  static constexpr bool kBound = true;
  using Row = const uint8_t *;
  __device__ Row row(uint32_t jp, uint32_t P) const { return s_cprod + __umul24(jp, P); }  // both < 2^24
  __device__ Row row0() const { return s_cls; }
  static constexpr bool kTwoLevel = true;
  __device__ int32_t val(Row r, uint32_t t) const { return at(idx(r, t)); }
  // the two levels of val: the class byte, then its value
  __device__ uint32_t idx(Row r, uint32_t t) const { return r[t]; }
  __device__ int32_t at(uint32_t off) const {
    return *reinterpret_cast<const int32_t *>(reinterpret_cast<const uint8_t *>(s_rcq) + off);
  }
  // the class value of t and its bound, from one class-index read
  __device__ int32_t val_md(uint32_t t, int32_t &md) const {
    const uint32_t off = s_cls[t];
    md = *reinterpret_cast<const int32_t *>(reinterpret_cast<const uint8_t *>(s_md) + off);
    return *reinterpret_cast<const int32_t *>(reinterpret_cast<const uint8_t *>(s_rcq) + off);
  }
};

struct ClsWide {
  const int16_t *s_cval;   // LDS, npool x P: class value of pool[j] * t in units of 2^-13
  const int16_t *s_cval0;  // LDS, P: class value of t
  static constexpr double kScale = 0x1p-13;
  static constexpr bool kBound = false;
  static constexpr bool kTwoLevel = false;
  using Row = const int16_t *;
  __device__ Row row(uint32_t jp, uint32_t P) const { return s_cval + __umul24(jp, P); }
  __device__ Row row0() const { return s_cval0; }
  __device__ int32_t val(Row r, uint32_t t) const { return r[t]; }
};

struct ClsCtx {
  const double *retr;  // global, element ReTr for exact decisions
  double margin;
};

// Uniforms drawn ahead by the block (Philox streams): hit h of the block's
// link l has proposal index jp[h * stride + l] and acceptance uniform
// u2[h * stride + l].
struct PreDraws {
  const uint16_t *jp;
  const double *u2;
  uint32_t stride, l;
};

// One link's state between the two halves of the class update: its
// staple-times-link products t_k = s_k b (plaquettes first) with their class
// sums, built by cls_prepare; its hits, made by cls_hits.
template <int D>
struct ClsLink {
  static constexpr int NS = 2 * (D - 1), NL = 8 * (D - 1);
  Site<D> s;
  size_t lp;
  uint32_t i, b, b0, ms;
  uint32_t t[NL];
  int32_t r1, m1;
  int32_t rb;  // sum over the rectangle terms of their classes' bounds (2^-24; 0 without bounds)
};

// Class sums of the plaquette (k < NS) and rectangle terms of t through `row`.
template <int D, class L>
__device__ inline void cls_sums(const L &Lk, typename L::Row row, const uint32_t *t, int32_t &r, int32_t &m) {
  constexpr int NS = 2 * (D - 1), NL = 8 * (D - 1);
  int32_t a = 0, c = 0;
  if constexpr (L::kTwoLevel) {
    // every class byte first, then the values: two LDS round trips per hit
    // instead of a chain of short ones
    uint32_t off[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) off[k] = Lk.idx(row, t[k]);
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int32_t v = Lk.at(off[k]);
      if (k < NS) a += v; else c += v;
    }
  } else {
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int32_t v = Lk.val(row, t[k]);
      if (k < NS) a += v; else c += v;
    }
  }
  r = a;
  m = c;
}

// Class sums of t through row 0 (r1, m1) and, with BOUND (the narrow form,
// its bounds in LDS), the rectangle terms' bound sum rb.
template <int D, bool BOUND, class L>
__device__ inline void cls_sums_bound(const L &Lk, ClsLink<D> &S) {
  constexpr int NS = 2 * (D - 1), NL = 8 * (D - 1);
  if constexpr (BOUND && L::kBound) {
    int32_t a = 0, c = 0, e = 0;
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      int32_t md;
      const int32_t v = Lk.val_md(S.t[k], md);
      if (k < NS) {
        a += v;
      } else {
        c += v;
        e += md;
      }
    }
    S.r1 = a;
    S.m1 = c;
    S.rb = e;
  } else {
    cls_sums<D>(Lk, Lk.row0(), S.t, S.r1, S.m1);
    S.rb = 0;
  }
}

// Class sum of the plaquette terms (k < NS) or of the rectangle terms of t.
template <int D, bool PLAQ, class L>
__device__ inline int32_t cls_part(const L &Lk, typename L::Row row, const uint32_t *t) {
  constexpr int NS = 2 * (D - 1), NL = 8 * (D - 1);
  int32_t a = 0;
#pragma unroll
  for (int k = PLAQ ? 0 : NS; k < (PLAQ ? NS : NL); ++k) a += Lk.val(row, t[k]);
  return a;
}

template <int D, bool BOUND, class L>
__device__ inline void cls_prepare(ClsLink<D> &S, const LinkCtx &C, const L &Lk, const Geom &G, const TabView &T,
                                   int mu, uint32_t i) {
  S.i = i;
  S.s = decode<D>(G, i);
  S.lp = (size_t)mu * G.V + lpos(G, i, S.s.par);
  S.b = S.b0 = gmem(C.links)[S.lp];
  build_t<D>(C.links, G, T, S.s, mu, S.b, S.t);
  cls_sums_bound<D, BOUND>(Lk, S);
  S.ms = C.rng == kRngMinstd ? gmem(C.mstate)[i] : 0u;
}

// cls_prepare with the products from the GF(4) form.
template <int D, bool BOUND, class L>
__device__ inline void cls_prepare_field(ClsLink<D> &S, const LinkCtx &C, const L &Lk, const Geom &G,
                                         const FieldView &F, int mu, uint32_t i) {
  S.i = i;
  S.s = decode<D>(G, i);
  S.lp = (size_t)mu * G.V + lpos(G, i, S.s.par);
  S.b = S.b0 = gmem(C.links)[S.lp];
  build_t_field<D>(C.links, G, F, S.s, mu, S.b, S.t);
  cls_sums_bound<D, BOUND>(Lk, S);
  S.ms = C.rng == kRngMinstd ? gmem(C.mstate)[i] : 0u;
}

// BOUND (the narrow form only): each hit is first tested on its plaquette
// terms with the rectangle terms bounded by |beta2| rb; a hit rejected even
// at the least action change that allows is rejected without the rectangle
// lookups.  It pays only where nearly every lane of a wavefront is rejected
// that way (a wavefront runs the rectangle terms if any lane needs them),
// which is why it is a separate instantiation chosen at very low acceptance.
//
// OOL: each hit decided in floats first, the double test and the element
// values (a few hits in 10^7) in the out-of-line decide_open.  It pays in
// the field sweep kernel (81.4 against 84.1 us per 32^3 x 64 pass); in the
// kernels that also carry the element-table path (batch) the call's saved
// registers spill, so they keep the double test inline.
template <int D, bool BOUND, class L, bool OOL = false>
__device__ inline uint32_t cls_hits(ClsLink<D> &S, const LinkCtx &C, const L &Lk, const ClsCtx &K, const TabView &T,
                                    int mu, uint32_t &nfall, const PreDraws *pre) {
  constexpr int NL = ClsLink<D>::NL;
  constexpr bool bounded = BOUND && L::kBound;
  const uint32_t link = S.i * (uint32_t)D + (uint32_t)mu;
  const double b1q = C.beta1 * L::kScale, b2q = C.beta2 * L::kScale, margin = K.margin;
  const double ab2q = fabs(C.beta2) * L::kScale;
  // the float tier's weights and margin (rounded up)
  const float b1f = (float)b1q, b2f = (float)b2q, mf = (float)margin * (1.f + 0x1p-22f);
  uint32_t acc = 0;
  U4 w{0, 0, 0, 0};
  for (int h = 0; h < kNHit; ++h) {
    double u2;  // the acceptance uniform; uf: the same rounded to float
    float uf;
    uint32_t jp;
    if (pre) {
      jp = pre->jp[h * pre->stride + pre->l];
      u2 = pre->u2[h * pre->stride + pre->l];
      uf = (float)u2;
    } else if (C.rng == kRngPhilox) {
      if ((h & 1) == 0) w = philox4x32_10(U4{link, (uint32_t)h >> 1, C.sweep_lo, C.sweep_hi}, C.seed, kKeySweep);
      jp = hit_pool_index((h & 1) ? w.z : w.x, C.npool);
      const uint32_t uw = (h & 1) ? w.w : w.y;
      u2 = hit_uniform(uw);
      uf = (float)uw * 0x1p-32f;
    } else {
      jp = (uint32_t)(C.npool_d * minstd_canonical(S.ms));
      u2 = minstd_canonical(S.ms);
      uf = (float)u2;
    }
    const typename L::Row row = Lk.row(jp, T.P);
    int32_t r1n, m1n;
    if constexpr (bounded) {
      // plaquette terms first: the rectangle terms change the action by at
      // most |beta2| rb, so the exact change is at least dp - bx - M; a hit
      // rejected even at that change is rejected without the rectangle terms
      r1n = cls_part<D, true>(Lk, row, S.t);
      const double dp = b1q * ((double)r1n - (double)S.r1), bx = ab2q * (double)S.rb;
      const double dl = dp - bx, Ml = margin + 0x1p-50 * (fabs(dp) + bx);
      if (metro_rejects(dl, Ml, u2)) continue;
      m1n = cls_part<D, false>(Lk, row, S.t);
    } else {
      cls_sums<D>(Lk, row, S.t, r1n, m1n);
    }
    // class-valued change of action; the exact one lies within margin of it
    int dec;
    const uint32_t p = C.s_pool[jp];
    if constexpr (OOL) {
      // In floats first: with exact integer differences d1, d2 (unless
      // they overflow), p1 + p2 is within 2^-22 (|p1| + |p2|) of it, so the
      // band [dcf - Mf, dcf + Mf] holds the exact change
      const int32_t d1 = (int32_t)((uint32_t)r1n - (uint32_t)S.r1), d2 = (int32_t)((uint32_t)m1n - (uint32_t)S.m1);
      const bool ovf = (((r1n ^ S.r1) & (r1n ^ d1)) | ((m1n ^ S.m1) & (m1n ^ d2))) < 0;
      const float p1 = b1f * (float)d1, p2 = b2f * (float)d2;
      dec = ovf ? -1 : metro_certified_f(p1 + p2, mf + 0x1p-21f * (fabsf(p1) + fabsf(p2)), uf);
      if (dec < 0) {
        const double dc = b1q * ((double)r1n - (double)S.r1) + b2q * ((double)m1n - (double)S.m1);
        const double M = margin + 0x1p-50 * fabs(dc);
        TermSet<D> ts;
#pragma unroll
        for (int k = 0; k < NL; ++k) ts.v[k] = S.t[k];
        const int r = decide_open<D>(K.retr, T.mult, T.inv, T.P, C.beta0, C.beta1, C.beta2, dc, M, S.b, p, ts, u2);
        dec = r & 1;
        nfall += (uint32_t)r >> 1;
      }
    } else {
      const double dc = b1q * ((double)r1n - (double)S.r1) + b2q * ((double)m1n - (double)S.m1);
      const double M = margin + 0x1p-50 * fabs(dc);
      dec = metro_certified(dc, M, u2);
      if (dec < 0) {
        dec = exact_hit<D>(C, K.retr, T, S.b, p, S.t, u2);
        ++nfall;
      }
    }
    if (dec) {
      S.b = T.M(S.b, p);
#pragma unroll
      for (int k = 0; k < NL; ++k) S.t[k] = T.M(S.t[k], p);
      if constexpr (bounded) {
        cls_sums_bound<D, true>(Lk, S);  // the new products' classes and bounds
      } else {
        S.r1 = r1n;
        S.m1 = m1n;
      }
      ++acc;
    }
  }
  if (S.b != S.b0) gmem(C.links)[S.lp] = (uint16_t)S.b;
  if (C.rng == kRngMinstd) gmem(C.mstate)[S.i] = S.ms;
  return acc;
}

template <int D, bool BOUND = false, class L>
__device__ inline uint32_t update_link_cls(const LinkCtx &C, const L &Lk, const ClsCtx &K, const Geom &G,
                                           const TabView &T, int mu, uint32_t i, uint32_t &nfall,
                                           const PreDraws *pre = nullptr) {
  ClsLink<D> S;
  cls_prepare<D, BOUND>(S, C, Lk, G, T, mu, i);
  return cls_hits<D, BOUND>(S, C, Lk, K, T, mu, nfall, pre);
}

// Matrix form of the hit loop.  rho(g) is the SU(3) matrix the group file
// appends to element g (ReTr(g) = -Re Tr rho(g), rho(a) rho(b) = rho(ab);
// both to ~1e-15, checked on the host).  The reference's action of link b is
// beta0 + sum_k w_k ReTr(b s_k) (w = beta1 for plaquettes, beta2 for
// rectangles) = beta0 - Re Tr M with M = sum_k w_k rho(t_k), t_k = s_k b (a
// trace is cyclic), so a proposal b -> b p changes it by
// dS = Re Tr M - Re Tr(rho(p) M), and an accepted p takes M to M rho(p).  The
// t_k come from table gathers exactly as in the class kernel, M from their
// matrices in LDS (24 FP32 axpys); each hit then costs one matrix read and 18
// FMAs instead of 24 class lookups, and an accepted hit one gather and a 3x3
// product instead of 25 gathers.  dS carries a rounding error below the
// host's bound (margin0, plus margin_acc per accepted hit), and a hit whose
// decision that error could change is decided from element values by
// exact_hit: the links and counters are update_link's.

struct Cm3 {
  float re[9], im[9];  // row-major entries
};

// rho(e) from LDS: 20 floats per element (re[9], im[9], 2 unused).
__device__ inline Cm3 mat_of(const float *s_mat, uint32_t e) {
  const float4 *q = reinterpret_cast<const float4 *>(s_mat + 20 * (size_t)e);
  const float4 a = q[0], b = q[1], c = q[2], d = q[3], f = q[4];
  return Cm3{{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x}, {c.y, c.z, c.w, d.x, d.y, d.z, d.w, f.x, f.y}};
}

// A <- A B, one row at a time (six temporaries instead of a third matrix);
// each entry's real and imaginary part a chain of six FMAs.
__device__ inline void mul_right(Cm3 &A, const Cm3 &B) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float r[3], m[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      r[j] = 0.f;
      m[j] = 0.f;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float ar = A.re[3 * i + k], ai = A.im[3 * i + k], br = B.re[3 * k + j], bi = B.im[3 * k + j];
        r[j] = fmaf(ar, br, r[j]);
        r[j] = fmaf(-ai, bi, r[j]);
        m[j] = fmaf(ar, bi, m[j]);
        m[j] = fmaf(ai, br, m[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      A.re[3 * i + j] = r[j];
      A.im[3 * i + j] = m[j];
    }
  }
}

// Y += a X
__device__ inline void axpy(Cm3 &Y, float a, const Cm3 &X) {
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    Y.re[q] = fmaf(a, X.re[q], Y.re[q]);
    Y.im[q] = fmaf(a, X.im[q], Y.im[q]);
  }
}

__device__ inline Cm3 scaled(float a, const Cm3 &X) {
  Cm3 Y;
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    Y.re[q] = a * X.re[q];
    Y.im[q] = a * X.im[q];
  }
  return Y;
}

__device__ inline void addto(Cm3 &Y, const Cm3 &X) {
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    Y.re[q] += X.re[q];
    Y.im[q] += X.im[q];
  }
}

// Re Tr(A B), a chain of 18 FMAs.
__device__ inline float retr_prod(const Cm3 &A, const Cm3 &B) {
  float t = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      t = fmaf(A.re[3 * i + j], B.re[3 * j + i], t);
      t = fmaf(-A.im[3 * i + j], B.im[3 * j + i], t);
    }
  return t;
}

struct MatCtx {
  const float *s_mat;         // LDS, P x 20 floats
  const double *retr;         // global, element ReTr for exact decisions
  double margin0, margin_acc; // |dS - dc| before any accepted hit, and its growth per accepted hit
};

template <int D>
__device__ inline uint32_t update_link_mat(const LinkCtx &C, const MatCtx &K, const Geom &G, const TabView &T, int mu,
                                           uint32_t i, uint32_t &nfall) {
  constexpr int NS = 2 * (D - 1), NL = 8 * (D - 1);
  const Site<D> s = decode<D>(G, i);
  const size_t lp = (size_t)mu * G.V + lpos(G, i, s.par);
  uint32_t b = gmem(C.links)[lp];
  const uint32_t b0 = b;
  uint32_t ms = 0;
  if (C.rng == kRngMinstd) ms = gmem(C.mstate)[i];
  const uint32_t link = i * (uint32_t)D + (uint32_t)mu;
  // hit h's proposal index and acceptance uniform, in stream order
  U4 w{0, 0, 0, 0};
  auto draw = [&](int h, uint32_t &jp, double &u2) {
    if (C.rng == kRngPhilox) {
      if ((h & 1) == 0) w = philox4x32_10(U4{link, (uint32_t)h >> 1, C.sweep_lo, C.sweep_hi}, C.seed, kKeySweep);
      jp = hit_pool_index((h & 1) ? w.z : w.x, C.npool);
      u2 = hit_uniform((h & 1) ? w.w : w.y);
    } else {
      jp = (uint32_t)(C.npool_d * minstd_canonical(ms));
      u2 = minstd_canonical(ms);
    }
  };
  uint32_t acc = 0;
  {
    Cm3 M;
    {
      uint32_t t[NL];
      build_t<D>(C.links, G, T, s, mu, b, t);
      // every gather issued before the first matrix is read; then one
      // matrix (20 registers) at a time
      __builtin_amdgcn_sched_barrier(0);
      const float w1 = (float)C.beta1, w2 = (float)C.beta2;
      M = scaled(w1, mat_of(K.s_mat, t[0]));
#pragma unroll
      for (int k = 1; k < NL; ++k) {
        __builtin_amdgcn_sched_barrier(0);
        axpy(M, k < NS ? w1 : w2, mat_of(K.s_mat, t[k]));
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    float v0 = M.re[0] + M.re[4] + M.re[8];
    double margin = K.margin0;
    for (int h = 0; h < kNHit; ++h) {
      uint32_t jp;
      double u2;
      draw(h, jp, u2);
      const uint32_t p = C.s_pool[jp];
      // dS = Re Tr M - Re Tr(rho(p) M); the difference of two floats is exact
      // in double.  rho(p) is read again for an accepted hit rather than
      // kept live across the element-value path.
      const double dc = (double)v0 - (double)retr_prod(mat_of(K.s_mat, p), M);
      int dec = metro_certified(dc, margin + 0x1p-50 * fabs(dc), u2);
      if (dec < 0) {
        // rare: this hit from element values (t_k = s_k b as the class
        // kernel builds them); M goes on estimating the later hits.  The
        // site is decoded again from an opaque copy of i: otherwise the
        // compiler keeps the first build_t's link addresses live through
        // the hit loop for this path (spilling them on the common one).
        uint32_t i2 = i;
        asm volatile("" : "+v"(i2));
        const Site<D> s2 = decode<D>(G, i2);
        uint32_t t[NL];
        build_t<D>(C.links, G, T, s2, mu, b, t);
        dec = exact_hit<D>(C, K.retr, T, b, p, t, u2);
        ++nfall;
      }
      if (dec) {
        b = T.M(b, p);
        mul_right(M, mat_of(K.s_mat, p));
        v0 = M.re[0] + M.re[4] + M.re[8];
        margin += K.margin_acc;
        ++acc;
      }
    }
  }
  if (b != b0) gmem(C.links)[lp] = (uint16_t)b;
  if (C.rng == kRngMinstd) gmem(C.mstate)[i] = ms;
  return acc;
}

// Block-wide staging of a table into LDS in two halves: load() issues a
// lane's global loads for its first K elements (t = lane + j * blockDim),
// store() writes them and copies any elements beyond K * blockDim with a
// plain loop.  Calling every table's load() before any store() puts a whole
// block's staging behind one round trip to L2; a load-store loop waits on
// each load before the store of the same iteration (one round trip per
// iteration, ~18 for the S(1080) class tables at 512 lanes).
// 16-byte unit of the staged copies: a native vector type, which the
// compiler keeps in registers (arrays of HIP's uint4 struct land in scratch).

template <int K, typename T>
struct Staged {
  T r[K];
  __device__ inline void load(const T *__restrict__ src, uint32_t n) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const uint32_t t = threadIdx.x + (uint32_t)j * blockDim.x;
      if (t < n) r[j] = gmem(src)[t];
    }
  }
  __device__ inline void store(T *dst, const T *__restrict__ src, uint32_t n) const {
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const uint32_t t = threadIdx.x + (uint32_t)j * blockDim.x;
      if (t < n) dst[t] = r[j];
    }
    for (uint32_t t = threadIdx.x + (uint32_t)K * blockDim.x; t < n; t += blockDim.x) dst[t] = gmem(src)[t];
  }
};

// The class tables of the narrow form (byte class of every element, the
// npool x P byte table of pool[j] * t, the class values), the inverse table
// and the pool, staged together.  cprod is read in 16-byte units up to its
// 16-byte padding.
struct ClsNarrowStage {
  Staged<3, uint16_t> inv;
  Staged<1, uint16_t> pool;
  Staged<3, uint8_t> cls;
  Staged<1, int32_t> rcq;
  Staged<1, int32_t> mdq;
  Staged<10, u32x4> tab;
  // s_md (64 entries) receives the per-class hit bounds where the group has them
  __device__ inline void run(const Tables &Tb, uint16_t *s_inv, uint16_t *s_pool, uint8_t *s_cls, int32_t *s_rcq,
                             uint8_t *s_tab, int32_t *s_md = nullptr) {
    const uint32_t n16 = (uint32_t)(pad16((size_t)Tb.npool * Tb.P) / 16);
    const u32x4 *src = reinterpret_cast<const u32x4 *>(Tb.cprod);
    const uint32_t nmd = s_md && Tb.mdq ? Tb.ncls : 0u;
    inv.load(Tb.inv, Tb.P);
    pool.load(Tb.pool, Tb.npool);
    cls.load(Tb.cls, Tb.P);
    rcq.load(Tb.rcq, Tb.ncls);
    mdq.load(Tb.mdq, nmd);
    tab.load(src, n16);
    inv.store(s_inv, Tb.inv, Tb.P);
    pool.store(s_pool, Tb.pool, Tb.npool);
    cls.store(s_cls, Tb.cls, Tb.P);
    rcq.store(s_rcq, Tb.rcq, Tb.ncls);
    mdq.store(s_md, Tb.mdq, nmd);
    tab.store(reinterpret_cast<u32x4 *>(s_tab), src, n16);
  }
};

// Loads ReTr, the inverse table and the proposal pool of one group into LDS.
__device__ inline void load_tables_lds(const Tables &Tb, double *s_retr, uint16_t *s_inv, uint16_t *s_pool) {
  Staged<5, double> retr;
  Staged<5, uint16_t> inv;
  Staged<2, uint16_t> pool;
  retr.load(Tb.retr, Tb.P);
  inv.load(Tb.inv, Tb.P);
  pool.load(Tb.pool, Tb.npool);
  retr.store(s_retr, Tb.retr, Tb.P);
  inv.store(s_inv, Tb.inv, Tb.P);
  pool.store(s_pool, Tb.pool, Tb.npool);
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

// ---------------------------------------------------------------- loops

__device__ inline uint32_t link_at(const uint16_t *links, const Geom &G, uint32_t site, int dir) {
  return gmem(links)[(size_t)dir * G.V + lpos(G, site, site_parity(G, site))];
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


// Normalisations of dym-mod-metro.cpp:291,303,317 and lattice.c:110-111, in
// the reference's unsigned integer arithmetic.
__device__ inline double obs_norm(const Geom &G, uint32_t q) {
  const uint32_t V = G.V, D = (uint32_t)G.D;
  if (q < 2) return (double)(V / G.ext[0]);
  if (q == 2) return (double)(V * D * (D - 1) / 2);
  if (q == 3) return (double)(V * D * (D - 1));
  return (double)(V * (D - 1));
}


}  // namespace
}  // namespace dsa
