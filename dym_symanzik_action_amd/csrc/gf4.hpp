// gf4.hpp -- 3x3 matrices over GF(4), packed for the group product of the
// sweep kernels (host and device share these functions bit for bit).
//
// GF(4) = {0, 1, w, w^2 = w + 1}; an entry e = e1 w + e0 splits into two
// GF(2) planes, lo (e0) and hi (e1), and a matrix A = Al + w Ah into two GF(2)
// matrices.  Then
//   A B = (Al Bl + Ah Bh) + w ((Al + Ah)(Bl + Bh) + Al Bl),
// three GF(2) matrix products (Karatsuba).  The packed word holds the three
// planes the products need, 9 bits each (entry (r, c) at bit 3r + c):
//   bits 0-8 lo plane, bits 10-18 lo ^ hi, bits 20-28 hi plane.
// One AND of a word whose entry (r, c) holds A[r][k] with one whose entry
// (r, c) holds B[k][c] forms the k-th outer-product term of all three GF(2)
// products at once.  Bits 9, 19 and 29-31 are don't-care: every consumer
// masks them.
#pragma once
#include <stdint.h>

#include "dsa_math.hpp"

namespace dsa {

constexpr uint32_t kGf4Col = 0x49u * (1u | 1u << 10 | 1u << 20);  // entries (r, 0) of each plane
constexpr uint32_t kGf4Row = 0x7u * (1u | 1u << 10 | 1u << 20);   // entries (0, c) of each plane
constexpr uint32_t kGf4Keys = 4096;                                 // first two rows: 6 entries x 2 bits

// A[r][k] across row r of every plane (the k-th term's left factor).
DSA_HD inline uint32_t gf4_col(uint32_t A, int k) {
  uint32_t a = (A >> k) & kGf4Col;  // A[r][k] at (r, 0)
#if defined(__HIP_DEVICE_COMPILE__)
  // hide a's known-zero bits: the compiler would otherwise turn the
  // replication into a * 7, a quarter-rate 32-bit multiply (a reaches bit
  // 26, past the full-rate 24-bit one); two shift-ors are full rate
  asm("" : "+v"(a));
#endif
  return a | (a << 1) | (a << 2);
}

// B[k][c] down column c of every plane (the k-th term's right factor).
DSA_HD inline uint32_t gf4_row(uint32_t B, int k) {
  const uint32_t b = (B >> (3 * k)) & kGf4Row;  // B[k][c] at (0, c)
  return b | (b << 3) | (b << 6);
}

// The product from the sum S of its three outer-product terms.  The planes
// of S are f0 = Al Bl, f1 = (Al + Ah)(Bl + Bh), f2 = Ah Bh; the product's lo
// plane is f0 + f2, its hi plane f0 + f1, and lo ^ hi = f1 + f2.
DSA_HD inline uint32_t gf4_fin(uint32_t S) {
  const uint32_t u = S ^ (S >> 10), v = S ^ (S >> 20);
  return (v & 0x1FFu) | (u & (0x1FFu << 10)) | (u << 20);
}

DSA_HD inline uint32_t gf4_mul3(uint32_t A, uint32_t B) {
  uint32_t S = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) S ^= gf4_col(A, k) & gf4_row(B, k);
  return gf4_fin(S);
}

// An operand with its three term factors formed ahead: as the left factor
// of products (columns) or as the right one (rows).  The sweep kernels read
// the links' matrices in these forms from tables.
struct Gf4L {
  uint32_t c[3];
};
struct Gf4R {
  uint32_t r[3];
};

DSA_HD inline Gf4L gf4_left(uint32_t A) { return Gf4L{{gf4_col(A, 0), gf4_col(A, 1), gf4_col(A, 2)}}; }
DSA_HD inline Gf4R gf4_right(uint32_t B) { return Gf4R{{gf4_row(B, 0), gf4_row(B, 1), gf4_row(B, 2)}}; }

DSA_HD inline uint32_t gf4_mul(const Gf4L &a, const Gf4R &b) {
  return gf4_fin((a.c[0] & b.r[0]) ^ (a.c[1] & b.r[1]) ^ (a.c[2] & b.r[2]));
}

// The first two rows of the packed matrix (lo and hi planes) as a 12-bit key.
DSA_HD inline uint32_t gf4_key(uint32_t W) { return (W & 0x3Fu) | ((W >> 14) & 0xFC0u); }

// Packed word of a matrix given as 2-bit entries, entry (r, c) at bits
// 2(3r + c) of `plain` (e = 2 e1 + e0).
DSA_HD inline uint32_t gf4_pack(uint32_t plain) {
  uint32_t w = 0;
  for (int q = 0; q < 9; ++q) {
    const uint32_t e = (plain >> (2 * q)) & 3u, lo = e & 1u, hi = e >> 1;
    w |= lo << q | (lo ^ hi) << (10 + q) | hi << (20 + q);
  }
  return w;
}

}  // namespace dsa
