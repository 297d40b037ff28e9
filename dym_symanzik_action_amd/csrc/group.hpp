// group.hpp -- finite-group tables as the Monte Carlo consumes them.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace dsa {

// Largest order whose P x P uint16 table the kernels address with 32-bit
// byte offsets (P^2 * 2 < 2^32); element indices then fit the uint16 links.
constexpr unsigned kMaxOrder = 46340;

struct Group {
  std::string path;
  unsigned P = 0;
  int id = -1;
  std::vector<double> retr, imtr;  // Re/Im of -Tr per element
  std::vector<uint16_t> mult;      // P*P, mult[n*P+m] = n*m
  std::vector<uint16_t> inv;       // P
  std::vector<uint16_t> pool;      // proposal elements (nearest the identity)
  // defining-representation (SU(3)) matrix of each element when the file
  // appends them: su3[18*e + part + 2*(c + 3*r)] (empty otherwise)
  std::vector<double> su3;
  double min_retr = 0, nn_retr = 0;
  // Conjugacy classes, for the class-product sweep kernel (empty when the
  // group has more than kMaxClasses classes or P > kMaxClassOrder).  rc[c]
  // is the ReTr of the lowest-index element of class c; dmax = max |retr[e] -
  // rc[class(e)]| (the table's ReTr is a class function up to rounding), rmax
  // = max |retr|; rcq[c] = rc[c] in units of 2^-24, rounded to nearest.
  // cls[e] = 4 * class(e), the byte offset of e's value in an int32 array
  // rcq, and cprod[j*P + t] = cls[pool[j] * t].  The wide form holds the
  // values themselves: cval[j*P + t] = rc[class(pool[j] * t)] and cval0[t] =
  // rc[class(t)] in units of 2^-13, rounded to nearest (|ReTr| < 4).
  std::vector<uint8_t> cls, cprod;
  std::vector<double> rc;
  std::vector<int32_t> rcq;
  // mdq[c]: max over elements t of class c and pool elements p of
  // |ReTr(p t) - ReTr(t)|, in units of 2^-24 rounded up: bounds how much one
  // hit can move a staple term whose staple-times-link product is in class c
  // (empty unless the pool is closed under conjugation, which makes the bound
  // hold for the reference's element values)
  std::vector<int32_t> mdq;
  std::vector<int16_t> cval, cval0;
  double dmax = 0, rmax = 0;
  // A faithful representation by 3x3 matrices over GF(4) (empty where none
  // was found): fmat[e] is e's matrix in the packed form of gf4.hpp, and
  // fkey[gf4_key(fmat[e])] = e (0xFFFF for keys of no element), so that
  // e f = fkey[gf4_key(gf4_mul3(fmat[e], fmat[f]))] for every pair.
  std::vector<uint32_t> fmat;
  std::vector<uint16_t> fkey;
};

constexpr unsigned kMaxClassOrder = 8192;
constexpr unsigned kMaxClasses = 63;

// Parses a groups/ table file (plain text or gzip): P, P ReTr, P ImTr, P*P
// table; trailing tokens are ignored (group.c:13-47).  The identity is the
// first element whose row and column are the identity map, the inverse of n
// is the first m with n*m = m*n = identity (group.c:49-80), and the proposal
// pool is every element with min ReTr < ReTr < nn ReTr + 1e-6
// (dym-mod-metro.cpp:239-245).  Returns an empty string on success, else a
// message; never aborts.
std::string load_group(const char *path, Group &g);

// Fills cls, cprod, rc, dmax, rmax (left empty when not applicable).
void build_classes(Group &g);

// Fills fmat and fkey when the group embeds in SL(3,4) and is generated by
// two elements; every product of the table is checked against them, and
// both stay empty otherwise.
void build_field(Group &g);

}  // namespace dsa
