// dsa_math.hpp -- arithmetic shared by the host and the gfx950 kernels.
//
// Everything here is basic IEEE double arithmetic and 32/64-bit integer work,
// compiled with -ffp-contract=off on both sides, so host and device produce
// the same bits.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define DSA_HD __host__ __device__
#else
#define DSA_HD
#endif

namespace dsa {

constexpr int kNHit = 20;                    // dym-mod-metro.cpp:49
constexpr uint32_t kKeySweep = 0x44534131u;  // "DSA1": Metropolis hits
constexpr uint32_t kKeyInit = 0x44534149u;   // "DSAI": hot start
constexpr uint32_t kKeyLLR = 0x4453414cu;    // "DSAL": constrained hits

DSA_HD inline double bits_to_double(uint64_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __longlong_as_double((long long)u);
#else
  double d;
  __builtin_memcpy(&d, &u, 8);
  return d;
#endif
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
struct U4 {
  uint32_t x, y, z, w;
};

DSA_HD inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    U4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// 53-bit uniform in [0,1) from two 32-bit words.
DSA_HD inline double u53(uint32_t hi, uint32_t lo) {
  return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6)) * 0x1.0p-53;
}

// The Metropolis sweep's Philox stream: one Philox4x32-10 call, counter
// (link, h >> 1, sweep_lo, sweep_hi) and key (seed, kKeySweep), serves hits h
// and h + 1 (h even); hit h takes words 2 (h & 1) and 2 (h & 1) + 1 as its
// proposal index floor(npool w / 2^32) and its acceptance uniform w' / 2^32.
DSA_HD inline uint32_t hit_pool_index(uint32_t w, uint32_t npool) {
  return (uint32_t)(((uint64_t)w * npool) >> 32);
}
DSA_HD inline double hit_uniform(uint32_t w) { return (double)w * 0x1.0p-32; }
static_assert(kNHit % 2 == 0, "hits are drawn in pairs");

// std::minstd_rand0 step: x <- 16807 x mod (2^31-1), via the Mersenne fold.
DSA_HD inline uint32_t minstd_next(uint32_t &x) {
  const uint64_t p = (uint64_t)x * 16807u;
  uint32_t r = (uint32_t)(p & 0x7fffffffu) + (uint32_t)(p >> 31);
  if (r >= 0x7fffffffu) r -= 0x7fffffffu;
  x = r;
  return r;
}

// generate_canonical<double,53>(minstd_rand0) as libstdc++ 11 computes it:
// two draws, weights 1 and R = 2^31-2, divided by (double)(R*R) = 2^62-2^33.
DSA_HD inline double minstd_canonical(uint32_t &x) {
  double sum = 0.0;
  sum += (double)(minstd_next(x) - 1u) * 1.0;
  sum += (double)(minstd_next(x) - 1u) * 2147483646.0;
  double ret = sum / 4611686009837453312.0;
  if (ret >= 1.0) ret = 0x1.fffffffffffffp-1;
  return ret;
}

// exp(x): Cody-Waite reduction x = k ln2 + r with a 32-bit head of ln2 (k*hi
// exact for |k| < 2^21) and a tail; e^r by its Taylor series through r^13 in
// Horner order (truncation below 5e-18 relative for |r| <= ln2/2); 2^k put in
// through the exponent field, split in two factors near the range ends.
DSA_HD inline double dsa_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return __builtin_huge_val();
  if (x < -745.1332191019412) return 0.0;
  const double log2e = 0x1.71547652b82fep+0;
  const double ln2hi = 0x1.62e42fee00000p-1;
  const double ln2lo = 0x1.a39ef35793c76p-33;
  const double t = x * log2e;
  const int k = (int)(t < 0.0 ? t - 0.5 : t + 0.5);
  const double kd = (double)k;
  const double r = (x - kd * ln2hi) - kd * ln2lo;
  double p = 0x1.6124613a86d09p-33;
  p = p * r + 0x1.1eed8eff8d898p-29;
  p = p * r + 0x1.ae64567f544e4p-26;
  p = p * r + 0x1.27e4fb7789f5cp-22;
  p = p * r + 0x1.71de3a556c734p-19;
  p = p * r + 0x1.a01a01a01a01ap-16;
  p = p * r + 0x1.a01a01a01a01ap-13;
  p = p * r + 0x1.6c16c16c16c17p-10;
  p = p * r + 0x1.1111111111111p-7;
  p = p * r + 0x1.5555555555555p-5;
  p = p * r + 0x1.5555555555555p-3;
  p = p * r + 0x1.0000000000000p-1;
  const double er = 1.0 + (r + (r * r) * p);
  if (k > 1000)
    return (er * bits_to_double((uint64_t)(k - 1000 + 1023) << 52)) * bits_to_double((uint64_t)(2023) << 52);
  if (k < -1000)
    return (er * bits_to_double((uint64_t)(k + 1000 + 1023) << 52)) * bits_to_double((uint64_t)(23) << 52);
  return er * bits_to_double((uint64_t)(k + 1023) << 52);
}

// Energies of the LLR window are kept as integers in units of 2^-32, so the
// sum of a pass's changes does not depend on the order of the additions.
constexpr double kFixScale = 4294967296.0;
DSA_HD inline long long to_fix(double x) { return (long long)(x * kFixScale + (x >= 0.0 ? 0.5 : -0.5)); }
DSA_HD inline double from_fix(long long v) { return (double)v / kFixScale; }

}  // namespace dsa
