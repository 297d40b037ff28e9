// group.hpp -- finite-group tables as the Monte Carlo consumes them.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace dsa {

// Largest order whose element indices fit the uint16 link encoding.
constexpr unsigned kMaxOrder = 65535;

struct Group {
  std::string path;
  unsigned P = 0;
  int id = -1;
  std::vector<double> retr, imtr;  // Re/Im of -Tr per element
  std::vector<uint16_t> mult;      // P*P, mult[n*P+m] = n*m
  std::vector<uint16_t> inv;       // P
  std::vector<uint16_t> pool;      // proposal elements (nearest the identity)
  double min_retr = 0, nn_retr = 0;
};

// Parses a groups/ table file (plain text or gzip): P, P ReTr, P ImTr, P*P
// table; trailing tokens are ignored (group.c:13-47).  The identity is the
// first element whose row and column are the identity map, the inverse of n
// is the first m with n*m = m*n = identity (group.c:49-80), and the proposal
// pool is every element with min ReTr < ReTr < nn ReTr + 1e-6
// (dym-mod-metro.cpp:239-245).  Returns an empty string on success, else a
// message; never aborts.
std::string load_group(const char *path, Group &g);

}  // namespace dsa
