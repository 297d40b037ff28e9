// internal.hpp -- state shared by the C ABI translation units (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>

#include "../../include/dsa.h"
#include "group.hpp"
#include "kernels.hpp"

namespace dsa {

// Records the calling thread's error message; returns code.
int fail(int code, const std::string &msg);

struct DeviceTables {
  uint16_t *mult = nullptr, *inv = nullptr, *pool = nullptr;
  double *retr = nullptr, *imtr = nullptr;
};

}  // namespace dsa

struct dsa_group {
  dsa::Group g;
  std::mutex m;
  std::map<int, dsa::DeviceTables> dev;  // one upload per device
};

namespace dsa {

// The group's tables on `device`, uploaded on first use.
int upload_tables(dsa_group *g, int device, DeviceTables &out);
Tables tables_view(const dsa_group *g, const DeviceTables &t);
// Shape and option checks shared by chains and batches.
int check_params(const dsa_chain_params *p);
// minstd-compat start (dym-mod-metro.cpp:227-228, 253-255): per-site states and
// the hot links drawn from site 0's engine, in the reference's link order.
void minstd_start(const Group &g, const dsa_chain_params &p, uint32_t V, std::vector<uint32_t> &states,
                  std::vector<int32_t> &aos);

}  // namespace dsa

#define HIP_TRY(expr)                                                                                   \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess) return dsa::fail(DSA_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)
