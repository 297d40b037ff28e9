// internal.hpp -- state shared by the C ABI translation units (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <vector>
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
  uint8_t *cls = nullptr, *cprod = nullptr;  // class tables (null if the group has none)
  int32_t *rcq = nullptr, *mdq = nullptr;
  int16_t *cval = nullptr, *cval0 = nullptr;
  float *mat = nullptr;  // FP32 SU(3) matrices, 20 floats per element (null if none)
  uint32_t *fmat = nullptr, *finv = nullptr;  // GF(4) form (null if none)
  uint16_t *fkey = nullptr;
};

}  // namespace dsa

struct dsa_group {
  dsa::Group g;
  std::mutex m;
  // the matrix sweep kernel's FP32 matrices and the worst deviation of the
  // file's matrices from the tables (su3_check), computed on first upload
  int mat_state = 0;  // 0 unchecked, 1 usable, -1 not usable
  std::vector<float> mat32;
  double mat_dev = 0;
  std::map<int, dsa::DeviceTables> dev;  // one upload per device
};

namespace dsa {

// The group's tables on `device`, uploaded on first use.
int upload_tables(dsa_group *g, int device, DeviceTables &out);
Tables tables_view(const dsa_group *g, const DeviceTables &t);
// Bound on |dS_exact - dS_class| for one hit of the class-product kernel:
// both trial actions' class-value error (D-dependent staple counts, the
// table's within-class ReTr spread dmax) and the float rounding of both sums
// and of beta0 + beta1 r1 + beta2 m1, with a safety factor of 4.
// q: the quantisation of the class values (half their unit).
double class_margin(const Group &g, int D, double beta0, double beta1, double beta2, double q);
// The terms of class_margin per coupling: value spread and quantisation,
// sum rounding and magnitude of the plaquette (1) and rectangle (2) sums.
void class_margin_terms(const Group &g, int D, double q, double &e1, double &e2, double &f1, double &f2, double &r1,
                        double &r2);
// The matrix kernel's FP32 matrices (20 floats per element: re[9], im[9],
// 2 zero) and the largest deviation of the file's double matrices from the
// tables: |ReTr(g) + Re Tr rho(g)| over g and every entry of
// rho(a) rho(b) - rho(ab) over all pairs.  False when the file has none.
bool su3_check(const Group &g, std::vector<float> &mat32, double &dev);
// Bound on |dS_exact - dS_matrix| for the matrix kernel's first hit, and its
// growth per accepted hit (the FP32 rounding of the matrices, their weighted
// sum, the traces and the products M rho(p)).
void mat_margin(int D, double beta0, double beta1, double beta2, double dev, double &m0, double &macc);
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
