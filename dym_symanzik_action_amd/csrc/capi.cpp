// capi.cpp -- the C ABI of libdsa.so (include/dsa.h): group and chain handles,
// device memory, and the sweep / measurement drivers.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <cmath>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dsa.h"
#include "dsa_math.hpp"
#include "gf4.hpp"
#include "group.hpp"
#include "internal.hpp"
#include "kernels.hpp"
#include "probe.hpp"
#include "schedule.hpp"

namespace {
thread_local std::string g_err;
}  // namespace

using dsa::DeviceTables;
using dsa::fail;

int dsa::fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

struct dsa_chain {
  dsa_group *grp = nullptr;
  DeviceTables dt;
  dsa::Tables T{};
  dsa::Geom G{};
  dsa_chain_params p{};
  dsa::Schedule S;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint16_t *d_links = nullptr;
  uint32_t *d_sites = nullptr;
  uint32_t *d_mstate = nullptr;
  unsigned long long *d_acc = nullptr;
  unsigned long long *d_fall = nullptr;  // re-decided hits (class kernel)
  int32_t *d_aos = nullptr;
  uint16_t *d_lines = nullptr;
  size_t lines_cap = 0;
  uint32_t lines_t = 0, lines_x = 0;
  bool lines_valid = false;
  double *d_partials = nullptr;
  size_t partials_cap = 0;
  double *d_out = nullptr;
  size_t out_cap = 0;
  uint64_t sweep = 0, hit = 0;
  double last_ms = 0;
  // the narrow class form's plaquette-first test for the next call, chosen
  // from the acceptance of the previous one (dsa::kBoundAcceptance)
  bool bound_next = false;
  unsigned long long acc_host = 0;
  int max_blocks = 2048;
  // sweep kernel: DSA_KERNEL_*, the class-kernel forms that can run here
  // (dsa::kCls*, kClsNone if neither), their decision margins
  int kernel = DSA_KERNEL_AUTO;
  int cls_best = dsa::kClsNone;
  bool narrow_ok = false;
  bool field_ok = false;  // the narrow form with GF(4) products fits here
  size_t lds_max = 0;  // LDS bytes per block on the chain's device
  uint32_t ncu = 0;    // its compute units
  double margin8 = 0, margin16 = 0, margin_scale = 1;
  // the matrix kernel: runs here (the group has matrices, they fit LDS),
  // its decision bound and the bound's growth per accepted hit
  bool mat_ok = false;
  double mat_m0 = 0, mat_macc = 0;
  // per-launch timing (measurement mode): one event pair per pass
  int profiling = 0;  // 0 off, 1 an event pair per pass launch, 2 one pair per dsa_chain_sweep call
  std::vector<hipEvent_t> pev;
  double kern_ms = 0;
  uint64_t kern_launches = 0;
};

dsa::Tables dsa::tables_view(const dsa_group *g, const DeviceTables &t) {
  dsa::Tables T{};
  T.mult = t.mult;
  T.inv = t.inv;
  T.retr = t.retr;
  T.imtr = t.imtr;
  T.pool = t.pool;
  T.P = g->g.P;
  T.npool = (uint32_t)g->g.pool.size();
  T.id = (uint32_t)g->g.id;
  T.cls = t.cls;
  T.cprod = t.cprod;
  T.rcq = t.rcq;
  T.mdq = t.mdq;
  T.ncls = (uint32_t)g->g.rc.size();
  T.cval = t.cval;
  T.cval0 = t.cval0;
  T.mat = t.mat;
  T.fmat = t.fmat;
  T.finv = t.finv;
  T.fkey = t.fkey;
  return T;
}

bool dsa::su3_check(const Group &g, std::vector<float> &mat32, double &dev) {
  const size_t P = g.P;
  if (g.su3.size() != 18 * P) return false;
  mat32.assign(20 * P, 0.0f);
  dev = 0;
  for (size_t e = 0; e < P; ++e) {
    double tr = 0;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        const double re = g.su3[18 * e + 2 * (c + 3 * r)], im = g.su3[18 * e + 2 * (c + 3 * r) + 1];
        mat32[20 * e + 3 * r + c] = (float)re;
        mat32[20 * e + 9 + 3 * r + c] = (float)im;
        if (r == c) tr += re;
      }
    dev = std::max(dev, std::fabs(g.retr[e] + tr));  // the table's ReTr is minus the real trace
  }
  // a representation: rho(a) rho(b) = rho(ab) entry by entry, every pair
  for (size_t a = 0; a < P; ++a) {
    const double *A = &g.su3[18 * a];
    for (size_t b = 0; b < P; ++b) {
      const double *B = &g.su3[18 * b], *Q = &g.su3[18 * (size_t)g.mult[a * P + b]];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double xr = 0, xi = 0;
          for (int k = 0; k < 3; ++k) {
            const double ar = A[2 * (k + 3 * r)], ai = A[2 * (k + 3 * r) + 1];
            const double br = B[2 * (c + 3 * k)], bi = B[2 * (c + 3 * k) + 1];
            xr += ar * br - ai * bi;
            xi += ar * bi + ai * br;
          }
          dev = std::max(dev, std::max(std::fabs(xr - Q[2 * (c + 3 * r)]), std::fabs(xi - Q[2 * (c + 3 * r) + 1])));
        }
    }
  }
  return true;
}

// The matrix kernel's bound.  With u = 2^-24: M = sum_k w_k rho(t_k) is
// accumulated with 24 FMAs per entry from matrices whose entries (|.| <= 1,
// unitary) are rounded from the file, with weights rounded to float, so each
// real and imaginary entry errs by <= 26 u W, W = sum |w_k| =
// 2(D-1)|beta1| + 6(D-1)|beta2|.  Re Tr M (three adds) then errs by <= 84 u W;
// Re Tr(rho(p) M), an 18-FMA chain over terms whose magnitudes sum to
// <= 3 sqrt(6) W, by <= 3 sqrt(6) 26 u W (M's error) + 132 u W (the chain)
// + 8 u W (rho(p)'s rounding): dS errs by < 420 u W.  Each accepted hit's
// product M rho(p) (six-FMA chains per entry) adds < 100 u W.  The bound below
// takes 1024 u W and 256 u W (safety >= 2.4), plus the file's own deviation
// from the tables (su3_check) and the rounding of the reference's double sums.
void dsa::mat_margin(int D, double beta0, double beta1, double beta2, double dev, double &m0, double &macc) {
  const double u = 0x1p-24, ud = 0x1p-53;
  const double W = 2.0 * (D - 1) * std::fabs(beta1) + 6.0 * (D - 1) * std::fabs(beta2);
  m0 = W * (1024.0 * u + 64.0 * dev + 1e-12) + 16.0 * ud * (std::fabs(beta0) + 3.0 * W);
  macc = W * 256.0 * u;
}

void dsa::class_margin_terms(const Group &g, int D, double q, double &e1, double &e2, double &f1, double &f2,
                             double &r1, double &r2) {
  const double u = 0x1p-53;
  const int ns = 2 * (D - 1), nr = 6 * (D - 1);
  auto gam = [u](int n) { return n * u / (1.0 - n * u); };
  // |r1_exact - r1_class| <= e1 + f1: value differences (exact bounds) plus
  // the rounding of the reference's float sum; likewise m1
  e1 = ns * (g.dmax + q);
  f1 = gam(ns) * ns * g.rmax;
  e2 = nr * (g.dmax + q);
  f2 = gam(nr) * nr * g.rmax;
  r1 = ns * g.rmax + e1;
  r2 = nr * g.rmax + e2;
}

double dsa::class_margin(const Group &g, int D, double beta0, double beta1, double beta2, double q) {
  const double u = 0x1p-53;
  double e1, e2, f1, f2, r1, r2;
  class_margin_terms(g, D, q, e1, e2, f1, f2, r1, r2);
  const double a = std::fabs(beta0) + std::fabs(beta1) * r1 + std::fabs(beta2) * r2;
  const double vals = std::fabs(beta1) * e1 + std::fabs(beta2) * e2;                   // per action
  const double rounding = std::fabs(beta1) * f1 + std::fabs(beta2) * f2 + 8.0 * u * a;  // per action
  // both actions of a hit; the rounding terms with a safety factor of 4
  return 2.0 * vals * (1.0 + 1e-6) + 4.0 * (2.0 * rounding);
}

int dsa::check_params(const dsa_chain_params *p) {
  if (p->D < 2 || p->D > (uint32_t)dsa::kDMax)
    return fail(DSA_E_RANGE, "D = " + std::to_string(p->D) + " outside the supported 2..5");
  if (p->Nt < 1 || p->Nx < 1) return fail(DSA_E_ARG, "Nt and Nx must be positive");
  if (p->Nt > 4095 || p->Nx > 4095) return fail(DSA_E_RANGE, "Nt and Nx must be below 4096");
  uint64_t V = p->Nt;
  for (uint32_t d = 1; d < p->D; ++d) V *= p->Nx;
  if (V * p->D >= (1ull << 31)) return fail(DSA_E_RANGE, "V*D must be below 2^31 links");
  if (p->rng != DSA_RNG_PHILOX && p->rng != DSA_RNG_MINSTD) return fail(DSA_E_ARG, "unknown rng");
  if (p->order != DSA_ORDER_COLOUR && p->order != DSA_ORDER_REFERENCE) return fail(DSA_E_ARG, "unknown order");
  if (p->start != DSA_START_HOT && p->start != DSA_START_COLD) return fail(DSA_E_ARG, "unknown start");
  return 0;
}

void dsa::minstd_start(const Group &g, const dsa_chain_params &p, uint32_t V, std::vector<uint32_t> &ms,
                       std::vector<int32_t> &a) {
  ms.resize(V);
  for (uint32_t i = 0; i < V; ++i) {
    const uint64_t s = (uint64_t)(int64_t)(int32_t)((uint32_t)p.seed + i);
    uint64_t x = s % 2147483647ull;
    ms[i] = (uint32_t)(x == 0 ? 1 : x);
  }
  const size_t nl = (size_t)V * p.D;
  a.assign(nl, g.id);
  if (p.start == DSA_START_COLD) return;
  // uniform_int_distribution(0, P-1) on minstd_rand0: two-division downscaling
  const uint64_t urngrange = 2147483645ull, uerange = g.P;
  const uint64_t scaling = urngrange / uerange, past = uerange * scaling;
  for (size_t l = 0; l < nl; ++l) {
    uint64_t r;
    do r = (uint64_t)dsa::minstd_next(ms[0]) - 1;
    while (r >= past);
    a[l] = (int32_t)(r / scaling);
  }
}

namespace {

void free_tables(DeviceTables &t) {
  for (void *p : {(void *)t.mult, (void *)t.inv, (void *)t.pool, (void *)t.retr, (void *)t.imtr, (void *)t.cls,
                  (void *)t.cprod, (void *)t.rcq, (void *)t.mdq, (void *)t.cval, (void *)t.cval0, (void *)t.mat,
                  (void *)t.fmat, (void *)t.finv, (void *)t.fkey})
    if (p) (void)hipFree(p);
  t = DeviceTables();
}

// Allocates and fills every device table of `G` in `t`; on failure the
// caller frees what was allocated.  Arrays the class kernels stage into LDS
// in 16-byte units (cprod, cval, cval0) are padded to 16 bytes with zeros.
int fill_tables(const dsa::Group &G, dsa_group *g, DeviceTables &t) {
  const size_t P = G.P;
  HIP_TRY(hipMalloc(&t.mult, P * P * sizeof(uint16_t)));
  HIP_TRY(hipMalloc(&t.inv, P * sizeof(uint16_t)));
  HIP_TRY(hipMalloc(&t.pool, G.pool.size() * sizeof(uint16_t)));
  HIP_TRY(hipMalloc(&t.retr, P * sizeof(double)));
  HIP_TRY(hipMalloc(&t.imtr, P * sizeof(double)));
  HIP_TRY(hipMemcpy(t.mult, G.mult.data(), P * P * sizeof(uint16_t), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(t.inv, G.inv.data(), P * sizeof(uint16_t), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(t.pool, G.pool.data(), G.pool.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(t.retr, G.retr.data(), P * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(t.imtr, G.imtr.data(), P * sizeof(double), hipMemcpyHostToDevice));
  if (!G.cprod.empty()) {
    const size_t nb = dsa::pad16(G.cprod.size());
    HIP_TRY(hipMalloc(&t.cls, P));
    HIP_TRY(hipMalloc(&t.cprod, nb));
    HIP_TRY(hipMalloc(&t.rcq, G.rcq.size() * sizeof(int32_t)));
    HIP_TRY(hipMemset(t.cprod, 0, nb));
    HIP_TRY(hipMemcpy(t.cls, G.cls.data(), P, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.cprod, G.cprod.data(), G.cprod.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.rcq, G.rcq.data(), G.rcq.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!G.mdq.empty()) {
      HIP_TRY(hipMalloc(&t.mdq, G.mdq.size() * sizeof(int32_t)));
      HIP_TRY(hipMemcpy(t.mdq, G.mdq.data(), G.mdq.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
  }
  if (!G.cval.empty()) {
    const size_t nb = dsa::pad16(G.cval.size() * sizeof(int16_t)), nb0 = dsa::pad16(P * sizeof(int16_t));
    HIP_TRY(hipMalloc(&t.cval, nb));
    HIP_TRY(hipMalloc(&t.cval0, nb0));
    HIP_TRY(hipMemset(t.cval, 0, nb));
    HIP_TRY(hipMemset(t.cval0, 0, nb0));
    HIP_TRY(hipMemcpy(t.cval, G.cval.data(), G.cval.size() * sizeof(int16_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.cval0, G.cval0.data(), P * sizeof(int16_t), hipMemcpyHostToDevice));
  }
  if (!G.fmat.empty()) {
    std::vector<uint32_t> fi(P);
    for (size_t e = 0; e < P; ++e) fi[e] = G.fmat[G.inv[e]];
    HIP_TRY(hipMalloc(&t.fmat, P * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&t.finv, P * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&t.fkey, G.fkey.size() * sizeof(uint16_t)));
    HIP_TRY(hipMemcpy(t.fmat, G.fmat.data(), P * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.finv, fi.data(), P * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.fkey, G.fkey.data(), G.fkey.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  }
  if (g->mat_state == 0) g->mat_state = dsa::su3_check(G, g->mat32, g->mat_dev) && g->mat_dev < 1e-9 ? 1 : -1;
  if (g->mat_state > 0) {
    HIP_TRY(hipMalloc(&t.mat, g->mat32.size() * sizeof(float)));
    HIP_TRY(hipMemcpy(t.mat, g->mat32.data(), g->mat32.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  return 0;
}

}  // namespace

int dsa::upload_tables(dsa_group *g, int device, DeviceTables &out) {
  std::lock_guard<std::mutex> lk(g->m);
  auto it = g->dev.find(device);
  if (it != g->dev.end()) {
    out = it->second;
    return 0;
  }
  DeviceTables t;
  if (int rc = fill_tables(g->g, g, t)) {
    free_tables(t);
    return rc;
  }
  g->dev[device] = t;
  out = t;
  return 0;
}

namespace {

// The class form a chain's large passes run (kClsNone: the table kernel).
// CLASS takes the int16-value form where it fits, else the byte-index one;
// AUTO prefers the byte-index form (half the LDS of the int16 one), which
// both certifies more decisions and measures faster on S(1080), with its
// products formed in the group's GF(4) form where the group has one.
int cls_mode(const dsa_chain *c) {
  switch (c->kernel) {
    case DSA_KERNEL_CLASS: return c->cls_best;
    case DSA_KERNEL_CLASS_NARROW: return dsa::kClsNarrow;
    case DSA_KERNEL_MATRIX: return dsa::kClsMatrix;
    case DSA_KERNEL_FIELD: return dsa::kClsField;
    case DSA_KERNEL_AUTO:
      if (c->narrow_ok && c->margin8 < 1e-3) return c->field_ok ? (int)dsa::kClsField : (int)dsa::kClsNarrow;
      return c->cls_best;
    default: return dsa::kClsNone;
  }
}

// The form the next sweep call runs: cls_mode, except that under AUTO a
// frozen chain (acceptance below dsa::kBoundAcceptance over the previous
// call, its links near one another) runs the class kernel with table
// gathers, which then hit the caches, and its plaquette-first test.
int run_mode(const dsa_chain *c) {
  const int m = cls_mode(c);
  return c->kernel == DSA_KERNEL_AUTO && c->bound_next && m == dsa::kClsField ? (int)dsa::kClsNarrow : m;
}

int ensure_lines(dsa_chain *c, uint32_t nt, uint32_t nx) {
  if (c->lines_valid && c->lines_t == nt && c->lines_x == nx) return 0;
  const size_t rows = (size_t)nt + (size_t)(c->G.D - 1) * nx;
  const size_t need = rows * c->G.V * sizeof(uint16_t);
  if (need > c->lines_cap) {
    if (c->d_lines) HIP_TRY(hipFree(c->d_lines));
    c->d_lines = nullptr;
    c->lines_cap = 0;
    HIP_TRY(hipMalloc(&c->d_lines, need));
    c->lines_cap = need;
  }
  HIP_TRY(dsa::launch_lines(c->d_lines, c->d_links, c->T, c->G, nt, nx, c->stream));
  c->lines_t = nt;
  c->lines_x = nx;
  c->lines_valid = true;
  return 0;
}

int set_device(dsa_chain *c) {
  HIP_TRY(hipSetDevice(c->device));
  return 0;
}

int check_links(const dsa_chain *c, const int32_t *a) {
  const size_t n = (size_t)c->G.V * c->G.D;
  for (size_t l = 0; l < n; ++l)
    if (a[l] < 0 || (uint32_t)a[l] >= c->T.P)
      return fail(DSA_E_ARG, "link " + std::to_string(l) + " holds " + std::to_string(a[l]) + ", outside [0, P)");
  return 0;
}

int upload_links(dsa_chain *c, const int32_t *a) {
  const size_t n = (size_t)c->G.V * c->G.D;
  HIP_TRY(hipMemcpyAsync(c->d_aos, a, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(dsa::launch_pack(c->d_links, c->d_aos, c->G, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->lines_valid = false;
  return 0;
}

}  // namespace

extern "C" {

const char *dsa_last_error(void) { return g_err.c_str(); }
const char *dsa_version(void) { return "dsa-mi355x 0.1 (gfx950)"; }

int dsa_group_load(const char *path, dsa_group **out) {
  if (!out) return fail(DSA_E_ARG, "dsa_group_load: out is NULL");
  *out = nullptr;
  auto *g = new dsa_group();
  std::string err = dsa::load_group(path, g->g);
  if (!err.empty()) {
    delete g;
    return fail(err.rfind("cannot open", 0) == 0 ? DSA_E_IO : DSA_E_FORMAT, err);
  }
  *out = g;
  return 0;
}

int dsa_group_info_get(const dsa_group *g, dsa_group_info *info) {
  if (!g || !info) return fail(DSA_E_ARG, "dsa_group_info_get: NULL argument");
  info->P = g->g.P;
  info->id = g->g.id;
  info->npool = (uint32_t)g->g.pool.size();
  info->min_retr = g->g.min_retr;
  info->nn_retr = g->g.nn_retr;
  return 0;
}

int dsa_group_tables(const dsa_group *g, int32_t *mult, int32_t *inv, double *retr, double *imtr, int32_t *pool) {
  if (!g) return fail(DSA_E_ARG, "dsa_group_tables: NULL group");
  const size_t P = g->g.P;
  if (mult)
    for (size_t n = 0; n < P * P; ++n) mult[n] = g->g.mult[n];
  if (inv)
    for (size_t n = 0; n < P; ++n) inv[n] = g->g.inv[n];
  if (retr) memcpy(retr, g->g.retr.data(), P * sizeof(double));
  if (imtr) memcpy(imtr, g->g.imtr.data(), P * sizeof(double));
  if (pool)
    for (size_t n = 0; n < g->g.pool.size(); ++n) pool[n] = g->g.pool[n];
  return 0;
}

void dsa_group_free(dsa_group *g) {
  if (!g) return;
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto &kv : g->dev) {
    (void)hipSetDevice(kv.first);
    free_tables(kv.second);
  }
  (void)hipSetDevice(cur);
  delete g;
}

int dsa_chain_create(const dsa_group *gc, const dsa_chain_params *p, dsa_chain **out) {
  if (!gc || !p || !out) return fail(DSA_E_ARG, "dsa_chain_create: NULL argument");
  *out = nullptr;
  dsa_group *g = const_cast<dsa_group *>(gc);
  if (int rc0 = dsa::check_params(p)) return rc0;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(DSA_E_DEVICE, "no HIP device visible");

  auto *c = new dsa_chain();
  c->grp = g;
  c->p = *p;
  c->G = dsa::make_geom(p->D, p->Nt, p->Nx);
  auto bail = [&](int code) {
    dsa_chain_destroy(c);
    return code;
  };
  if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(DSA_E_DEVICE, "hipGetDevice failed"));
  int rc = dsa::upload_tables(g, c->device, c->dt);
  if (rc) return bail(rc);
  c->T = dsa::tables_view(g, c->dt);
  if (dsa::sweep_lds_bytes(c->T) > 65536)
    return bail(fail(DSA_E_RANGE, "group order " + std::to_string(c->T.P) + " needs more than 64 KB of LDS tables"));

  std::string err = p->order == DSA_ORDER_COLOUR ? dsa::build_colour_schedule(c->G, c->S)
                                                 : dsa::build_reference_schedule(c->G, c->S);
  if (!err.empty()) return bail(fail(DSA_E_STATE, err));

  const size_t nl = (size_t)c->G.V * c->G.D;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    return bail(fail(DSA_E_DEVICE, "stream/event creation failed"));
  if (hipMalloc(&c->d_links, nl * sizeof(uint16_t)) != hipSuccess ||
      hipMalloc(&c->d_aos, nl * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&c->d_sites, c->S.sites.size() * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&c->d_acc, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&c->d_fall, sizeof(unsigned long long)) != hipSuccess)
    return bail(fail(DSA_E_DEVICE, "device allocation failed"));
  if (hipMemcpy(c->d_sites, c->S.sites.data(), c->S.sites.size() * sizeof(uint32_t), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemset(c->d_acc, 0, sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(c->d_fall, 0, sizeof(unsigned long long)) != hipSuccess)
    return bail(fail(DSA_E_DEVICE, "device upload failed"));
  int lds_max = 0;
  if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess)
    lds_max = 0;
  c->lds_max = (size_t)lds_max;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess) ncu = 0;
  c->ncu = ncu > 0 ? (uint32_t)ncu : 0u;
  if (c->T.mat) {
    c->mat_ok = dsa::sweep_cls_lds_bytes(c->T, dsa::kClsMatrix) <= (size_t)lds_max;
    dsa::mat_margin((int)p->D, p->beta0, p->beta1, p->beta2, g->mat_dev, c->mat_m0, c->mat_macc);
  }
  if (c->T.cprod) {
    // quantisation: half a unit of the class values (2^-24 narrow, 2^-13 wide)
    c->margin8 = dsa::class_margin(g->g, (int)p->D, p->beta0, p->beta1, p->beta2, 0x1p-25);
    c->margin16 = dsa::class_margin(g->g, (int)p->D, p->beta0, p->beta1, p->beta2, 0x1p-14);
    // + the kernels' static class-value array
    c->narrow_ok = dsa::sweep_cls_lds_bytes(c->T, dsa::kClsNarrow) + 256 <= (size_t)lds_max;
    c->field_ok = c->T.fmat && dsa::sweep_cls_lds_bytes(c->T, dsa::kClsField) + 256 <= (size_t)lds_max;
    const bool wide_ok = c->T.cval && dsa::sweep_cls_lds_bytes(c->T, dsa::kClsWide) + 256 <= (size_t)lds_max;
    if (wide_ok && c->margin16 < 1e-2)
      c->cls_best = dsa::kClsWide;
    else if (c->narrow_ok && c->margin8 < 1e-3)
      c->cls_best = dsa::kClsNarrow;
  }

  const bool cold = p->start == DSA_START_COLD;
  if (p->rng == DSA_RNG_MINSTD) {
    // dym-mod-metro.cpp:227-228, 253-255: per-site engines seeded iseed+i; hot
    // links drawn in link order from site 0's engine (host: it is sequential).
    std::vector<uint32_t> ms;
    std::vector<int32_t> a;
    dsa::minstd_start(g->g, *p, c->G.V, ms, a);
    if (hipMalloc(&c->d_mstate, c->G.V * sizeof(uint32_t)) != hipSuccess ||
        hipMemcpy(c->d_mstate, ms.data(), c->G.V * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(DSA_E_DEVICE, "minstd state upload failed"));
    rc = upload_links(c, a.data());
    if (rc) return bail(rc);
  } else {
    hipError_t e = cold ? dsa::launch_fill(c->d_links, c->G, (uint16_t)g->g.id, c->stream)
                        : dsa::launch_init_philox(c->d_links, c->G, c->T.P, (uint32_t)p->seed, c->stream);
    if (e != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
      return bail(fail(DSA_E_DEVICE, "hot/cold start kernel failed"));
  }
  *out = c;
  return 0;
}

void dsa_chain_destroy(dsa_chain *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void *ptr : {(void *)c->d_links, (void *)c->d_sites, (void *)c->d_mstate, (void *)c->d_acc,
                    (void *)c->d_fall, (void *)c->d_aos, (void *)c->d_lines, (void *)c->d_partials, (void *)c->d_out})
    if (ptr) (void)hipFree(ptr);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (hipEvent_t e : c->pev) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int dsa_chain_sweep(dsa_chain *c, uint32_t n) {
  if (!c) return fail(DSA_E_ARG, "dsa_chain_sweep: NULL chain");
  if (set_device(c)) return DSA_E_DEVICE;
  dsa::SweepArgs A{};
  A.links = c->d_links;
  A.mstate = c->d_mstate;
  A.acc = c->d_acc;
  A.T = c->T;
  A.G = c->G;
  A.rng = c->p.rng == DSA_RNG_MINSTD ? dsa::kRngMinstd : dsa::kRngPhilox;
  A.beta0 = c->p.beta0;
  A.beta1 = c->p.beta1;
  A.beta2 = c->p.beta2;
  A.seed = (uint32_t)c->p.seed;
  A.fall = c->d_fall;
  A.lds_max = c->lds_max;
  A.ncu = c->ncu;
  // the class form this chain runs (kClsNone: the table kernel)
  const int mode = run_mode(c);
  A.cmargin = (mode == dsa::kClsMatrix ? c->mat_m0 : mode == dsa::kClsWide ? c->margin16 : c->margin8) * c->margin_scale;
  A.cmargin_acc = c->mat_macc * c->margin_scale;
  A.bound = c->bound_next ? 1 : 0;
  A.field_all = c->kernel == DSA_KERNEL_FIELD ? 1 : 0;
  const size_t np = c->S.passes.size();
  if (c->profiling == 1 && c->pev.size() < 2 * np) {
    for (size_t q = c->pev.size(); q < 2 * np; ++q) {
      hipEvent_t e;
      HIP_TRY(hipEventCreate(&e));
      c->pev.push_back(e);
    }
  }
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  for (uint32_t it = 0; it < n; ++it) {
    A.sweep_lo = (uint32_t)c->sweep;
    A.sweep_hi = (uint32_t)(c->sweep >> 32);
    for (size_t q = 0; q < np; ++q) {
      const dsa::Pass &ps = c->S.passes[q];
      A.mu = ps.mu;
      A.sites = c->d_sites + ps.offset;
      A.count = ps.count;
      A.use_cls = mode;
      if (c->profiling == 1) HIP_TRY(hipEventRecord(c->pev[2 * q], c->stream));
      HIP_TRY(dsa::launch_sweep(A, c->max_blocks, c->stream));
      if (c->profiling == 1) HIP_TRY(hipEventRecord(c->pev[2 * q + 1], c->stream));
      c->hit += (uint64_t)ps.count * dsa::kNHit;
    }
    if (c->profiling == 1) {
      HIP_TRY(hipEventSynchronize(c->pev[2 * np - 1]));
      for (size_t q = 0; q < np; ++q) {
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, c->pev[2 * q], c->pev[2 * q + 1]));
        c->kern_ms += t;
      }
      c->kern_launches += np;
    }
    c->sweep++;
  }
  const unsigned long long acc_before = c->acc_host;
  HIP_TRY(hipMemcpyAsync(&c->acc_host, c->d_acc, sizeof c->acc_host, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  HIP_TRY(hipEventSynchronize(c->ev1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  if (n) {
    const double hits = (double)n * (double)c->G.V * c->G.D * dsa::kNHit;
    c->bound_next = (double)(c->acc_host - acc_before) < dsa::kBoundAcceptance * hits;
  }
  if (c->profiling == 2) {
    c->kern_ms += ms;
    c->kern_launches += (uint64_t)n * np;
  }
  if (n) c->lines_valid = false;
  return 0;
}

int dsa_chain_kernel_form(dsa_chain *c, int32_t *form) {
  if (!c || !form) return fail(DSA_E_ARG, "dsa_chain_kernel_form: NULL argument");
  const int mode = run_mode(c);
  *form = mode == dsa::kClsMatrix  ? DSA_KERNEL_MATRIX
          : mode == dsa::kClsField ? DSA_KERNEL_FIELD
          : mode == dsa::kClsWide  ? DSA_KERNEL_CLASS
          : mode == dsa::kClsNarrow ? DSA_KERNEL_CLASS_NARROW
                                    : DSA_KERNEL_TABLE;
  return 0;
}

int dsa_chain_set_kernel(dsa_chain *c, int32_t kernel, double margin_scale) {
  if (!c) return fail(DSA_E_ARG, "dsa_chain_set_kernel: NULL chain");
  if (kernel != DSA_KERNEL_AUTO && kernel != DSA_KERNEL_TABLE && kernel != DSA_KERNEL_CLASS &&
      kernel != DSA_KERNEL_CLASS_NARROW && kernel != DSA_KERNEL_MATRIX && kernel != DSA_KERNEL_FIELD)
    return fail(DSA_E_ARG, "dsa_chain_set_kernel: unknown kernel " + std::to_string(kernel));
  if (kernel == DSA_KERNEL_MATRIX && !c->mat_ok)
    return fail(DSA_E_STATE, "dsa_chain_set_kernel: the matrix kernel needs the group's SU(3) matrices (appended to "
                             "its file) in this device's LDS");
  if (!(margin_scale >= 1.0) || !std::isfinite(margin_scale))
    return fail(DSA_E_ARG, "dsa_chain_set_kernel: margin_scale must be finite and >= 1");
  if (kernel == DSA_KERNEL_CLASS && c->cls_best == dsa::kClsNone)
    return fail(DSA_E_STATE, "dsa_chain_set_kernel: no class kernel for this group on this device (order > 8192, "
                             "more than 63 classes, or class tables beyond the LDS)");
  if (kernel == DSA_KERNEL_CLASS_NARROW && !c->narrow_ok)
    return fail(DSA_E_STATE, "dsa_chain_set_kernel: the narrow class tables do not fit this device's LDS");
  if (kernel == DSA_KERNEL_FIELD && !(c->narrow_ok && c->field_ok))
    return fail(DSA_E_STATE, "dsa_chain_set_kernel: the field kernel needs the narrow class tables and a GF(4) "
                             "form of the group (a subgroup of SL(3,4)) in this device's LDS");
  c->kernel = kernel;
  c->margin_scale = margin_scale;
  return 0;
}

int dsa_chain_exact_decisions(dsa_chain *c, uint64_t *n) {
  if (!c || !n) return fail(DSA_E_ARG, "dsa_chain_exact_decisions: NULL argument");
  if (set_device(c)) return DSA_E_DEVICE;
  unsigned long long v = 0;
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(&v, c->d_fall, sizeof v, hipMemcpyDeviceToHost));
  *n = v;
  return 0;
}

int dsa_chain_counters(dsa_chain *c, uint64_t *hit, uint64_t *acc) {
  if (!c) return fail(DSA_E_ARG, "dsa_chain_counters: NULL chain");
  if (set_device(c)) return DSA_E_DEVICE;
  unsigned long long a = 0;
  HIP_TRY(hipMemcpy(&a, c->d_acc, sizeof a, hipMemcpyDeviceToHost));
  if (hit) *hit = c->hit;
  if (acc) *acc = a;
  return 0;
}

int dsa_chain_get_sweep(dsa_chain *c, uint64_t *sweep) {
  if (!c || !sweep) return fail(DSA_E_ARG, "dsa_chain_get_sweep: NULL argument");
  *sweep = c->sweep;
  return 0;
}

int dsa_chain_set_sweep(dsa_chain *c, uint64_t sweep) {
  if (!c) return fail(DSA_E_ARG, "dsa_chain_set_sweep: NULL chain");
  c->sweep = sweep;
  return 0;
}

int dsa_chain_measure(dsa_chain *c, uint32_t wk, uint32_t wl, int32_t sum_mode, double *out) {
  if (!c || !out) return fail(DSA_E_ARG, "dsa_chain_measure: NULL argument");
  if (wk > c->p.Nx || wl > c->p.Nt) return fail(DSA_E_ARG, "Wilson grid larger than Nx x Nt");
  if ((wk == 0) != (wl == 0)) return fail(DSA_E_ARG, "Wilson grid must be empty or at least 1x1");
  if (sum_mode != DSA_SUM_TREE && sum_mode != DSA_SUM_ORDERED) return fail(DSA_E_ARG, "unknown sum mode");
  if (set_device(c)) return DSA_E_DEVICE;
  if (wk && wl) {
    int rc = ensure_lines(c, wl, wk);
    if (rc) return rc;
  }
  const uint32_t nobs = 4 + wk * wl;
  if (c->out_cap < nobs) {
    if (c->d_out) HIP_TRY(hipFree(c->d_out));
    c->d_out = nullptr;
    HIP_TRY(hipMalloc(&c->d_out, nobs * sizeof(double)));
    c->out_cap = nobs;
  }
  const uint16_t *lines = (wk && wl) ? c->d_lines : nullptr;
  if (sum_mode == DSA_SUM_TREE) {
    const size_t need = (size_t)nobs * dsa::obs_blocks(c->G);
    if (c->partials_cap < need) {
      if (c->d_partials) HIP_TRY(hipFree(c->d_partials));
      c->d_partials = nullptr;
      HIP_TRY(hipMalloc(&c->d_partials, need * sizeof(double)));
      c->partials_cap = need;
    }
    HIP_TRY(dsa::launch_measure_tree(c->d_out, c->d_partials, c->d_links, lines, c->T, c->G, wk, wl, wl, wk,
                                     c->stream));
  } else {
    HIP_TRY(dsa::launch_measure_ordered(c->d_out, c->d_links, lines, c->T, c->G, wk, wl, wl, wk, c->stream));
  }
  HIP_TRY(hipMemcpyAsync(out, c->d_out, nobs * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int dsa_chain_measure_noalloc(dsa_chain *c, uint32_t wk, uint32_t wl, int32_t sum_mode, double *out) {
  static_assert(dsa::kScratchBytes == DSA_NOALLOC_SCRATCH_BYTES, "work space size differs from the header's");
  if (!c || !out) return fail(DSA_E_ARG, "dsa_chain_measure_noalloc: NULL argument");
  if (wk > c->p.Nx || wl > c->p.Nt) return fail(DSA_E_ARG, "Wilson grid larger than Nx x Nt");
  if ((wk == 0) != (wl == 0)) return fail(DSA_E_ARG, "Wilson grid must be empty or at least 1x1");
  if (sum_mode != DSA_SUM_TREE && sum_mode != DSA_SUM_ORDERED) return fail(DSA_E_ARG, "unknown sum mode");
  // the work space: result staging, per-block partial sums (tree), line
  // products (Wilson grid), each region 256-byte aligned
  auto pad256 = [](size_t n) { return (n + 255) & ~(size_t)255; };
  const uint32_t nobs = 4 + wk * wl;
  const size_t out_b = pad256((size_t)nobs * sizeof(double));
  const size_t part_b = sum_mode == DSA_SUM_TREE ? pad256((size_t)nobs * dsa::obs_blocks(c->G) * sizeof(double)) : 0;
  const size_t lines_b = (wk && wl) ? ((size_t)wl + (size_t)(c->G.D - 1) * wk) * c->G.V * sizeof(uint16_t) : 0;
  const size_t need = out_b + part_b + lines_b;
  if (need > dsa::kScratchBytes)
    return fail(DSA_E_RANGE, "dsa_chain_measure_noalloc: this lattice and Wilson grid need " + std::to_string(need) +
                                 " bytes of work space, above the " + std::to_string(dsa::kScratchBytes) +
                                 "-byte limit (use dsa_chain_measure)");
  if (set_device(c)) return DSA_E_DEVICE;
  void *base = nullptr;
  HIP_TRY(dsa::scratch_region(&base));
  unsigned char *ws = static_cast<unsigned char *>(base);
  double *d_out = reinterpret_cast<double *>(ws);
  double *partials = reinterpret_cast<double *>(ws + out_b);
  uint16_t *lines = (wk && wl) ? reinterpret_cast<uint16_t *>(ws + out_b + part_b) : nullptr;
  // every region is written before it is read: the line kernel fills all
  // rows x V products, every observable block writes its partial sums
  if (lines) HIP_TRY(dsa::launch_lines(lines, c->d_links, c->T, c->G, wl, wk, c->stream));
  if (sum_mode == DSA_SUM_TREE)
    HIP_TRY(dsa::launch_measure_tree(d_out, partials, c->d_links, lines, c->T, c->G, wk, wl, wl, wk, c->stream));
  else
    HIP_TRY(dsa::launch_measure_ordered(d_out, c->d_links, lines, c->T, c->G, wk, wl, wl, wk, c->stream));
  HIP_TRY(hipMemcpyAsync(out, d_out, nobs * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int dsa_chain_get_links(dsa_chain *c, int32_t *a) {
  if (!c || !a) return fail(DSA_E_ARG, "dsa_chain_get_links: NULL argument");
  if (set_device(c)) return DSA_E_DEVICE;
  const size_t n = (size_t)c->G.V * c->G.D;
  HIP_TRY(dsa::launch_unpack(c->d_aos, c->d_links, c->G, c->stream));
  HIP_TRY(hipMemcpyAsync(a, c->d_aos, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int dsa_chain_set_links(dsa_chain *c, const int32_t *a) {
  if (!c || !a) return fail(DSA_E_ARG, "dsa_chain_set_links: NULL argument");
  if (check_links(c, a)) return DSA_E_ARG;
  if (set_device(c)) return DSA_E_DEVICE;
  return upload_links(c, a);
}

int dsa_group_matrices(const dsa_group *g, double *out) {
  if (!g || !out) return fail(DSA_E_ARG, "dsa_group_matrices: NULL argument");
  if (g->g.su3.empty()) return fail(DSA_E_STATE, "group file " + g->g.path + " has no appended SU(3) matrices");
  memcpy(out, g->g.su3.data(), g->g.su3.size() * sizeof(double));
  return 0;
}

int dsa_group_field(const dsa_group *g, uint32_t *mat, uint16_t *key) {
  if (!g) return fail(DSA_E_ARG, "dsa_group_field: NULL group");
  if (g->g.fmat.empty()) return fail(DSA_E_STATE, "group " + g->g.path + " has no GF(4) form");
  const size_t P = g->g.P;
  // packed planes (lo bits 0-8, hi bits 20-28, entry (r, c) at 3r + c) to
  // 2-bit entries; the first two rows of the plain code are its low 12 bits
  // and key the same element as the packed word's key
  std::vector<uint16_t> k2(dsa::kGf4Keys, 0xFFFF);
  for (size_t e = 0; e < P; ++e) {
    const uint32_t w = g->g.fmat[e];
    uint32_t plain = 0;
    for (int q = 0; q < 9; ++q) plain |= (((w >> q) & 1u) | ((w >> (20 + q)) & 1u) << 1) << (2 * q);
    if (mat) mat[e] = plain;
    k2[plain & 0xFFFu] = (uint16_t)e;
  }
  if (key) memcpy(key, k2.data(), k2.size() * sizeof(uint16_t));
  return 0;
}

int dsa_cfg_save(const char *path, const int32_t *a, uint64_t n) {
  if (!path || (!a && n)) return fail(DSA_E_ARG, "dsa_cfg_save: NULL argument");
  FILE *f = fopen(path, "wb");
  if (!f) return fail(DSA_E_IO, std::string("dsa_cfg_save: cannot open ") + path);
  const size_t w = fwrite(a, sizeof(int32_t), (size_t)n, f);
  const int cl = fclose(f);
  if (w != (size_t)n || cl) return fail(DSA_E_IO, std::string("dsa_cfg_save: short write to ") + path);
  return 0;
}

int dsa_cfg_load(const char *path, int32_t *a, uint64_t n) {
  if (!path || (!a && n)) return fail(DSA_E_ARG, "dsa_cfg_load: NULL argument");
  FILE *f = fopen(path, "rb");
  if (!f) return fail(DSA_E_IO, std::string("dsa_cfg_load: cannot open ") + path);
  const size_t r = fread(a, sizeof(int32_t), (size_t)n, f);
  const bool more = fgetc(f) != EOF;
  fclose(f);
  if (r != (size_t)n || more)
    return fail(DSA_E_RANGE, std::string("dsa_cfg_load: ") + path + " does not hold exactly " + std::to_string(n) +
                                 " int32 links");
  return 0;
}

int dsa_chain_save_cfg(dsa_chain *c, const char *path) {
  if (!c) return fail(DSA_E_ARG, "dsa_chain_save_cfg: NULL chain");
  std::vector<int32_t> a((size_t)c->G.V * c->G.D);
  if (int rc = dsa_chain_get_links(c, a.data())) return rc;
  return dsa_cfg_save(path, a.data(), a.size());
}

int dsa_chain_load_cfg(dsa_chain *c, const char *path) {
  if (!c) return fail(DSA_E_ARG, "dsa_chain_load_cfg: NULL chain");
  std::vector<int32_t> a((size_t)c->G.V * c->G.D);
  if (int rc = dsa_cfg_load(path, a.data(), a.size())) return rc;
  return dsa_chain_set_links(c, a.data());
}

int dsa_chain_save_kentucky(dsa_chain *c, const char *prefix, int32_t num, char *name, uint64_t cap) {
  if (!c || !prefix) return fail(DSA_E_ARG, "dsa_chain_save_kentucky: NULL argument");
  if (c->G.D != 4) return fail(DSA_E_STATE, "dsa_chain_save_kentucky: the Kentucky layout is four-dimensional");
  const std::vector<double> &m = c->grp->g.su3;
  if (m.empty()) return fail(DSA_E_STATE, "group file " + c->grp->g.path + " has no appended SU(3) matrices");
  if (set_device(c)) return DSA_E_DEVICE;
  char fname[1024];
  snprintf(fname, sizeof fname, "%slat-beta1%.3f-beta2%.3f-nt%02d-nx%02d-num%04d", prefix, c->p.beta1, c->p.beta2,
           (int)c->p.Nt, (int)c->p.Nx, (int)num);
  const size_t nd = (size_t)c->G.V * 4 * 18;
  double *d_out = nullptr, *d_m = nullptr;
  std::vector<double> host(nd);
  hipError_t e = hipMalloc(&d_out, nd * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&d_m, m.size() * sizeof(double));
  if (e == hipSuccess) e = hipMemcpyAsync(d_m, m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = dsa::launch_kentucky(d_out, c->d_links, d_m, c->G, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(host.data(), d_out, nd * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (d_out) (void)hipFree(d_out);
  if (d_m) (void)hipFree(d_m);
  if (e != hipSuccess) return fail(DSA_E_DEVICE, std::string("dsa_chain_save_kentucky: ") + hipGetErrorString(e));
  FILE *f = fopen(fname, "wb");
  if (!f) return fail(DSA_E_IO, std::string("dsa_chain_save_kentucky: cannot open ") + fname);
  const size_t w = fwrite(host.data(), sizeof(double), nd, f);
  const int cl = fclose(f);
  if (w != nd || cl) return fail(DSA_E_IO, std::string("dsa_chain_save_kentucky: short write to ") + fname);
  if (name && cap) {
    strncpy(name, fname, (size_t)cap - 1);
    name[cap - 1] = 0;
  }
  return 0;
}

int dsa_chain_get_rng_state(dsa_chain *c, uint32_t *state) {
  if (!c || !state) return fail(DSA_E_ARG, "dsa_chain_get_rng_state: NULL argument");
  if (!c->d_mstate) return fail(DSA_E_STATE, "chain does not use the minstd streams");
  if (set_device(c)) return DSA_E_DEVICE;
  HIP_TRY(hipMemcpy(state, c->d_mstate, c->G.V * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return 0;
}

int dsa_chain_set_rng_state(dsa_chain *c, const uint32_t *state) {
  if (!c || !state) return fail(DSA_E_ARG, "dsa_chain_set_rng_state: NULL argument");
  if (!c->d_mstate) return fail(DSA_E_STATE, "chain does not use the minstd streams");
  for (uint32_t i = 0; i < c->G.V; ++i)
    if (state[i] == 0 || state[i] >= 2147483647u) return fail(DSA_E_ARG, "minstd state outside [1, 2^31-2]");
  if (set_device(c)) return DSA_E_DEVICE;
  HIP_TRY(hipMemcpy(c->d_mstate, state, c->G.V * sizeof(uint32_t), hipMemcpyHostToDevice));
  return 0;
}

int dsa_chain_staples(dsa_chain *c, int32_t *st, int32_t *re) {
  if (!c || !st || !re) return fail(DSA_E_ARG, "dsa_chain_staples: NULL argument");
  if (set_device(c)) return DSA_E_DEVICE;
  const size_t nl = (size_t)c->G.V * c->G.D, ns = 2 * (c->G.D - 1), nr = 6 * (c->G.D - 1);
  int32_t *dst = nullptr, *dre = nullptr;
  HIP_TRY(hipMalloc(&dst, nl * ns * sizeof(int32_t)));
  if (hipMalloc(&dre, nl * nr * sizeof(int32_t)) != hipSuccess) {
    (void)hipFree(dst);
    return fail(DSA_E_DEVICE, "allocation failed");
  }
  hipError_t e = dsa::launch_staples(dst, dre, c->d_links, c->T, c->G, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(st, dst, nl * ns * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(re, dre, nl * nr * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dst);
  (void)hipFree(dre);
  if (e != hipSuccess) return fail(DSA_E_DEVICE, std::string("staples: ") + hipGetErrorString(e));
  return 0;
}

int dsa_chain_loops(dsa_chain *c, int32_t kind, uint32_t k, uint32_t l, int32_t *out) {
  if (!c || !out) return fail(DSA_E_ARG, "dsa_chain_loops: NULL argument");
  if (kind < DSA_LOOP_PLAQ || kind > DSA_LOOP_WILSON) return fail(DSA_E_ARG, "unknown loop kind");
  if (kind == DSA_LOOP_WILSON && (k >= c->p.Nx || l >= c->p.Nt)) return fail(DSA_E_ARG, "Wilson loop too large");
  if (set_device(c)) return DSA_E_DEVICE;
  const uint32_t D = c->G.D, V = c->G.V;
  size_t n = kind == DSA_LOOP_PLAQ ? (size_t)V * D * (D - 1) / 2
             : kind == DSA_LOOP_RECT ? (size_t)V * D * (D - 1)
             : kind == DSA_LOOP_POLY ? (size_t)V / c->p.Nt
                                     : (size_t)V * (D - 1);
  if (kind == DSA_LOOP_WILSON) {
    int rc = ensure_lines(c, l + 1, k + 1);
    if (rc) return rc;
  }
  int32_t *d = nullptr;
  HIP_TRY(hipMalloc(&d, n * sizeof(int32_t)));
  hipError_t e = dsa::launch_loops(d, kind, (int)k, (int)l, c->d_links, c->d_lines, c->T, c->G, c->lines_t,
                                   c->lines_x, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(DSA_E_DEVICE, std::string("loops: ") + hipGetErrorString(e));
  return 0;
}

int dsa_chain_schedule_info(dsa_chain *c, uint32_t *npasses, uint32_t *ncolours) {
  if (!c) return fail(DSA_E_ARG, "dsa_chain_schedule_info: NULL chain");
  if (npasses) *npasses = (uint32_t)c->S.passes.size();
  if (ncolours)
    for (int d = 0; d < c->G.D; ++d) ncolours[d] = c->S.ncol[d];
  return 0;
}

int dsa_schedule_query(uint32_t D, uint32_t Nt, uint32_t Nx, int32_t order, uint32_t *npasses, uint32_t *nsites,
                       uint32_t *pass_mu, uint32_t *pass_count, uint32_t *sites) {
  if (!npasses || !nsites) return fail(DSA_E_ARG, "dsa_schedule_query: NULL counts");
  if (D < 2 || D > (uint32_t)dsa::kDMax || Nt < 1 || Nx < 1) return fail(DSA_E_RANGE, "unsupported lattice shape");
  uint64_t V = Nt;
  for (uint32_t d = 1; d < D; ++d) V *= Nx;
  if (V * D >= (1ull << 31)) return fail(DSA_E_RANGE, "V*D must be below 2^31 links");
  dsa::Geom G = dsa::make_geom(D, Nt, Nx);
  dsa::Schedule S;
  std::string err = order == DSA_ORDER_REFERENCE ? dsa::build_reference_schedule(G, S)
                                                 : dsa::build_colour_schedule(G, S);
  if (!err.empty()) return fail(DSA_E_STATE, err);
  const bool fill = pass_mu && pass_count && sites;
  if (fill) {
    if (*npasses < S.passes.size() || *nsites < S.sites.size()) return fail(DSA_E_ARG, "output arrays too small");
    for (size_t q = 0; q < S.passes.size(); ++q) {
      pass_mu[q] = (uint32_t)S.passes[q].mu;
      pass_count[q] = S.passes[q].count;
    }
    memcpy(sites, S.sites.data(), S.sites.size() * sizeof(uint32_t));
  }
  *npasses = (uint32_t)S.passes.size();
  *nsites = (uint32_t)S.sites.size();
  return 0;
}

int dsa_chain_last_sweep_ms(dsa_chain *c, double *ms) {
  if (!c || !ms) return fail(DSA_E_ARG, "dsa_chain_last_sweep_ms: NULL argument");
  *ms = c->last_ms;
  return 0;
}

int dsa_chain_set_profiling(dsa_chain *c, int32_t on) {
  if (!c) return fail(DSA_E_ARG, "dsa_chain_set_profiling: NULL chain");
  if (on < 0 || on > 2) return fail(DSA_E_ARG, "dsa_chain_set_profiling: mode must be 0, 1 or 2");
  c->profiling = on;
  c->kern_ms = 0;
  c->kern_launches = 0;
  return 0;
}

int dsa_chain_kernel_time(dsa_chain *c, double *ms, uint64_t *launches) {
  if (!c || !ms || !launches) return fail(DSA_E_ARG, "dsa_chain_kernel_time: NULL argument");
  *ms = c->kern_ms;
  *launches = c->kern_launches;
  return 0;
}

int dsa_probe_gather(const dsa_group *gc, uint32_t blocks, uint32_t rounds, double *gathers_per_s) {
  if (!gc || !gathers_per_s || !blocks || !rounds) return fail(DSA_E_ARG, "dsa_probe_gather: bad argument");
  dsa_group *g = const_cast<dsa_group *>(gc);
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  DeviceTables t;
  int rc = dsa::upload_tables(g, dev, t);
  if (rc) return rc;
  uint32_t *sink = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipMalloc(&sink, (size_t)blocks * 256 * sizeof(uint32_t)));
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  hipError_t e = dsa::launch_gather_probe(t.mult, g->g.P, blocks, 2, sink, s);  // warm the table into L2
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  if (e == hipSuccess) e = dsa::launch_gather_probe(t.mult, g->g.P, blocks, rounds, sink, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  (void)hipFree(sink);
  if (e != hipSuccess) return fail(DSA_E_DEVICE, std::string("gather probe: ") + hipGetErrorString(e));
  *gathers_per_s = (double)blocks * 256.0 * rounds * 24.0 / (ms * 1e-3);
  return 0;
}

int dsa_probe_copy(uint64_t bytes, uint32_t reps, double *ms_per_copy) {
  if (!ms_per_copy || bytes < 16 || !reps) return fail(DSA_E_ARG, "dsa_probe_copy: bad argument");
  bytes &= ~(uint64_t)15;
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipMalloc(&a, bytes));
  HIP_TRY(hipMalloc(&b, bytes));
  HIP_TRY(hipMemset(a, 1, bytes));
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  hipError_t e = dsa::launch_copy_probe(a, b, bytes, s);
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) e = dsa::launch_copy_probe(a, b, bytes, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return fail(DSA_E_DEVICE, std::string("copy probe: ") + hipGetErrorString(e));
  *ms_per_copy = ms / reps;
  return 0;
}

}  // extern "C"
