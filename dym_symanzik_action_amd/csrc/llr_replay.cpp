// llr_replay.cpp -- the C ABI of the device replay of the reference llr's serial
// chain (include/dsa.h, dsa_llr_ref_*): main's set-up and the steps of
// run_pass (llr.cpp:178-274, 296-326), each enqueued on the chain's stream.
#include <hip/hip_runtime.h>
#include <string.h>

#include <cmath>
#include <string>
#include <vector>

#include "dsa_math.hpp"
#include "internal.hpp"
#include "schedule.hpp"

using dsa::fail;

struct dsa_llr_ref {
  dsa_group *grp = nullptr;
  dsa::DeviceTables dt;
  dsa::LrefArgs A{};
  int device = 0;
  int32_t seed = 0;
  hipStream_t stream = nullptr;
  int32_t *d_aos = nullptr;
  double *d_sums = nullptr;  // [0..1] current sums, [2..3] the calibration's sums before the flip
  dsa::LrefState host{};
};

namespace {

constexpr uint32_t kCap = 20000;  // seed_window's CAP (llr.cpp:225)
constexpr uint32_t kDriveChunk = 16;

int pull(dsa_llr_ref *R) {
  HIP_TRY(hipMemcpyAsync(&R->host, R->A.st, sizeof R->host, hipMemcpyDeviceToHost, R->stream));
  HIP_TRY(hipStreamSynchronize(R->stream));
  return 0;
}

// resync_E (llr.cpp:175): both energies recomputed in the reference's order
int resync(dsa_llr_ref *R) {
  HIP_TRY(dsa::launch_lref_sums(R->A, R->d_sums, R->stream));
  HIP_TRY(dsa::launch_lref_resync(R->A, R->d_sums, 0, 0.0, 1.0, 0, 1, R->stream));
  return 0;
}

void release(dsa_llr_ref *R) {
  for (void *p : {(void *)R->A.links, (void *)R->A.ms, (void *)R->A.st, (void *)R->A.t1, (void *)R->A.t2,
                  (void *)R->d_aos, (void *)R->d_sums})
    if (p) (void)hipFree(p);
  if (R->stream) (void)hipStreamDestroy(R->stream);
}

}  // namespace

extern "C" {

int dsa_llr_ref_create(const dsa_group *gc, uint32_t D, uint32_t Nt, uint32_t Nx, int32_t axis, double bg,
                       int32_t seed, dsa_llr_ref **out) {
  if (!gc || !out) return fail(DSA_E_ARG, "dsa_llr_ref_create: NULL argument");
  *out = nullptr;
  dsa_chain_params cp{D, Nt, Nx, 0, 0, 0, seed, DSA_START_HOT, DSA_RNG_MINSTD, DSA_ORDER_REFERENCE};
  if (int rc = dsa::check_params(&cp)) return rc;
  if (axis != 1 && axis != 2) return fail(DSA_E_ARG, "dsa_llr_ref_create: axis must be 1 or 2");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(DSA_E_DEVICE, "no HIP device visible");
  auto *R = new dsa_llr_ref();
  auto bail = [&](int rc) {
    release(R);
    delete R;
    return rc;
  };
  R->grp = const_cast<dsa_group *>(gc);
  R->seed = seed;
  if (hipGetDevice(&R->device) != hipSuccess) return bail(fail(DSA_E_DEVICE, "hipGetDevice failed"));
  if (int rc = dsa::upload_tables(R->grp, R->device, R->dt)) return bail(rc);
  dsa::LrefArgs &A = R->A;
  A.T = dsa::tables_view(R->grp, R->dt);
  A.G = dsa::make_geom(D, Nt, Nx);
  A.axis = axis;
  A.bg = bg;
  const size_t V = A.G.V, nl = V * D;
  int lds_max = 0;
  if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, R->device) != hipSuccess ||
      (size_t)A.T.P * sizeof(double) > (size_t)lds_max)
    return bail(fail(DSA_E_RANGE, "dsa_llr_ref_create: the group's ReTr table does not fit the LDS"));
  if (hipStreamCreateWithFlags(&R->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&A.links, nl * sizeof(uint16_t)) != hipSuccess ||
      hipMalloc(&A.ms, V * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&A.st, sizeof(dsa::LrefState)) != hipSuccess ||
      hipMalloc(&A.t1, V * D * (D - 1) / 2 * sizeof(double)) != hipSuccess ||
      hipMalloc(&A.t2, nl * 6 * (D - 1) * sizeof(double)) != hipSuccess ||
      hipMalloc(&R->d_aos, nl * sizeof(int32_t)) != hipSuccess || hipMalloc(&R->d_sums, 4 * sizeof(double)) != hipSuccess)
    return bail(fail(DSA_E_DEVICE, "dsa_llr_ref_create: allocation failed"));
  dsa::LrefState st0{};
  st0.rect_mult = 6.0;
  // main (llr.cpp:314-320): streams seeded iseed + i, hot links from site 0's
  // stream, then calibrate_rect_mult on that configuration
  hipError_t e = hipMemcpyAsync(A.st, &st0, sizeof st0, hipMemcpyHostToDevice, R->stream);
  if (e == hipSuccess) e = dsa::launch_lref_seed(A, (uint32_t)seed, R->stream);
  if (e == hipSuccess) e = dsa::launch_lref_hot(A, R->stream);
  if (e == hipSuccess) e = dsa::launch_lref_sums(A, R->d_sums + 2, R->stream);          // T_old
  if (e == hipSuccess) e = dsa::launch_lref_calibrate(A, R->d_sums + 2, R->d_sums, 0, R->stream);  // flip
  if (e == hipSuccess) e = dsa::launch_lref_sums(A, R->d_sums, R->stream);              // T_new
  if (e == hipSuccess) e = dsa::launch_lref_calibrate(A, R->d_sums + 2, R->d_sums, 1, R->stream);  // restore
  if (e != hipSuccess) return bail(fail(DSA_E_DEVICE, std::string("dsa_llr_ref_create: ") + hipGetErrorString(e)));
  if (int rc = resync(R)) return bail(rc);
  if (int rc = pull(R)) return bail(rc);
  *out = R;
  return 0;
}

void dsa_llr_ref_destroy(dsa_llr_ref *R) {
  if (!R) return;
  (void)hipSetDevice(R->device);
  if (R->stream) (void)hipStreamSynchronize(R->stream);
  release(R);
  delete R;
}

int dsa_llr_ref_rect_mult(dsa_llr_ref *R, double *rect_mult) {
  if (!R || !rect_mult) return fail(DSA_E_ARG, "dsa_llr_ref_rect_mult: NULL argument");
  *rect_mult = R->host.rect_mult;
  return 0;
}

int dsa_llr_ref_reseed(dsa_llr_ref *R, uint32_t repeat) {
  if (!R) return fail(DSA_E_ARG, "dsa_llr_ref_reseed: NULL chain");
  HIP_TRY(hipSetDevice(R->device));
  HIP_TRY(dsa::launch_lref_seed(R->A, (uint32_t)R->seed + repeat * 1000003u, R->stream));
  HIP_TRY(hipStreamSynchronize(R->stream));
  return 0;
}

int dsa_llr_ref_start(dsa_llr_ref *R, int32_t start) {
  if (!R) return fail(DSA_E_ARG, "dsa_llr_ref_start: NULL chain");
  if (start != DSA_START_HOT && start != DSA_START_COLD) return fail(DSA_E_ARG, "dsa_llr_ref_start: unknown start");
  HIP_TRY(hipSetDevice(R->device));
  if (start == DSA_START_HOT)
    HIP_TRY(dsa::launch_lref_hot(R->A, R->stream));
  else
    HIP_TRY(dsa::launch_fill(R->A.links, R->A.G, (uint16_t)R->grp->g.id, R->stream));
  if (int rc = resync(R)) return rc;
  return pull(R);
}

int dsa_llr_ref_seed_window(dsa_llr_ref *R, double Elo, double Ehi, double a_guess, int32_t *reached) {
  if (!R || !reached) return fail(DSA_E_ARG, "dsa_llr_ref_seed_window: NULL argument");
  HIP_TRY(hipSetDevice(R->device));
  HIP_TRY(dsa::launch_lref_set(R->A, R->host.a, 0, R->stream));
  // drive-in sweeps enqueued in chunks: each one is skipped on the device once
  // the running energy is inside the window or the cap is reached
  for (;;) {
    for (uint32_t k = 0; k < kDriveChunk; ++k)
      HIP_TRY(dsa::launch_lref_update(R->A, a_guess, Elo, Ehi, 1, kCap, R->stream));
    if (int rc = pull(R)) return rc;
    const double Ec = R->A.axis == 1 ? R->host.S1 : R->host.S2;
    if ((Ec >= Elo && Ec <= Ehi) || R->host.drive >= kCap) break;
  }
  if (int rc = resync(R)) return rc;
  if (int rc = pull(R)) return rc;
  const double Ec = R->A.axis == 1 ? R->host.S1 : R->host.S2;
  *reached = Ec >= Elo && Ec <= Ehi;
  return 0;
}

int dsa_llr_ref_rm_solve(dsa_llr_ref *R, double Ewin, double hw, double a0, uint32_t K, uint32_t NRM, double *a) {
  if (!R || !a) return fail(DSA_E_ARG, "dsa_llr_ref_rm_solve: NULL argument");
  if (!K) return fail(DSA_E_ARG, "dsa_llr_ref_rm_solve: K must be positive");
  HIP_TRY(hipSetDevice(R->device));
  const double Elo = Ewin - hw, Ehi = Ewin + hw, W = 2 * hw;
  HIP_TRY(dsa::launch_lref_set(R->A, a0, 0, R->stream));
  for (uint32_t m = 0; m < NRM; ++m) {
    for (uint32_t k = 0; k < K; ++k)
      HIP_TRY(dsa::launch_lref_update(R->A, std::nan(""), Elo, Ehi, 0, 0, R->stream));
    HIP_TRY(dsa::launch_lref_sums(R->A, R->d_sums, R->stream));
    HIP_TRY(dsa::launch_lref_resync(R->A, R->d_sums, 1, Ewin, W, m, K, R->stream));
  }
  if (int rc = pull(R)) return rc;
  *a = R->host.a;
  return 0;
}

int dsa_llr_ref_state(dsa_llr_ref *R, double *S1, double *S2, double *maxdrift, uint64_t *hit, uint64_t *acc) {
  if (!R) return fail(DSA_E_ARG, "dsa_llr_ref_state: NULL chain");
  HIP_TRY(hipSetDevice(R->device));
  if (int rc = pull(R)) return rc;
  if (S1) *S1 = R->host.S1;
  if (S2) *S2 = R->host.S2;
  if (maxdrift) *maxdrift = R->host.maxdrift;
  if (hit) *hit = R->host.hit;
  if (acc) *acc = R->host.acc;
  return 0;
}

int dsa_llr_ref_get_links(dsa_llr_ref *R, int32_t *a) {
  if (!R || !a) return fail(DSA_E_ARG, "dsa_llr_ref_get_links: NULL argument");
  HIP_TRY(hipSetDevice(R->device));
  const size_t n = (size_t)R->A.G.V * R->A.G.D;
  HIP_TRY(dsa::launch_unpack(R->d_aos, R->A.links, R->A.G, R->stream));
  HIP_TRY(hipMemcpyAsync(a, R->d_aos, n * sizeof(int32_t), hipMemcpyDeviceToHost, R->stream));
  HIP_TRY(hipStreamSynchronize(R->stream));
  return 0;
}

}  // extern "C"
