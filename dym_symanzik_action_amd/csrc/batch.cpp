// batch.cpp -- many independent chains advanced together (include/dsa.h,
// dsa_batch_*).
//
// The reference scans a phase diagram by running one process per (beta1,
// beta2) point (evermore.sh, jobs2run).  Here every chain of a scan lives on
// the device at once; one kernel launch per pass index sweeps that pass of
// every chain, so small lattices (a few thousand links per pass) still fill
// the 256 CUs.  Chains never interact: each keeps its own Philox key and
// visiting order, so a chain's links are the same alone and in any batch.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "dsa_math.hpp"
#include "internal.hpp"
#include "schedule.hpp"

using dsa::fail;

namespace {

// links per sweep block (one per lane): the candidates the tuner times
constexpr uint32_t kTaskCandidates[] = {256, 512, 1024};
constexpr uint32_t kObsSites = 1024;   // sites per measurement block

struct HostChain {
  dsa_group *grp;
  dsa_chain_params p;
  dsa::Geom G;
  dsa::Schedule S;
  dsa::DeviceTables dt;
  size_t links_off = 0, sites_off = 0, ms_off = 0;
  bool cls = false;     // class-product update
  bool field = false;   // ... with GF(4) products
  double margin = 0;    // its decision bound at margin scale 1
};

}  // namespace

struct dsa_batch {
  std::vector<HostChain> chains;
  bool final = false;
  int device = 0;
  int D = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint16_t *d_links = nullptr;
  uint32_t *d_sites = nullptr;
  uint32_t *d_mstate = nullptr;
  dsa::BatchChain *d_chains = nullptr;
  dsa::BatchTask *d_tasks = nullptr;
  dsa::ObsTask *d_obs = nullptr;
  uint32_t *d_obs_first = nullptr, *d_obs_n = nullptr;
  unsigned long long *d_acc = nullptr, *d_sweep = nullptr;
  double *d_partials = nullptr, *d_out = nullptr;
  int32_t *d_aos = nullptr;
  std::vector<uint32_t> pass_off, pass_n;  // tasks of pass q: [pass_off[q], pass_off[q] + pass_n[q])
  uint32_t nobs = 0;
  size_t lds = 0;
  int kernel = DSA_KERNEL_AUTO;
  double margin_scale = 1.0;
  uint32_t task_links = 0;  // requested links per block (0: tuned at finalize)
  uint32_t block = 256;     // in use
  std::vector<double> tune_ms;  // ms per sweep of each candidate, when tuned
  unsigned long long *d_fall = nullptr;
  uint64_t sweeps = 0;
  double last_ms = 0;
  hipGraphExec_t graph = nullptr;
  hipGraph_t graph_def = nullptr;
  uint32_t graph_sweeps = 0;
  // per chain: the descriptors as uploaded and the accepted hits seen so
  // far, for choosing each class chain's plaquette-first test from its
  // acceptance over the previous call (dsa::kBoundAcceptance)
  std::vector<dsa::BatchChain> desc;
  std::vector<unsigned long long> acc_seen;
};

namespace {

int enqueue_sweeps(dsa_batch *b, uint32_t n) {
  for (uint32_t it = 0; it < n; ++it) {
    for (size_t q = 0; q < b->pass_off.size(); ++q)
      HIP_TRY(dsa::launch_sweep_batch(b->D, b->d_chains, b->d_tasks + b->pass_off[q], b->pass_n[q], b->block,
                                      b->lds, b->d_sweep, b->d_acc, b->stream));
    HIP_TRY(dsa::launch_count_sweep(b->d_sweep, b->stream));
  }
  return 0;
}

// Pass q of every chain, split into blocks of `links` links of one chain,
// uploaded to d_tasks (sized for the smallest candidate).
int set_tasks(dsa_batch *b, uint32_t links) {
  std::vector<dsa::BatchTask> tasks;
  size_t maxpass = 0;
  for (const HostChain &hc : b->chains) maxpass = std::max(maxpass, hc.S.passes.size());
  b->pass_off.clear();
  b->pass_n.clear();
  for (size_t q = 0; q < maxpass; ++q) {
    b->pass_off.push_back((uint32_t)tasks.size());
    for (size_t c = 0; c < b->chains.size(); ++c) {
      const HostChain &hc = b->chains[c];
      if (q >= hc.S.passes.size()) continue;
      const dsa::Pass &ps = hc.S.passes[q];
      for (uint32_t o = 0; o < ps.count; o += links)
        tasks.push_back(dsa::BatchTask{(uint32_t)c, ps.offset + o, std::min(links, ps.count - o), ps.mu});
    }
    b->pass_n.push_back((uint32_t)tasks.size() - b->pass_off.back());
  }
  if (!tasks.empty())
    HIP_TRY(hipMemcpy(b->d_tasks, tasks.data(), tasks.size() * sizeof(dsa::BatchTask), hipMemcpyHostToDevice));
  b->block = links;
  return 0;
}

// Times two sweeps with each block size (after one untimed sweep) and keeps
// the fastest.  The state the sweeps change (links, minstd states, counters)
// is saved before and restored after each candidate, so tuning leaves no
// trace in the results.
int tune_tasks(dsa_batch *b, size_t nlinks, size_t nms) {
  const size_t nc = b->chains.size();
  uint16_t *links0 = nullptr;
  uint32_t *ms0 = nullptr;
  unsigned long long *acc0 = nullptr, *misc0 = nullptr;
  auto release = [&]() {
    for (void *p : {(void *)links0, (void *)ms0, (void *)acc0, (void *)misc0})
      if (p) (void)hipFree(p);
  };
  auto copy = [&](void *dst, const void *src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, b->stream) : hipSuccess;
  };
  hipError_t e = hipMalloc(&links0, nlinks * sizeof(uint16_t));
  if (e == hipSuccess && nms) e = hipMalloc(&ms0, nms * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&acc0, nc * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc(&misc0, 2 * sizeof(unsigned long long));
  if (e == hipSuccess) e = copy(links0, b->d_links, nlinks * sizeof(uint16_t));
  if (e == hipSuccess) e = copy(ms0, b->d_mstate, nms * sizeof(uint32_t));
  if (e == hipSuccess) e = copy(acc0, b->d_acc, nc * sizeof(unsigned long long));
  if (e == hipSuccess) e = copy(misc0, b->d_sweep, sizeof(unsigned long long));
  if (e == hipSuccess) e = copy(misc0 + 1, b->d_fall, sizeof(unsigned long long));
  if (e != hipSuccess) {
    release();
    return fail(DSA_E_DEVICE, std::string("batch tuning: ") + hipGetErrorString(e));
  }
  uint32_t best = kTaskCandidates[0];
  double best_ms = 0;
  b->tune_ms.clear();
  for (uint32_t links : kTaskCandidates) {
    int rc = set_tasks(b, links);
    if (!rc) rc = enqueue_sweeps(b, 1);
    if (!rc && hipEventRecord(b->ev0, b->stream) != hipSuccess) rc = fail(DSA_E_DEVICE, "event record failed");
    if (!rc) rc = enqueue_sweeps(b, 2);
    float ms = 0;
    if (!rc && (hipEventRecord(b->ev1, b->stream) != hipSuccess || hipEventSynchronize(b->ev1) != hipSuccess ||
                hipEventElapsedTime(&ms, b->ev0, b->ev1) != hipSuccess))
      rc = fail(DSA_E_DEVICE, "event timing failed");
    e = copy(b->d_links, links0, nlinks * sizeof(uint16_t));
    if (e == hipSuccess) e = copy(b->d_mstate, ms0, nms * sizeof(uint32_t));
    if (e == hipSuccess) e = copy(b->d_acc, acc0, nc * sizeof(unsigned long long));
    if (e == hipSuccess) e = copy(b->d_sweep, misc0, sizeof(unsigned long long));
    if (e == hipSuccess) e = copy(b->d_fall, misc0 + 1, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (!rc && e != hipSuccess) rc = fail(DSA_E_DEVICE, std::string("batch tuning: ") + hipGetErrorString(e));
    if (rc) {
      release();
      return rc;
    }
    b->tune_ms.push_back(ms / 2.0);
    if (b->tune_ms.size() == 1 || ms / 2.0 < best_ms) {
      best = links;
      best_ms = ms / 2.0;
    }
  }
  release();
  return set_tasks(b, best);
}

// After a call of n sweeps: every class chain whose acceptance over the call
// was below dsa::kBoundAcceptance runs the plaquette-first test in the next
// call (descriptors re-uploaded when a choice changes; the stream is idle).
int choose_bounds(dsa_batch *b, uint32_t n) {
  if (!n) return 0;
  const size_t nc = b->chains.size();
  std::vector<unsigned long long> a(nc);
  HIP_TRY(hipMemcpy(a.data(), b->d_acc, nc * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  bool changed = false;
  for (size_t c = 0; c < nc; ++c) {
    const HostChain &hc = b->chains[c];
    const double hits = (double)n * hc.G.V * hc.G.D * dsa::kNHit;
    const int want = hc.cls && (double)(a[c] - b->acc_seen[c]) < dsa::kBoundAcceptance * hits ? 1 : 0;
    b->acc_seen[c] = a[c];
    if (want != b->desc[c].bound) {
      b->desc[c].bound = want;
      changed = true;
    }
  }
  if (changed)
    HIP_TRY(hipMemcpy(b->d_chains, b->desc.data(), nc * sizeof(dsa::BatchChain), hipMemcpyHostToDevice));
  return 0;
}

int enqueue_measure(dsa_batch *b) {
  HIP_TRY(dsa::launch_measure_batch(b->d_out, b->d_partials, b->d_chains, b->d_obs, b->nobs, b->d_obs_first,
                                    b->d_obs_n, (uint32_t)b->chains.size(), b->stream));
  return 0;
}

}  // namespace

extern "C" {

int dsa_batch_create(dsa_batch **out) {
  if (!out) return fail(DSA_E_ARG, "dsa_batch_create: out is NULL");
  *out = new dsa_batch();
  return 0;
}

int dsa_batch_add(dsa_batch *b, const dsa_group *g, const dsa_chain_params *p, uint32_t *index) {
  if (!b || !g || !p) return fail(DSA_E_ARG, "dsa_batch_add: NULL argument");
  if (b->final) return fail(DSA_E_STATE, "dsa_batch_add after dsa_batch_finalize");
  if (int rc = dsa::check_params(p)) return rc;
  if (b->D && (int)p->D != b->D) return fail(DSA_E_ARG, "all chains of a batch must have the same D");
  b->D = (int)p->D;
  HostChain hc{};
  hc.grp = const_cast<dsa_group *>(g);
  hc.p = *p;
  hc.G = dsa::make_geom(p->D, p->Nt, p->Nx);
  if (index) *index = (uint32_t)b->chains.size();
  b->chains.push_back(hc);
  return 0;
}

int dsa_batch_finalize(dsa_batch *b) {
  if (!b) return fail(DSA_E_ARG, "dsa_batch_finalize: NULL batch");
  if (b->final) return fail(DSA_E_STATE, "batch already finalized");
  if (b->chains.empty()) return fail(DSA_E_STATE, "empty batch");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(DSA_E_DEVICE, "no HIP device visible");
  HIP_TRY(hipGetDevice(&b->device));
  int lds_max = 0;
  if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device) != hipSuccess) lds_max = 0;
  size_t nlinks = 0, nsites = 0, nms = 0, maxlinks = 0;
  size_t maxpass = 0;
  for (HostChain &hc : b->chains) {
    if (int rc = dsa::upload_tables(hc.grp, b->device, hc.dt)) return rc;
    const dsa::Tables T = dsa::tables_view(hc.grp, hc.dt);
    if (dsa::sweep_lds_bytes(T) > 65536) return fail(DSA_E_RANGE, "group order too large for the LDS tables");
    // the byte-index class update where its tables fit beside the static class
    // values and its decision bound is small, as a lone chain's AUTO choice
    hc.margin = T.cprod ? dsa::class_margin(hc.grp->g, (int)hc.p.D, hc.p.beta0, hc.p.beta1, hc.p.beta2, 0x1p-25) : 0;
    hc.cls = b->kernel != DSA_KERNEL_TABLE && T.cprod &&
             dsa::sweep_cls_lds_bytes(T, dsa::kClsNarrow) + 256 <= (size_t)lds_max &&
             (b->kernel != DSA_KERNEL_AUTO || hc.margin < 1e-3);
    if (b->kernel != DSA_KERNEL_TABLE && b->kernel != DSA_KERNEL_AUTO && !hc.cls)
      return fail(DSA_E_STATE, "no class tables for a chain's group on this device");
    // with the GF(4) products where the group has that form and it fits
    hc.field = hc.cls && T.fmat && (b->kernel == DSA_KERNEL_AUTO || b->kernel == DSA_KERNEL_FIELD) &&
               dsa::sweep_cls_lds_bytes(T, dsa::kClsField) + 256 <= (size_t)lds_max;
    if (b->kernel == DSA_KERNEL_FIELD && !hc.field)
      return fail(DSA_E_STATE, "no GF(4) form of a chain's group in this device's LDS");
    b->lds = std::max(b->lds, hc.field ? dsa::sweep_cls_lds_bytes(T, dsa::kClsField)
                              : hc.cls ? dsa::sweep_cls_lds_bytes(T, dsa::kClsNarrow)
                                       : dsa::sweep_lds_bytes(T));
    std::string err = hc.p.order == DSA_ORDER_COLOUR ? dsa::build_colour_schedule(hc.G, hc.S)
                                                     : dsa::build_reference_schedule(hc.G, hc.S);
    if (!err.empty()) return fail(DSA_E_STATE, err);
    hc.links_off = nlinks;
    hc.sites_off = nsites;
    hc.ms_off = nms;
    const size_t nl = (size_t)hc.G.V * hc.G.D;
    nlinks += nl;
    maxlinks = std::max(maxlinks, nl);
    nsites += hc.S.sites.size();
    if (hc.p.rng == DSA_RNG_MINSTD) nms += hc.G.V;
    maxpass = std::max(maxpass, hc.S.passes.size());
  }
  HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&b->ev0));
  HIP_TRY(hipEventCreate(&b->ev1));
  HIP_TRY(hipMalloc(&b->d_links, nlinks * sizeof(uint16_t)));
  HIP_TRY(hipMalloc(&b->d_sites, nsites * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&b->d_aos, maxlinks * sizeof(int32_t)));
  if (nms) HIP_TRY(hipMalloc(&b->d_mstate, nms * sizeof(uint32_t)));
  const size_t nc = b->chains.size();
  HIP_TRY(hipMalloc(&b->d_acc, nc * sizeof(unsigned long long)));
  HIP_TRY(hipMalloc(&b->d_sweep, sizeof(unsigned long long)));
  HIP_TRY(hipMalloc(&b->d_fall, sizeof(unsigned long long)));
  HIP_TRY(hipMemset(b->d_fall, 0, sizeof(unsigned long long)));
  HIP_TRY(hipMemset(b->d_acc, 0, nc * sizeof(unsigned long long)));
  HIP_TRY(hipMemset(b->d_sweep, 0, sizeof(unsigned long long)));

  std::vector<dsa::BatchChain> desc(nc);
  std::vector<dsa::ObsTask> obs;
  std::vector<uint32_t> ofirst(nc), on(nc);
  for (size_t c = 0; c < nc; ++c) {
    HostChain &hc = b->chains[c];
    dsa::BatchChain &d = desc[c];
    d.links = b->d_links + hc.links_off;
    d.sites = b->d_sites + hc.sites_off;
    d.mstate = hc.p.rng == DSA_RNG_MINSTD ? b->d_mstate + hc.ms_off : nullptr;
    d.T = dsa::tables_view(hc.grp, hc.dt);
    d.G = hc.G;
    d.beta0 = hc.p.beta0;
    d.beta1 = hc.p.beta1;
    d.beta2 = hc.p.beta2;
    d.seed = (uint32_t)hc.p.seed;
    d.rng = hc.p.rng == DSA_RNG_MINSTD ? dsa::kRngMinstd : dsa::kRngPhilox;
    d.sweep_base = 0;
    d.cls = hc.cls ? 1 : 0;
    d.field = hc.field ? 1 : 0;
    d.cmargin = hc.margin * b->margin_scale;
    d.fall = b->d_fall;
    d.bound = 0;
    HIP_TRY(hipMemcpy(b->d_sites + hc.sites_off, hc.S.sites.data(), hc.S.sites.size() * sizeof(uint32_t),
                      hipMemcpyHostToDevice));
    // start: the same kernels and host draws as a lone chain
    if (hc.p.rng == DSA_RNG_MINSTD) {
      std::vector<uint32_t> ms;
      std::vector<int32_t> a;
      dsa::minstd_start(hc.grp->g, hc.p, hc.G.V, ms, a);
      HIP_TRY(hipMemcpy(b->d_mstate + hc.ms_off, ms.data(), ms.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(b->d_aos, a.data(), a.size() * sizeof(int32_t), hipMemcpyHostToDevice));
      HIP_TRY(dsa::launch_pack(d.links, b->d_aos, hc.G, b->stream));
    } else if (hc.p.start == DSA_START_COLD) {
      HIP_TRY(dsa::launch_fill(d.links, hc.G, (uint16_t)hc.grp->g.id, b->stream));
    } else {
      HIP_TRY(dsa::launch_init_philox(d.links, hc.G, hc.grp->g.P, (uint32_t)hc.p.seed, b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    ofirst[c] = (uint32_t)obs.size();
    for (uint32_t s0 = 0; s0 < hc.G.V; s0 += kObsSites)
      obs.push_back(dsa::ObsTask{(uint32_t)c, s0, std::min(kObsSites, hc.G.V - s0)});
    on[c] = (uint32_t)obs.size() - ofirst[c];
  }
  // tasks: at most one per kTaskCandidates[0] links of every pass of every chain
  size_t ntasks_max = 0;
  for (const HostChain &hc : b->chains)
    for (const dsa::Pass &ps : hc.S.passes) ntasks_max += (ps.count + kTaskCandidates[0] - 1) / kTaskCandidates[0];
  b->nobs = (uint32_t)obs.size();
  HIP_TRY(hipMalloc(&b->d_chains, nc * sizeof(dsa::BatchChain)));
  HIP_TRY(hipMalloc(&b->d_tasks, std::max<size_t>(1, ntasks_max) * sizeof(dsa::BatchTask)));
  HIP_TRY(hipMalloc(&b->d_obs, obs.size() * sizeof(dsa::ObsTask)));
  HIP_TRY(hipMalloc(&b->d_obs_first, nc * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&b->d_obs_n, nc * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&b->d_partials, 4 * obs.size() * sizeof(double)));
  HIP_TRY(hipMalloc(&b->d_out, 4 * nc * sizeof(double)));
  HIP_TRY(hipMemcpy(b->d_chains, desc.data(), nc * sizeof(dsa::BatchChain), hipMemcpyHostToDevice));
  b->desc = desc;
  b->acc_seen.assign(nc, 0);
  HIP_TRY(hipMemcpy(b->d_obs, obs.data(), obs.size() * sizeof(dsa::ObsTask), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b->d_obs_first, ofirst.data(), nc * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b->d_obs_n, on.data(), nc * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (int rc = b->task_links ? set_tasks(b, b->task_links) : tune_tasks(b, nlinks, nms)) return rc;
  b->final = true;
  return 0;
}

int dsa_batch_set_task_links(dsa_batch *b, uint32_t links) {
  if (!b) return fail(DSA_E_ARG, "dsa_batch_set_task_links: NULL batch");
  if (b->final) return fail(DSA_E_STATE, "dsa_batch_set_task_links after dsa_batch_finalize");
  bool ok = links == 0;
  for (uint32_t c : kTaskCandidates) ok = ok || links == c;
  if (!ok) return fail(DSA_E_ARG, "dsa_batch_set_task_links: links must be 0, 256, 512 or 1024");
  b->task_links = links;
  return 0;
}

int dsa_batch_task_links(dsa_batch *b, uint32_t *links, double *ms_per_sweep, uint32_t ncand) {
  if (!b || !links) return fail(DSA_E_ARG, "dsa_batch_task_links: NULL argument");
  if (!b->final) return fail(DSA_E_STATE, "batch not finalized");
  *links = b->block;
  if (ms_per_sweep)
    for (uint32_t k = 0; k < ncand; ++k) ms_per_sweep[k] = k < b->tune_ms.size() ? b->tune_ms[k] : 0.0;
  return 0;
}

void dsa_batch_destroy(dsa_batch *b) {
  if (!b) return;
  if (b->final) {
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
  }
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->graph_def) (void)hipGraphDestroy(b->graph_def);
  for (void *p : {(void *)b->d_links, (void *)b->d_sites, (void *)b->d_mstate, (void *)b->d_chains,
                  (void *)b->d_tasks, (void *)b->d_obs, (void *)b->d_obs_first, (void *)b->d_obs_n,
                  (void *)b->d_acc, (void *)b->d_sweep, (void *)b->d_partials, (void *)b->d_out, (void *)b->d_aos,
                  (void *)b->d_fall})
    if (p) (void)hipFree(p);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int dsa_batch_set_kernel(dsa_batch *b, int32_t kernel, double margin_scale) {
  if (!b) return fail(DSA_E_ARG, "dsa_batch_set_kernel: NULL batch");
  if (b->final) return fail(DSA_E_STATE, "dsa_batch_set_kernel after dsa_batch_finalize");
  if (kernel != DSA_KERNEL_AUTO && kernel != DSA_KERNEL_TABLE && kernel != DSA_KERNEL_CLASS &&
      kernel != DSA_KERNEL_CLASS_NARROW && kernel != DSA_KERNEL_FIELD)
    return fail(DSA_E_ARG, "dsa_batch_set_kernel: unknown kernel " + std::to_string(kernel));
  if (!(margin_scale >= 1.0) || !std::isfinite(margin_scale))
    return fail(DSA_E_ARG, "dsa_batch_set_kernel: margin_scale must be finite and >= 1");
  b->kernel = kernel;
  b->margin_scale = margin_scale;
  return 0;
}

int dsa_batch_exact_decisions(dsa_batch *b, uint64_t *n) {
  if (!b || !n) return fail(DSA_E_ARG, "dsa_batch_exact_decisions: NULL argument");
  if (!b->final) return fail(DSA_E_STATE, "batch not finalized");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));
  unsigned long long v = 0;
  HIP_TRY(hipMemcpy(&v, b->d_fall, sizeof v, hipMemcpyDeviceToHost));
  *n = v;
  return 0;
}

int dsa_batch_sweep(dsa_batch *b, uint32_t n) {
  if (!b) return fail(DSA_E_ARG, "dsa_batch_sweep: NULL batch");
  if (!b->final) return fail(DSA_E_STATE, "batch not finalized");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipEventRecord(b->ev0, b->stream));
  if (int rc = enqueue_sweeps(b, n)) return rc;
  HIP_TRY(hipEventRecord(b->ev1, b->stream));
  HIP_TRY(hipEventSynchronize(b->ev1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  b->sweeps += n;
  return choose_bounds(b, n);
}

int dsa_batch_measure(dsa_batch *b, double *out) {
  if (!b || !out) return fail(DSA_E_ARG, "dsa_batch_measure: NULL argument");
  if (!b->final) return fail(DSA_E_STATE, "batch not finalized");
  HIP_TRY(hipSetDevice(b->device));
  if (int rc = enqueue_measure(b)) return rc;
  HIP_TRY(hipMemcpyAsync(out, b->d_out, 4 * b->chains.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return 0;
}

int dsa_batch_run_cycles(dsa_batch *b, uint32_t ncycles, uint32_t sweeps_per_cycle, double *out) {
  if (!b || !out || !sweeps_per_cycle) return fail(DSA_E_ARG, "dsa_batch_run_cycles: bad argument");
  if (!b->final) return fail(DSA_E_STATE, "batch not finalized");
  HIP_TRY(hipSetDevice(b->device));
  if (!b->graph || b->graph_sweeps != sweeps_per_cycle) {
    if (b->graph) HIP_TRY(hipGraphExecDestroy(b->graph));
    if (b->graph_def) HIP_TRY(hipGraphDestroy(b->graph_def));
    b->graph = nullptr;
    b->graph_def = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_sweeps(b, sweeps_per_cycle);
    if (!rc) rc = enqueue_measure(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return fail(DSA_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    b->graph_def = g;
    HIP_TRY(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
    b->graph_sweeps = sweeps_per_cycle;
  }
  const size_t nc = b->chains.size();
  HIP_TRY(hipEventRecord(b->ev0, b->stream));
  for (uint32_t k = 0; k < ncycles; ++k) {
    HIP_TRY(hipGraphLaunch(b->graph, b->stream));
    HIP_TRY(hipMemcpyAsync(out + k * nc * 4, b->d_out, 4 * nc * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  }
  HIP_TRY(hipEventRecord(b->ev1, b->stream));
  HIP_TRY(hipEventSynchronize(b->ev1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  b->sweeps += (uint64_t)ncycles * sweeps_per_cycle;
  return choose_bounds(b, ncycles * sweeps_per_cycle);
}

int dsa_batch_info(dsa_batch *b, uint32_t *nchains, uint32_t *npasses, uint32_t *nblocks_per_sweep) {
  if (!b) return fail(DSA_E_ARG, "dsa_batch_info: NULL batch");
  if (nchains) *nchains = (uint32_t)b->chains.size();
  if (npasses) *npasses = (uint32_t)b->pass_off.size();
  if (nblocks_per_sweep) {
    uint32_t s = 0;
    for (uint32_t n : b->pass_n) s += n;
    *nblocks_per_sweep = s;
  }
  return 0;
}

int dsa_batch_get_links(dsa_batch *b, uint32_t chain, int32_t *a) {
  if (!b || !a) return fail(DSA_E_ARG, "dsa_batch_get_links: NULL argument");
  if (!b->final) return fail(DSA_E_STATE, "batch not finalized");
  if (chain >= b->chains.size()) return fail(DSA_E_ARG, "chain index out of range");
  HIP_TRY(hipSetDevice(b->device));
  const HostChain &hc = b->chains[chain];
  const size_t n = (size_t)hc.G.V * hc.G.D;
  HIP_TRY(dsa::launch_unpack(b->d_aos, b->d_links + hc.links_off, hc.G, b->stream));
  HIP_TRY(hipMemcpyAsync(a, b->d_aos, n * sizeof(int32_t), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return 0;
}

int dsa_batch_counters(dsa_batch *b, uint64_t *hit, uint64_t *acc) {
  if (!b) return fail(DSA_E_ARG, "dsa_batch_counters: NULL batch");
  if (!b->final) return fail(DSA_E_STATE, "batch not finalized");
  HIP_TRY(hipSetDevice(b->device));
  const size_t nc = b->chains.size();
  std::vector<unsigned long long> a(nc);
  HIP_TRY(hipMemcpy(a.data(), b->d_acc, nc * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for (size_t c = 0; c < nc; ++c) {
    const HostChain &hc = b->chains[c];
    if (hit) hit[c] = b->sweeps * hc.G.V * hc.G.D * (uint64_t)dsa::kNHit;
    if (acc) acc[c] = a[c];
  }
  return 0;
}

int dsa_batch_last_ms(dsa_batch *b, double *ms) {
  if (!b || !ms) return fail(DSA_E_ARG, "dsa_batch_last_ms: NULL argument");
  *ms = b->last_ms;
  return 0;
}

}  // extern "C"
