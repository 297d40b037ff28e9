// llr_engine.cpp -- device-resident LLR chains (include/dsa.h, dsa_llr_*):
// admission into slots, the active chains split into three groups of about
// equal links, each group's sweeps launched on a stream of its own (one
// proposal launch and one decision launch per pass over the group's chains,
// then the group's sweep end and energy resync; the groups meet once a step),
// the per-chain drive-in / Robbins-Monro state machine on the device, and a
// continuous-batching serving loop for streams of window requests.
//
// The decision launch walks each chain's proposals in order with one
// wavefront per chain, which leaves most of the device idle; with three
// groups another group's proposal launch runs beside it.  Chains are
// independent, so the grouping does not change any chain's results.  The
// engine has no stream besides the groups' (set-up and copies go to the
// first group's): HIP gives a process a few hardware queues (four by
// default) and hands them to streams as they are created, and two groups
// whose streams share a queue wait for each other's launches.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <chrono>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "dsa_math.hpp"
#include "internal.hpp"
#include "schedule.hpp"

using dsa::fail;

namespace {

constexpr uint32_t kTaskLinks = 256;
constexpr uint32_t kTaskLinksSmall = 64;  // for passes below kSmallPass links in all
constexpr uint64_t kSmallPass = 32768;
constexpr uint32_t kObsSites = 1024;
constexpr int kGroups = 3;

// What every chain of one lattice shape shares: the colour schedule (a
// 16^4 one takes tens of milliseconds to build and verify on the host), the
// position of each scheduled link in the link array, and their device
// copies, which the kernels only read.  Built at the first admission of the
// shape and kept until the engine is destroyed.
struct Layout {
  std::array<unsigned, 3> shape{};  // D, Nt, Nx
  dsa::Schedule S;
  uint32_t *sites = nullptr, *pos = nullptr;
  dsa::PassRange *passes = nullptr;
  ~Layout() {
    for (void *p : {(void *)sites, (void *)pos, (void *)passes})
      if (p) (void)hipFree(p);
  }
};

struct Slot {
  bool used = false;
  dsa_group *grp = nullptr;
  dsa_llr_params p{};
  dsa::Geom G{};
  std::shared_ptr<const Layout> lay;
  dsa::DeviceTables dt;
  uint16_t *links = nullptr, *old = nullptr;
  long long *dE1 = nullptr, *dE2 = nullptr;
  uint8_t *nacc = nullptr;
  uint64_t bytes = 0;
};

// links, previous values, site lists, link positions, two energy changes and
// a hit count per link (the site lists and positions are shared by the chains
// of one shape, so this overstates what a chain adds once its shape is in)
uint64_t footprint(const dsa_llr_params &p) {
  uint64_t V = p.Nt;
  for (uint32_t d = 1; d < p.D; ++d) V *= p.Nx;
  const uint64_t nl = V * p.D;
  return nl * (2 + 2 + 4 + 4 + 8 + 8 + 1) + 64 * sizeof(dsa::PassRange) + sizeof(dsa::LlrChain) +
         sizeof(dsa::LlrState);
}

}  // namespace

struct dsa_llr {
  int D = 0;
  int device = 0;
  uint32_t cap = 0;
  uint64_t budget = 0, used_bytes = 0;
  std::vector<Slot> slots;
  hipStream_t stream = nullptr;         // = gstream[0]: set-up, a step's first resync, copies
  hipStream_t gstream[kGroups] = {};    // each group's passes
  hipEvent_t fork = nullptr, stagger[kGroups] = {}, join[kGroups] = {};
  dsa::LlrChain *d_chains = nullptr;
  dsa::LlrState *d_state = nullptr;
  dsa::BatchTask *d_tasks = nullptr;
  dsa::ObsTask *d_obs = nullptr;
  size_t tasks_cap = 0, obs_cap = 0;
  // group g's tasks of pass q: [pass_off[g][q], + pass_n[g][q]) of d_tasks;
  // its chains: d_idx[idx_off[g] .. + idx_n[g])
  std::vector<uint32_t> pass_off[kGroups], pass_n[kGroups];
  uint32_t *d_idx = nullptr;
  size_t idx_cap = 0;
  uint32_t idx_off[kGroups] = {}, idx_n[kGroups] = {};
  // the energy-resync tasks, group by group: group g's are
  // d_obs[obs_off[g] .. + obs_n[g]), all of them d_obs[0 .. nobs)
  uint32_t obs_off[kGroups] = {}, obs_n[kGroups] = {};
  uint32_t nobs = 0;
  size_t lds = 0, lds_max = 0;
  int kernel = DSA_KERNEL_AUTO;  // for chains added from now on
  double margin_scale = 1.0;
  bool dirty = true;
  std::vector<dsa::LlrState> host_state;
  std::vector<std::shared_ptr<const Layout>> layouts;
};

namespace {

int rebuild_tasks(dsa_llr *L) {
  std::vector<dsa::BatchTask> tasks;
  std::vector<dsa::ObsTask> obs;
  // the groups: chains by links, largest first, each to the group with
  // fewer links so far
  std::vector<uint32_t> act;
  for (uint32_t c = 0; c < L->slots.size(); ++c)
    if (L->slots[c].used) act.push_back(c);
  auto links_of = [&](uint32_t c) { return (uint64_t)L->slots[c].G.V * L->slots[c].G.D; };
  std::stable_sort(act.begin(), act.end(), [&](uint32_t x, uint32_t y) { return links_of(x) > links_of(y); });
  std::vector<uint32_t> mem[kGroups];
  uint64_t load[kGroups] = {};
  for (uint32_t c : act) {
    int g = 0;
    for (int h = 1; h < kGroups; ++h)
      if (load[h] < load[g]) g = h;
    mem[g].push_back(c);
    load[g] += links_of(c);
  }
  std::vector<uint32_t> idx;
  for (int g = 0; g < kGroups; ++g) {
    std::sort(mem[g].begin(), mem[g].end());
    L->idx_off[g] = (uint32_t)idx.size();
    L->idx_n[g] = (uint32_t)mem[g].size();
    idx.insert(idx.end(), mem[g].begin(), mem[g].end());
    size_t maxpass = 0;
    for (uint32_t c : mem[g]) maxpass = std::max(maxpass, L->slots[c].lay->S.passes.size());
    L->pass_off[g].clear();
    L->pass_n[g].clear();
    for (size_t q = 0; q < maxpass; ++q) {
      L->pass_off[g].push_back((uint32_t)tasks.size());
      // a pass that would not fill the device is cut into blocks of fewer
      // links, so that more lanes share each link's random draws
      uint64_t total = 0;
      for (uint32_t c : mem[g])
        if (q < L->slots[c].lay->S.passes.size()) total += L->slots[c].lay->S.passes[q].count;
      const uint32_t per = total < kSmallPass ? kTaskLinksSmall : kTaskLinks;
      for (uint32_t c : mem[g]) {
        const Slot &s = L->slots[c];
        if (q >= s.lay->S.passes.size()) continue;
        const dsa::Pass &ps = s.lay->S.passes[q];
        for (uint32_t o = 0; o < ps.count; o += per)
          tasks.push_back(dsa::BatchTask{c, ps.offset + o, std::min(per, ps.count - o), ps.mu});
      }
      L->pass_n[g].push_back((uint32_t)tasks.size() - L->pass_off[g].back());
    }
  }
  for (int g = 0; g < kGroups; ++g) {
    L->obs_off[g] = (uint32_t)obs.size();
    for (uint32_t c : mem[g]) {
      const Slot &s = L->slots[c];
      for (uint32_t s0 = 0; s0 < s.G.V; s0 += kObsSites)
        obs.push_back(dsa::ObsTask{c, s0, std::min(kObsSites, s.G.V - s0)});
    }
    L->obs_n[g] = (uint32_t)obs.size() - L->obs_off[g];
  }
  if (tasks.size() > L->tasks_cap) {
    if (L->d_tasks) HIP_TRY(hipFree(L->d_tasks));
    L->d_tasks = nullptr;
    L->tasks_cap = tasks.size() * 2;
    HIP_TRY(hipMalloc(&L->d_tasks, L->tasks_cap * sizeof(dsa::BatchTask)));
  }
  if (idx.size() > L->idx_cap) {
    if (L->d_idx) HIP_TRY(hipFree(L->d_idx));
    L->d_idx = nullptr;
    L->idx_cap = std::max<size_t>(L->slots.size(), 1);
    HIP_TRY(hipMalloc(&L->d_idx, L->idx_cap * sizeof(uint32_t)));
  }
  if (obs.size() > L->obs_cap) {
    if (L->d_obs) HIP_TRY(hipFree(L->d_obs));
    L->d_obs = nullptr;
    L->obs_cap = obs.size() * 2;
    HIP_TRY(hipMalloc(&L->d_obs, L->obs_cap * sizeof(dsa::ObsTask)));
  }
  if (!tasks.empty())
    HIP_TRY(hipMemcpyAsync(L->d_tasks, tasks.data(), tasks.size() * sizeof(dsa::BatchTask), hipMemcpyHostToDevice,
                           L->stream));
  if (!idx.empty())
    HIP_TRY(hipMemcpyAsync(L->d_idx, idx.data(), idx.size() * sizeof(uint32_t), hipMemcpyHostToDevice, L->stream));
  if (!obs.empty())
    HIP_TRY(hipMemcpyAsync(L->d_obs, obs.data(), obs.size() * sizeof(dsa::ObsTask), hipMemcpyHostToDevice,
                           L->stream));
  HIP_TRY(hipStreamSynchronize(L->stream));
  L->nobs = (uint32_t)obs.size();
  L->dirty = false;
  return 0;
}

void release(Slot &s) {
  for (void *p : {(void *)s.links, (void *)s.old, (void *)s.dE1, (void *)s.dE2, (void *)s.nacc})
    if (p) (void)hipFree(p);
  s = Slot();
}

int step(dsa_llr *L, uint32_t n) {
  if (L->dirty)
    if (int rc = rebuild_tasks(L)) return rc;
  const uint32_t nc = (uint32_t)L->slots.size();
  HIP_TRY(dsa::launch_llr_resync(L->d_chains, L->d_state, nullptr, nc, L->d_obs, L->nobs, L->stream));
  HIP_TRY(hipEventRecord(L->fork, L->stream));
  // Each group runs its n sweeps on its own stream: 32 (or so) passes, each
  // a proposal and a decision launch, then the sweep's end and the energy
  // resync of the group's own chains, which no other group's kernel reads
  // or writes.  A group starts once the group before it has done its first
  // proposal launch, so that a group's decision launch tends to run beside
  // another group's proposal launch; the groups meet only at the end.
  int prev[kGroups], last = -1;  // the non-empty group before each one
  for (int g = 0; g < kGroups; ++g) {
    prev[g] = last;
    if (!L->pass_off[g].empty()) last = g;
  }
  size_t np = 0;
  for (int g = 0; g < kGroups; ++g) np = std::max(np, L->pass_off[g].size());
  for (uint32_t it = 0; it < n; ++it) {
    for (size_t q = 0; q < np; ++q)
      for (int g = 0; g < kGroups; ++g) {
        if (q >= L->pass_off[g].size()) continue;
        const dsa::BatchTask *tk = L->d_tasks + L->pass_off[g][q];
        const bool first = it == 0 && q == 0;
        if (first) HIP_TRY(hipStreamWaitEvent(L->gstream[g], prev[g] < 0 ? L->fork : L->stagger[prev[g]], 0));
        HIP_TRY(dsa::launch_llr_pass(L->D, L->d_chains, L->d_state, tk, L->pass_n[g][q], L->lds, L->gstream[g]));
        if (first) HIP_TRY(hipEventRecord(L->stagger[g], L->gstream[g]));
        HIP_TRY(dsa::launch_llr_decide(L->d_chains, L->d_state, L->d_idx + L->idx_off[g], L->idx_n[g], (uint32_t)q,
                                       L->gstream[g]));
      }
    for (int g = 0; g < kGroups; ++g) {
      if (L->pass_off[g].empty()) continue;
      const uint32_t *ix = L->d_idx + L->idx_off[g];
      HIP_TRY(dsa::launch_llr_sweep_end(L->d_chains, L->d_state, ix, L->idx_n[g], L->gstream[g]));
      HIP_TRY(dsa::launch_llr_resync(L->d_chains, L->d_state, ix, L->idx_n[g], L->d_obs + L->obs_off[g],
                                     L->obs_n[g], L->gstream[g]));
    }
  }
  for (int g = 0; g < kGroups; ++g) {
    HIP_TRY(hipEventRecord(L->join[g], L->gstream[g]));
    HIP_TRY(hipStreamWaitEvent(L->stream, L->join[g], 0));
  }
  HIP_TRY(hipStreamSynchronize(L->stream));
  return 0;
}

int pull_state(dsa_llr *L) {
  L->host_state.resize(L->slots.size());
  HIP_TRY(hipMemcpy(L->host_state.data(), L->d_state, L->slots.size() * sizeof(dsa::LlrState),
                    hipMemcpyDeviceToHost));
  return 0;
}

}  // namespace

extern "C" {

int dsa_llr_footprint(const dsa_llr_params *p, uint64_t *bytes) {
  if (!p || !bytes) return fail(DSA_E_ARG, "dsa_llr_footprint: NULL argument");
  *bytes = footprint(*p);
  return 0;
}

int dsa_llr_create(uint32_t D, uint32_t max_chains, uint64_t hbm_budget, dsa_llr **out) {
  if (!out || !max_chains || D < 2 || D > (uint32_t)dsa::kDMax) return fail(DSA_E_ARG, "dsa_llr_create: bad argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(DSA_E_DEVICE, "no HIP device visible");
  auto *L = new dsa_llr();
  L->D = (int)D;
  L->cap = max_chains;
  L->budget = hbm_budget;
  L->slots.resize(max_chains);
  auto bail = [&](int rc) {
    dsa_llr_destroy(L);
    return rc;
  };
  if (hipGetDevice(&L->device) != hipSuccess) return bail(fail(DSA_E_DEVICE, "hipGetDevice failed"));
  int lds_max = 0;
  if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, L->device) != hipSuccess)
    lds_max = 0;
  L->lds_max = (size_t)lds_max;
  bool streams_ok = hipEventCreateWithFlags(&L->fork, hipEventDisableTiming) == hipSuccess;
  for (int g = 0; g < kGroups && streams_ok; ++g)
    streams_ok = hipStreamCreateWithFlags(&L->gstream[g], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&L->join[g], hipEventDisableTiming) == hipSuccess &&
                 hipEventCreateWithFlags(&L->stagger[g], hipEventDisableTiming) == hipSuccess;
  L->stream = L->gstream[0];
  if (!streams_ok || hipMalloc(&L->d_chains, max_chains * sizeof(dsa::LlrChain)) != hipSuccess ||
      hipMalloc(&L->d_state, max_chains * sizeof(dsa::LlrState)) != hipSuccess)
    return bail(fail(DSA_E_DEVICE, "allocation failed"));
  std::vector<dsa::LlrState> idle(max_chains);
  memset(idle.data(), 0, idle.size() * sizeof(dsa::LlrState));
  for (auto &s : idle) s.phase = dsa::kLlrIdle;
  std::vector<dsa::LlrChain> zero(max_chains);
  memset(zero.data(), 0, zero.size() * sizeof(dsa::LlrChain));
  if (hipMemcpy(L->d_state, idle.data(), idle.size() * sizeof(dsa::LlrState), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(L->d_chains, zero.data(), zero.size() * sizeof(dsa::LlrChain), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(DSA_E_DEVICE, "upload failed"));
  *out = L;
  return 0;
}

void dsa_llr_destroy(dsa_llr *L) {
  if (!L) return;
  (void)hipSetDevice(L->device);
  if (L->stream) (void)hipStreamSynchronize(L->stream);
  for (int g = 0; g < kGroups; ++g)
    if (L->gstream[g]) (void)hipStreamSynchronize(L->gstream[g]);
  for (Slot &s : L->slots) release(s);
  for (void *p : {(void *)L->d_chains, (void *)L->d_state, (void *)L->d_tasks, (void *)L->d_obs, (void *)L->d_idx})
    if (p) (void)hipFree(p);
  for (int g = 0; g < kGroups; ++g) {
    if (L->gstream[g]) (void)hipStreamDestroy(L->gstream[g]);
    if (L->join[g]) (void)hipEventDestroy(L->join[g]);
    if (L->stagger[g]) (void)hipEventDestroy(L->stagger[g]);
  }
  if (L->fork) (void)hipEventDestroy(L->fork);
  delete L;
}

int dsa_llr_add(dsa_llr *L, const dsa_group *gc, const dsa_llr_params *p, const int32_t *init_links,
                uint32_t *slot_out) {
  if (!L || !gc || !p) return fail(DSA_E_ARG, "dsa_llr_add: NULL argument");
  if ((int)p->D != L->D) return fail(DSA_E_ARG, "chain dimension differs from the engine's");
  dsa_chain_params cp{p->D, p->Nt, p->Nx, 0, 0, 0, p->seed, p->start, DSA_RNG_PHILOX, DSA_ORDER_COLOUR};
  if (int rc = dsa::check_params(&cp)) return rc;
  if (p->axis != 1 && p->axis != 2 && p->axis != 3) return fail(DSA_E_ARG, "axis must be 1, 2 or 3");
  if (!(p->hw > 0) || !p->K) return fail(DSA_E_ARG, "need hw > 0 and K > 0");
  if (p->axis == 3 && !(p->hw2 > 0)) return fail(DSA_E_ARG, "axis 3 needs hw2 > 0");
  const uint64_t bytes = footprint(*p);
  if (L->budget && L->used_bytes + bytes > L->budget) return fail(DSA_E_RANGE, "HBM budget exhausted");
  uint32_t c = 0;
  while (c < L->slots.size() && L->slots[c].used) ++c;
  if (c == L->slots.size()) return fail(DSA_E_RANGE, "no free slot");
  HIP_TRY(hipSetDevice(L->device));
  dsa_group *g = const_cast<dsa_group *>(gc);
  Slot s;
  s.grp = g;
  s.p = *p;
  s.G = dsa::make_geom(p->D, p->Nt, p->Nx);
  s.bytes = bytes;
  if (int rc = dsa::upload_tables(g, L->device, s.dt)) return rc;
  const dsa::Tables T = dsa::tables_view(g, s.dt);
  if (dsa::llr_lds_bytes(T, false) > 65536) return fail(DSA_E_RANGE, "group order too large for the LDS tables");
  // the class path where the byte class tables exist and fit beside the
  // kernel's static arrays
  const bool cls = L->kernel != DSA_KERNEL_TABLE && T.cprod &&
                   dsa::llr_lds_bytes(T, true) + dsa::llr_pass_static_lds() <= L->lds_max;
  if ((L->kernel == DSA_KERNEL_CLASS || L->kernel == DSA_KERNEL_CLASS_NARROW || L->kernel == DSA_KERNEL_FIELD) &&
      !cls)
    return fail(DSA_E_STATE, "no class tables for this group on this device");
  // with the GF(4) products where the group has that form and it fits too
  const bool field = cls && T.fmat && (L->kernel == DSA_KERNEL_AUTO || L->kernel == DSA_KERNEL_FIELD) &&
                     dsa::llr_lds_bytes(T, true, true) + dsa::llr_pass_static_lds() <= L->lds_max;
  if (L->kernel == DSA_KERNEL_FIELD && !field)
    return fail(DSA_E_STATE, "no GF(4) form of this group in this device's LDS");
  const std::array<unsigned, 3> shape{p->D, p->Nt, p->Nx};
  for (const auto &l : L->layouts)
    if (l->shape == shape) s.lay = l;
  if (!s.lay) {
    auto lay = std::make_shared<Layout>();
    lay->shape = shape;
    std::string err = dsa::build_colour_schedule(s.G, lay->S);
    if (!err.empty()) return fail(DSA_E_STATE, err);
    const dsa::Schedule &S = lay->S;
    const size_t ns = S.sites.size();
    std::vector<dsa::PassRange> pr;
    for (const dsa::Pass &ps : S.passes) pr.push_back(dsa::PassRange{ps.offset, ps.count, ps.mu, 0});
    // index of each scheduled link in the per-direction, per-parity link array
    std::vector<uint32_t> lpos(ns);
    for (const dsa::Pass &ps : S.passes)
      for (uint32_t q = ps.offset; q < ps.offset + ps.count; ++q) {
        const uint32_t site = S.sites[q];
        uint32_t par = 0, r = site;
        for (int d = 0; d < s.G.D; ++d) {
          par += r % s.G.ext[d];
          r /= s.G.ext[d];
        }
        const uint32_t at = s.G.eo ? ((par & 1u) ? s.G.halfV : 0u) + (site >> 1) : site;
        lpos[q] = (uint32_t)ps.mu * s.G.V + at;
      }
    if (hipMalloc(&lay->sites, ns * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&lay->pos, ns * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&lay->passes, pr.size() * sizeof(dsa::PassRange)) != hipSuccess)
      return fail(DSA_E_DEVICE, "layout allocation failed");
    if (hipMemcpy(lay->sites, S.sites.data(), ns * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(lay->pos, lpos.data(), ns * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(lay->passes, pr.data(), pr.size() * sizeof(dsa::PassRange), hipMemcpyHostToDevice) != hipSuccess)
      return fail(DSA_E_DEVICE, "layout upload failed");
    L->layouts.push_back(lay);
    s.lay = lay;
  }
  const size_t nl = (size_t)s.G.V * s.G.D;
  auto cleanup = [&](int rc) {
    release(s);
    return rc;
  };
  const size_t ns = s.lay->S.sites.size();
  if (hipMalloc(&s.links, nl * sizeof(uint16_t)) != hipSuccess || hipMalloc(&s.old, ns * sizeof(uint16_t)) != hipSuccess ||
      hipMalloc(&s.dE1, ns * sizeof(long long)) != hipSuccess ||
      hipMalloc(&s.dE2, ns * sizeof(long long)) != hipSuccess || hipMalloc(&s.nacc, ns) != hipSuccess)
    return cleanup(fail(DSA_E_DEVICE, "chain allocation failed"));
  hipError_t e;
  int32_t *tmp = nullptr;
  if (init_links) {
    for (size_t l = 0; l < nl; ++l)
      if (init_links[l] < 0 || (uint32_t)init_links[l] >= g->g.P) return cleanup(fail(DSA_E_ARG, "link out of range"));
    e = hipMalloc(&tmp, nl * sizeof(int32_t));
    if (e == hipSuccess) e = hipMemcpyAsync(tmp, init_links, nl * sizeof(int32_t), hipMemcpyHostToDevice, L->stream);
    if (e == hipSuccess) e = dsa::launch_pack(s.links, tmp, s.G, L->stream);
  } else if (p->start == DSA_START_COLD) {
    e = dsa::launch_fill(s.links, s.G, (uint16_t)g->g.id, L->stream);
  } else {
    e = dsa::launch_init_philox(s.links, s.G, g->g.P, (uint32_t)p->seed, L->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
  if (tmp) (void)hipFree(tmp);
  if (e != hipSuccess) return cleanup(fail(DSA_E_DEVICE, std::string("chain start: ") + hipGetErrorString(e)));

  dsa::LlrChain d{};
  d.base.links = s.links;
  d.base.sites = s.lay->sites;
  d.base.mstate = nullptr;
  d.base.T = T;
  d.base.G = s.G;
  d.base.seed = (uint32_t)p->seed;
  d.base.rng = dsa::kRngPhilox;
  d.old = s.old;
  d.dE1 = s.dE1;
  d.dE2 = s.dE2;
  d.nacc = s.nacc;
  d.pos = s.lay->pos;
  d.passes = s.lay->passes;
  d.axis = p->axis;
  d.bg = p->bg;
  d.lo = dsa::to_fix(p->Ewin - p->hw);
  d.hi = dsa::to_fix(p->Ewin + p->hw);
  d.Ewin = p->Ewin;
  d.W = 2 * p->hw;
  if (p->axis == 3) {
    d.lo2 = dsa::to_fix(p->Ewin2 - p->hw2);
    d.hi2 = dsa::to_fix(p->Ewin2 + p->hw2);
    d.Ewin2 = p->Ewin2;
    d.W2 = 2 * p->hw2;
  } else {
    d.W2 = 1.0;  // unused
  }
  d.K = p->K;
  d.NRM = p->NRM;
  d.tol = p->tol;
  d.cap = p->cap ? p->cap : 20000;
  d.npass = (uint32_t)s.lay->S.passes.size();
  d.cls = cls ? 1 : 0;
  d.field = field ? 1 : 0;
  d.cmscale = L->margin_scale;
  if (cls)  // byte tables: class values in units of 2^-24, half a unit of quantisation
    dsa::class_margin_terms(g->g, (int)p->D, 0x1p-25, d.ce1, d.ce2, d.cf1, d.cf2, d.cr1, d.cr2);
  dsa::LlrState st{};
  st.a = p->a0;
  st.a2 = p->axis == 3 ? p->a20 : 0.0;
  st.phase = dsa::kLlrSeed;
  st.resync = 2;  // initial energies: computed, not compared
  if (hipMemcpy(L->d_chains + c, &d, sizeof d, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(L->d_state + c, &st, sizeof st, hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(DSA_E_DEVICE, "descriptor upload failed"));
  s.used = true;
  L->slots[c] = s;
  L->used_bytes += bytes;
  L->lds = std::max(L->lds, dsa::llr_lds_bytes(T, cls, field));
  L->dirty = true;
  if (slot_out) *slot_out = c;
  return 0;
}

int dsa_llr_set_kernel(dsa_llr *L, int32_t kernel, double margin_scale) {
  if (!L) return fail(DSA_E_ARG, "dsa_llr_set_kernel: NULL engine");
  if (kernel != DSA_KERNEL_AUTO && kernel != DSA_KERNEL_TABLE && kernel != DSA_KERNEL_CLASS &&
      kernel != DSA_KERNEL_CLASS_NARROW && kernel != DSA_KERNEL_FIELD)
    return fail(DSA_E_ARG, "dsa_llr_set_kernel: unknown kernel " + std::to_string(kernel));
  if (!(margin_scale >= 1.0) || !std::isfinite(margin_scale))
    return fail(DSA_E_ARG, "dsa_llr_set_kernel: margin_scale must be finite and >= 1");
  L->kernel = kernel;
  L->margin_scale = margin_scale;
  return 0;
}

int dsa_llr_remove(dsa_llr *L, uint32_t slot) {
  if (!L || slot >= L->slots.size() || !L->slots[slot].used) return fail(DSA_E_ARG, "dsa_llr_remove: bad slot");
  HIP_TRY(hipSetDevice(L->device));
  HIP_TRY(hipStreamSynchronize(L->stream));
  dsa::LlrState idle{};
  idle.phase = dsa::kLlrIdle;
  HIP_TRY(hipMemcpy(L->d_state + slot, &idle, sizeof idle, hipMemcpyHostToDevice));
  L->used_bytes -= L->slots[slot].bytes;
  release(L->slots[slot]);
  L->dirty = true;
  return 0;
}

int dsa_llr_step(dsa_llr *L, uint32_t nsweeps) {
  if (!L) return fail(DSA_E_ARG, "dsa_llr_step: NULL engine");
  HIP_TRY(hipSetDevice(L->device));
  return step(L, nsweeps);
}

int dsa_llr_status_get(dsa_llr *L, uint32_t slot, dsa_llr_status *out) {
  if (!L || !out || slot >= L->slots.size() || !L->slots[slot].used)
    return fail(DSA_E_ARG, "dsa_llr_status_get: bad slot");
  HIP_TRY(hipSetDevice(L->device));
  dsa::LlrState s;
  HIP_TRY(hipMemcpy(&s, L->d_state + slot, sizeof s, hipMemcpyDeviceToHost));
  out->a = s.a;
  out->S1 = dsa::from_fix(s.S1);
  out->S2 = dsa::from_fix(s.S2);
  out->last_da = s.last_da;
  out->maxdrift = s.maxdrift;
  out->drift1 = s.drift1;
  out->drift2 = s.drift2;
  out->phase = s.phase;
  out->iterations = s.m;
  out->seed_sweeps = s.seed_sweeps;
  out->reached = s.reached;
  out->sweeps = s.sweep;
  out->hit = s.hit;
  out->acc = s.acc;
  out->a2 = s.a2;
  return 0;
}

int dsa_llr_get_links(dsa_llr *L, uint32_t slot, int32_t *a) {
  if (!L || !a || slot >= L->slots.size() || !L->slots[slot].used)
    return fail(DSA_E_ARG, "dsa_llr_get_links: bad slot");
  HIP_TRY(hipSetDevice(L->device));
  const Slot &s = L->slots[slot];
  const size_t nl = (size_t)s.G.V * s.G.D;
  int32_t *tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, nl * sizeof(int32_t)));
  hipError_t e = dsa::launch_unpack(tmp, s.links, s.G, L->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(a, tmp, nl * sizeof(int32_t), hipMemcpyDeviceToHost, L->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(DSA_E_DEVICE, std::string("get links: ") + hipGetErrorString(e));
  return 0;
}

// Serving: requests arrive at arrival_s[r] seconds after the call starts; each
// is admitted into a free slot once it has arrived and its footprint fits the
// budget (in arrival order: a request that must wait holds back the ones
// behind it), advanced `chunk` sweeps per round together with every other
// active window, and retired as soon as its Robbins-Monro iteration converges
// (or its drive-in fails).  done_s[r] is the retirement time, status[r] its
// final state; a request whose footprint exceeds the whole budget, or whose
// parameters are refused, is retired on arrival with phase DSA_LLR_REJECTED.
int dsa_llr_serve(dsa_llr *L, const dsa_group *const *groups, const dsa_llr_params *req, const double *arrival_s,
                  uint32_t nreq, uint32_t chunk, double *admit_s, double *done_s, dsa_llr_status *status) {
  if (!L || !groups || !req || !arrival_s || !done_s || !status || !nreq)
    return fail(DSA_E_ARG, "dsa_llr_serve: NULL argument");
  if (!chunk) chunk = 1;
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  auto now = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };
  std::vector<int> slot_of(nreq, -1), req_of(L->slots.size(), -1);
  std::vector<uint32_t> order(nreq);
  for (uint32_t r = 0; r < nreq; ++r) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return arrival_s[x] < arrival_s[y]; });
  size_t next = 0;
  uint32_t finished = 0, active = 0;
  // a request that no retirement can make room for (its footprint above the
  // budget, or parameters the engine refuses) retires at once as REJECTED
  auto reject = [&](uint32_t r) {
    memset(&status[r], 0, sizeof status[r]);
    status[r].phase = DSA_LLR_REJECTED;
    const double t = now();
    if (admit_s) admit_s[r] = t;
    done_s[r] = t;
    ++finished;
  };
  while (finished < nreq) {
    // admit, in arrival order, everything that has arrived and fits
    while (next < nreq && arrival_s[order[next]] <= now()) {
      const uint32_t r = order[next];
      if (!groups[r] || (L->budget && footprint(req[r]) > L->budget)) {
        reject(r);
        ++next;
        continue;
      }
      uint32_t s;
      int rc = dsa_llr_add(L, groups[r], &req[r], nullptr, &s);
      if (rc == DSA_E_RANGE && active) break;  // slots or budget full: wait for a retirement
      if (rc == DSA_E_ARG || rc == DSA_E_RANGE || rc == DSA_E_STATE) {
        reject(r);
        ++next;
        continue;
      }
      if (rc) return rc;  // device failure
      slot_of[r] = (int)s;
      req_of[s] = (int)r;
      if (admit_s) admit_s[r] = now();
      ++active;
      ++next;
    }
    if (!active) {
      if (next < nreq) {
        const double wait = arrival_s[order[next]] - now();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      }
      continue;
    }
    if (int rc = step(L, chunk)) return rc;
    if (int rc = pull_state(L)) return rc;
    for (uint32_t s = 0; s < L->slots.size(); ++s) {
      if (req_of[s] < 0) continue;
      const dsa::LlrState &st = L->host_state[s];
      if (st.phase != dsa::kLlrDone) continue;
      const uint32_t r = (uint32_t)req_of[s];
      dsa_llr_status_get(L, s, &status[r]);
      done_s[r] = now();
      dsa_llr_remove(L, s);
      req_of[s] = -1;
      --active;
      ++finished;
    }
  }
  return 0;
}

}  // extern "C"
