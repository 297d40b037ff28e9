// kernels.hip -- gfx950 kernels for the discrete-group Symanzik Metropolis
// sweep and its observables.
//
// The sweep is integer gather work: every staple product and every trial
// action is a lookup in the P x P multiplication table (uint16, 2.3 MB for
// P = 1080, resident in each XCD's 4 MB L2), with ReTr, the inverse table and
// the proposal pool in LDS.  No MFMA: there is no dense contraction here.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <type_traits>

#include "device_common.hpp"
#include "dsa_math.hpp"
#include "kernels.hpp"

namespace dsa {

namespace {

// One pass: every listed link of direction mu gets NHIT Metropolis hits.
// Links of one pass share no plaquette and no rectangle, so their updates are
// independent.
template <int D>
__global__ __launch_bounds__(kBlock) void k_sweep(SweepArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t P = A.T.P;
  double *s_retr = reinterpret_cast<double *>(smem);
  uint16_t *s_inv = reinterpret_cast<uint16_t *>(s_retr + P);
  uint16_t *s_pool = s_inv + P;
  load_tables_lds(A.T, s_retr, s_inv, s_pool);
  __syncthreads();

  const TabView T{A.T.mult, s_inv, P};
  const LinkCtx C{A.links, A.mstate, s_retr, s_pool, (double)A.T.npool, A.T.npool, A.beta0, A.beta1, A.beta2,
                  A.seed, A.sweep_lo, A.sweep_hi, A.rng};
  unsigned long long acc = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = xcd_block() * blockDim.x + threadIdx.x; k < A.count; k += stride)
    acc += update_link<D>(C, A.G, T, A.mu, A.sites[k]);
  acc = wave_sum_u64(acc);
  if ((threadIdx.x & 63u) == 0 && acc) atomicAdd(A.acc, acc);
}

// The same pass with the class-product update.  Narrow: the class-index table
// (npool x P bytes, 78 KB for S(1080)) in LDS, two 512-lane blocks per CU.
// Wide: the class values themselves (npool x P int16, 155 KB for S(1080)),
// one lookup per term, one 1024-lane block per CU.  The staples and accepted
// moves still gather from the global table.
constexpr int kClsBlockNarrow = 512, kClsBlockWide = 1024;
// The field form: 512 lanes per block (one block per CU by its LDS), and
// where the pass fills every CU, two links per lane one after the other.
// A wave's second link then overlaps other waves' first ones; with one link
// per lane all waves of a CU start in the same phase and the last ones to
// finish run alone (1024-lane blocks, one link per lane: 81.6 against 77.2
// us per pass at 32^3 x 64).
constexpr int kFieldBlock = 512;

// F (narrow only): the staple-times-link products from the group's GF(4)
// form instead of table gathers; its tables (16.6 KB for S(1080)) push the
// block's LDS past half a CU's, so one block per CU (kFieldBlock lanes).
template <int D, bool W, bool B, bool F = false>
__global__ __launch_bounds__(F ? kFieldBlock : W ? kClsBlockWide : kClsBlockNarrow)
__attribute__((amdgpu_waves_per_eu(4, 8)))
void k_sweep_cls(SweepArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  // class values at a fixed LDS address: the narrow table's bytes index them
  __shared__ int32_t s_rcq[64];
  const uint32_t P = A.T.P, np = A.T.npool;
  uint16_t *s_inv = reinterpret_cast<uint16_t *>(smem);
  uint16_t *s_pool = reinterpret_cast<uint16_t *>(reinterpret_cast<unsigned char *>(s_inv) + pad16(2 * (size_t)P));
  unsigned char *s_tab0 = reinterpret_cast<unsigned char *>(s_pool) + pad16(2 * (size_t)np);
  unsigned char *s_tab = s_tab0 + pad16((W ? 2 : 1) * (size_t)P);
  // narrow form: the per-class hit bounds after the class table (null if none)
  int32_t *s_md = !W && A.T.mdq ? reinterpret_cast<int32_t *>(s_tab + pad16((size_t)np * P)) : nullptr;
  // field form: the packed matrices, their inverses' and the key table next
  unsigned char *s_f = s_tab + pad16((size_t)np * P) + (A.T.mdq ? pad16(4 * (size_t)A.T.ncls) : 0);
  uint32_t *s_fm = reinterpret_cast<uint32_t *>(s_f);
  uint32_t *s_fmi = reinterpret_cast<uint32_t *>(s_f + pad16(4 * (size_t)P));
  uint16_t *s_fkey = reinterpret_cast<uint16_t *>(s_f + 2 * pad16(4 * (size_t)P));
  if constexpr (W) {
    // 1024 lanes: every table in at most 10 rounds
    Staged<2, uint16_t> g_inv;
    Staged<1, uint16_t> g_pool;
    Staged<1, u32x4> g_tab0;
    Staged<10, u32x4> g_tab;
    const uint32_t n0 = (uint32_t)(pad16(2 * (size_t)P) / 16), n1 = (uint32_t)(pad16(2 * (size_t)np * P) / 16);
    const u32x4 *c0 = reinterpret_cast<const u32x4 *>(A.T.cval0), *c1 = reinterpret_cast<const u32x4 *>(A.T.cval);
    g_inv.load(A.T.inv, P);
    g_pool.load(A.T.pool, np);
    g_tab0.load(c0, n0);
    g_tab.load(c1, n1);
    g_inv.store(s_inv, A.T.inv, P);
    g_pool.store(s_pool, A.T.pool, np);
    g_tab0.store(reinterpret_cast<u32x4 *>(s_tab0), c0, n0);
    g_tab.store(reinterpret_cast<u32x4 *>(s_tab), c1, n1);
  } else if constexpr (F) {
    constexpr uint32_t nk = kGf4Keys * sizeof(uint16_t) / 16;
    const u32x4 *ks = reinterpret_cast<const u32x4 *>(A.T.fkey);
    Staged<2, uint32_t> fm, fmi;
    Staged<1, u32x4> fk;
    fm.load(A.T.fmat, P);
    fmi.load(A.T.finv, P);
    fk.load(ks, nk);
    ClsNarrowStage g;
    g.run(A.T, s_inv, s_pool, s_tab0, s_rcq, s_tab, s_md);
    fm.store(s_fm, A.T.fmat, P);
    fmi.store(s_fmi, A.T.finv, P);
    fk.store(reinterpret_cast<u32x4 *>(s_fkey), ks, nk);
  } else {
    ClsNarrowStage g;
    g.run(A.T, s_inv, s_pool, s_tab0, s_rcq, s_tab, s_md);
  }
  __syncthreads();

  const TabView T{A.T.mult, s_inv, P};
  const LinkCtx C{A.links, A.mstate, nullptr, s_pool, (double)np, np, A.beta0, A.beta1, A.beta2,
                  A.seed, A.sweep_lo, A.sweep_hi, A.rng};
  const ClsCtx K{A.T.retr, A.cmargin};
  using L = typename std::conditional<W, ClsWide, ClsNarrow>::type;
  L Lk;
  if constexpr (W) {
    Lk = ClsWide{reinterpret_cast<const int16_t *>(s_tab), reinterpret_cast<const int16_t *>(s_tab0)};
  } else {
    Lk = ClsNarrow{s_tab, s_tab0, s_rcq, s_md};
  }
  unsigned long long acc = 0;
  uint32_t nfall = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  if constexpr (F) {
    const FieldView FV{s_fm, s_fmi, s_fkey};
    for (uint32_t k = xcd_block() * blockDim.x + threadIdx.x; k < A.count; k += stride) {
      ClsLink<D> S;
      cls_prepare_field<D, B>(S, C, Lk, A.G, FV, A.mu, A.sites[k]);
      acc += cls_hits<D, B, L, true>(S, C, Lk, K, T, A.mu, nfall, nullptr);
    }
  } else {
    for (uint32_t k = xcd_block() * blockDim.x + threadIdx.x; k < A.count; k += stride)
      acc += update_link_cls<D, B>(C, Lk, K, A.G, T, A.mu, A.sites[k], nfall);
  }
  acc = wave_sum_u64(acc);
  const unsigned long long nf = wave_sum_u64(nfall);
  if ((threadIdx.x & 63u) == 0) {
    if (acc) atomicAdd(A.acc, acc);
    if (nf) atomicAdd(A.fall, nf);
  }
}

// The same pass with the matrix update: the group's SU(3) matrices (80 B
// per element, 86 KB for S(1080)) in LDS, one 1024-lane block per CU; only
// the links (and accepted hits' new links) come from global memory.
constexpr int kMatBlock = 1024;

template <int D>
__global__ __launch_bounds__(kMatBlock) void k_sweep_mat(SweepArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t P = A.T.P, np = A.T.npool;
  float *s_mat = reinterpret_cast<float *>(smem);
  uint16_t *s_inv = reinterpret_cast<uint16_t *>(smem + 80 * (size_t)P);
  uint16_t *s_pool = reinterpret_cast<uint16_t *>(reinterpret_cast<unsigned char *>(s_inv) + pad16(2 * (size_t)P));
  {
    Staged<6, u32x4> g_mat;
    Staged<2, uint16_t> g_inv;
    Staged<1, uint16_t> g_pool;
    const uint32_t n16 = 5 * P;
    const u32x4 *src = reinterpret_cast<const u32x4 *>(A.T.mat);
    g_mat.load(src, n16);
    g_inv.load(A.T.inv, P);
    g_pool.load(A.T.pool, np);
    g_mat.store(reinterpret_cast<u32x4 *>(s_mat), src, n16);
    g_inv.store(s_inv, A.T.inv, P);
    g_pool.store(s_pool, A.T.pool, np);
  }
  __syncthreads();

  const TabView T{A.T.mult, s_inv, P};
  const LinkCtx C{A.links, A.mstate, nullptr, s_pool, (double)np, np, A.beta0, A.beta1, A.beta2,
                  A.seed, A.sweep_lo, A.sweep_hi, A.rng};
  const MatCtx K{s_mat, A.T.retr, A.cmargin, A.cmargin_acc};
  unsigned long long acc = 0;
  uint32_t nfall = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = xcd_block() * blockDim.x + threadIdx.x; k < A.count; k += stride)
    acc += update_link_mat<D>(C, K, A.G, T, A.mu, A.sites[k], nfall);
  acc = wave_sum_u64(acc);
  const unsigned long long nf = wave_sum_u64(nfall);
  if ((threadIdx.x & 63u) == 0) {
    if (acc) atomicAdd(A.acc, acc);
    if (nf) atomicAdd(A.fall, nf);
  }
}

// The narrow form for passes too small to fill the device: a 512-lane block
// takes NL links; all its lanes first draw the links' Philox uniforms
// (512 / NL lanes per link), so a link's hits wait only on its lookups.  NL
// is 128 below kClsSmallPass links a pass and 256 below kClsMidPass
// (measured on 8^4 ... 32^3x64 lattices); larger passes fill the device
// with k_sweep_cls, whose 16 waves a CU hide the draws.
constexpr uint32_t kClsSmallPass = 32768, kClsMidPass = 65536;
// static LDS of k_sweep_cls_small<D, NL>: class values, then the draws
constexpr size_t cls_small_static_lds(int nl) {
  return 64 * sizeof(int32_t) + (size_t)kNHit * nl * (sizeof(double) + sizeof(uint16_t));
}

// F: the staple-times-link products from the group's GF(4) form, its
// tables after the narrow ones (as in k_sweep_cls).
template <int D, int NL, bool B, bool F = false>
__global__ __launch_bounds__(kClsBlockNarrow) void k_sweep_cls_small(SweepArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int32_t s_rcq[64];
  __shared__ double s_u2[kNHit * NL];
  __shared__ uint16_t s_jp[kNHit * NL];
  const uint32_t P = A.T.P, np = A.T.npool;
  uint16_t *s_inv = reinterpret_cast<uint16_t *>(smem);
  uint16_t *s_pool = reinterpret_cast<uint16_t *>(smem + pad16(2 * (size_t)P));
  unsigned char *s_tab0 = reinterpret_cast<unsigned char *>(s_pool) + pad16(2 * (size_t)np);
  unsigned char *s_tab = s_tab0 + pad16(P);
  int32_t *s_md = A.T.mdq ? reinterpret_cast<int32_t *>(s_tab + pad16((size_t)np * P)) : nullptr;
  unsigned char *s_f = s_tab + pad16((size_t)np * P) + (A.T.mdq ? pad16(4 * (size_t)A.T.ncls) : 0);
  uint32_t *s_fm = reinterpret_cast<uint32_t *>(s_f);
  uint32_t *s_fmi = reinterpret_cast<uint32_t *>(s_f + pad16(4 * (size_t)P));
  uint16_t *s_fkey = reinterpret_cast<uint16_t *>(s_f + 2 * pad16(4 * (size_t)P));
  if constexpr (F) {
    constexpr uint32_t nk = kGf4Keys * sizeof(uint16_t) / 16;
    const u32x4 *ks = reinterpret_cast<const u32x4 *>(A.T.fkey);
    Staged<3, uint32_t> fm, fmi;
    Staged<1, u32x4> fk;
    fm.load(A.T.fmat, P);
    fmi.load(A.T.finv, P);
    fk.load(ks, nk);
    ClsNarrowStage g;
    g.run(A.T, s_inv, s_pool, s_tab0, s_rcq, s_tab, s_md);
    fm.store(s_fm, A.T.fmat, P);
    fmi.store(s_fmi, A.T.finv, P);
    fk.store(reinterpret_cast<u32x4 *>(s_fkey), ks, nk);
  } else {
    ClsNarrowStage g;
    g.run(A.T, s_inv, s_pool, s_tab0, s_rcq, s_tab, s_md);
  }
  const uint32_t first = blockIdx.x * NL;
  const uint32_t n = A.count - first < (uint32_t)NL ? A.count - first : (uint32_t)NL;
  const double npd = (double)np;
  for (uint32_t j = threadIdx.x; j < n * (uint32_t)(kNHit / 2); j += blockDim.x) {
    const uint32_t l = j % n, q = j / n;  // hits 2q and 2q + 1 of link l
    const uint32_t lk = A.sites[first + l] * (uint32_t)D + (uint32_t)A.mu;
    const U4 w = philox4x32_10(U4{lk, q, A.sweep_lo, A.sweep_hi}, A.seed, kKeySweep);
    s_jp[2 * q * NL + l] = (uint16_t)hit_pool_index(w.x, np);
    s_u2[2 * q * NL + l] = hit_uniform(w.y);
    s_jp[(2 * q + 1) * NL + l] = (uint16_t)hit_pool_index(w.z, np);
    s_u2[(2 * q + 1) * NL + l] = hit_uniform(w.w);
  }
  __syncthreads();

  const TabView T{A.T.mult, s_inv, P};
  const LinkCtx C{A.links, A.mstate, nullptr, s_pool, npd, np, A.beta0, A.beta1, A.beta2,
                  A.seed, A.sweep_lo, A.sweep_hi, A.rng};
  const ClsCtx K{A.T.retr, A.cmargin};
  const ClsNarrow Lk{s_tab, s_tab0, s_rcq, s_md};
  unsigned long long acc = 0;
  uint32_t nfall = 0;
  if (threadIdx.x < n) {
    const PreDraws pre{s_jp, s_u2, (uint32_t)NL, threadIdx.x};
    if constexpr (F) {
      const FieldView FV{s_fm, s_fmi, s_fkey};
      ClsLink<D> S;
      cls_prepare_field<D, B>(S, C, Lk, A.G, FV, A.mu, A.sites[first + threadIdx.x]);
      acc = cls_hits<D, B>(S, C, Lk, K, T, A.mu, nfall, &pre);
    } else {
      acc = update_link_cls<D, B>(C, Lk, K, A.G, T, A.mu, A.sites[first + threadIdx.x], nfall, &pre);
    }
  }
  acc = wave_sum_u64(acc);
  const unsigned long long nf = wave_sum_u64(nfall);
  if ((threadIdx.x & 63u) == 0) {
    if (acc) atomicAdd(A.acc, acc);
    if (nf) atomicAdd(A.fall, nf);
  }
}

// Many chains, one pass index: block b serves task[b] = up to blockDim.x
// links of one chain's pass (one per lane), with that chain's tables in LDS.  The sweep number of
// chain c is its base plus the batch counter kept in device memory (so a
// captured graph can be replayed unchanged).
template <int D>
__global__ __launch_bounds__(kBatchBlockMax) void k_sweep_batch(const BatchChain *chains, const BatchTask *tasks,
                                                        const unsigned long long *sweep_ctr,
                                                        unsigned long long *acc_out) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int32_t s_rcq[64];
  const BatchTask tk = tasks[blockIdx.x];
  const BatchChain &BC = chains[tk.chain];
  const uint32_t P = BC.T.P, np = BC.T.npool;
  const uint64_t sweep = BC.sweep_base + *sweep_ctr;
  unsigned long long acc = 0;
  if (BC.cls) {  // the layout of k_sweep_cls's narrow form (uniform over the block)
    uint16_t *s_inv = reinterpret_cast<uint16_t *>(smem);
    uint16_t *s_pool = reinterpret_cast<uint16_t *>(smem + pad16(2 * (size_t)P));
    unsigned char *s_tab0 = reinterpret_cast<unsigned char *>(s_pool) + pad16(2 * (size_t)np);
    unsigned char *s_tab = s_tab0 + pad16(P);
    int32_t *s_md = BC.T.mdq ? reinterpret_cast<int32_t *>(s_tab + pad16((size_t)np * P)) : nullptr;
    unsigned char *s_f = s_tab + pad16((size_t)np * P) + (BC.T.mdq ? pad16(4 * (size_t)BC.T.ncls) : 0);
    uint32_t *s_fm = reinterpret_cast<uint32_t *>(s_f);
    uint32_t *s_fmi = reinterpret_cast<uint32_t *>(s_f + pad16(4 * (size_t)P));
    uint16_t *s_fkey = reinterpret_cast<uint16_t *>(s_f + 2 * pad16(4 * (size_t)P));
    // a frozen chain (plaquette-first test on) keeps the table gathers, which
    // hit the caches for links near one another
    const bool fld = BC.field && !(BC.bound && s_md);
    if (fld) {  // uniform over the block
      constexpr uint32_t nk = kGf4Keys * sizeof(uint16_t) / 16;
      const u32x4 *ks = reinterpret_cast<const u32x4 *>(BC.T.fkey);
      Staged<5, uint32_t> fm, fmi;
      Staged<2, u32x4> fk;
      fm.load(BC.T.fmat, P);
      fmi.load(BC.T.finv, P);
      fk.load(ks, nk);
      ClsNarrowStage g;
      g.run(BC.T, s_inv, s_pool, s_tab0, s_rcq, s_tab, s_md);
      fm.store(s_fm, BC.T.fmat, P);
      fmi.store(s_fmi, BC.T.finv, P);
      fk.store(reinterpret_cast<u32x4 *>(s_fkey), ks, nk);
    } else {
      ClsNarrowStage g;
      g.run(BC.T, s_inv, s_pool, s_tab0, s_rcq, s_tab, s_md);
    }
    __syncthreads();
    const TabView T{BC.T.mult, s_inv, P};
    const LinkCtx C{BC.links, BC.mstate, nullptr, s_pool, (double)np, np, BC.beta0, BC.beta1, BC.beta2,
                    BC.seed, (uint32_t)sweep, (uint32_t)(sweep >> 32), BC.rng};
    const ClsCtx K{BC.T.retr, BC.cmargin};
    const ClsNarrow Lk{s_tab, s_tab0, s_rcq, s_md};
    uint32_t nfall = 0;
    if (fld) {
      const FieldView FV{s_fm, s_fmi, s_fkey};
      for (uint32_t t = threadIdx.x; t < tk.count; t += blockDim.x) {
        ClsLink<D> S;
        cls_prepare_field<D, false>(S, C, Lk, BC.G, FV, tk.mu, gmem(BC.sites)[tk.offset + t]);
        acc += cls_hits<D, false>(S, C, Lk, K, T, tk.mu, nfall, nullptr);
      }
    } else if (BC.bound && s_md) {  // uniform over the block
      for (uint32_t t = threadIdx.x; t < tk.count; t += blockDim.x)
        acc += update_link_cls<D, true>(C, Lk, K, BC.G, T, tk.mu, gmem(BC.sites)[tk.offset + t], nfall);
    } else {
      for (uint32_t t = threadIdx.x; t < tk.count; t += blockDim.x)
        acc += update_link_cls<D, false>(C, Lk, K, BC.G, T, tk.mu, gmem(BC.sites)[tk.offset + t], nfall);
    }
    const unsigned long long nf = wave_sum_u64(nfall);
    if ((threadIdx.x & 63u) == 0 && nf) atomicAdd(BC.fall, nf);
  } else {
    double *s_retr = reinterpret_cast<double *>(smem);
    uint16_t *s_inv = reinterpret_cast<uint16_t *>(s_retr + P);
    uint16_t *s_pool = s_inv + P;
    load_tables_lds(BC.T, s_retr, s_inv, s_pool);
    __syncthreads();
    const TabView T{BC.T.mult, s_inv, P};
    const LinkCtx C{BC.links, BC.mstate, s_retr, s_pool, (double)np, np, BC.beta0, BC.beta1, BC.beta2,
                    BC.seed, (uint32_t)sweep, (uint32_t)(sweep >> 32), BC.rng};
    for (uint32_t t = threadIdx.x; t < tk.count; t += blockDim.x)
      acc += update_link<D>(C, BC.G, T, tk.mu, gmem(BC.sites)[tk.offset + t]);
  }
  acc = wave_sum_u64(acc);
  if ((threadIdx.x & 63u) == 0 && acc) atomicAdd(&acc_out[tk.chain], acc);
}

__global__ void k_count_sweep(unsigned long long *sweep_ctr) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *sweep_ctr += 1;
}

__global__ void k_init_philox(uint16_t *links, Geom G, uint32_t P, uint32_t seed) {
  const uint32_t nl = G.V * (uint32_t)G.D;
  for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < nl; l += gridDim.x * blockDim.x) {
    const U4 w = philox4x32_10(U4{l, 0u, 0u, 0u}, seed, kKeyInit);
    const uint32_t e = (uint32_t)((double)P * u53(w.x, w.y));
    const uint32_t i = l / (uint32_t)G.D, mu = l - i * (uint32_t)G.D;
    uint32_t par = 0, r = i;
    for (int d = 0; d < G.D; ++d) {
      const uint32_t q = r / G.ext[d];
      par += r - q * G.ext[d];
      r = q;
    }
    links[(size_t)mu * G.V + lpos(G, i, par & 1u)] = (uint16_t)e;
  }
}

__global__ void k_fill(uint16_t *links, uint32_t n, uint16_t v) {
  for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < n; l += gridDim.x * blockDim.x) links[l] = v;
}


__global__ void k_pack(uint16_t *links, const int32_t *aos, Geom G) {
  const uint32_t nl = G.V * (uint32_t)G.D;
  for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < nl; l += gridDim.x * blockDim.x) {
    const uint32_t i = l / (uint32_t)G.D, mu = l - i * (uint32_t)G.D;
    links[(size_t)mu * G.V + lpos(G, i, site_parity(G, i))] = (uint16_t)aos[l];
  }
}

__global__ void k_unpack(int32_t *aos, const uint16_t *links, Geom G) {
  const uint32_t nl = G.V * (uint32_t)G.D;
  for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < nl; l += gridDim.x * blockDim.x) {
    const uint32_t i = l / (uint32_t)G.D, mu = l - i * (uint32_t)G.D;
    aos[l] = links[(size_t)mu * G.V + lpos(G, i, site_parity(G, i))];
  }
}

__global__ void k_kentucky(unsigned long long *out, const uint16_t *links, const double *mats, Geom G) {
  const uint32_t Nt = G.ext[0], Nx = G.ext[1];
  for (uint32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < G.V; o += gridDim.x * blockDim.x) {
    // o = x1 + Nx*(x2 + Nx*(x3 + Nx*x0)); site i = x0 + Nt*(x1 + Nx*(x2 + Nx*x3))
    const uint32_t x1 = o % Nx, x2 = (o / Nx) % Nx, x3 = (o / (Nx * Nx)) % Nx, x0 = o / (Nx * Nx * Nx);
    const uint32_t i = x0 + Nt * (x1 + Nx * (x2 + Nx * x3));
    const uint32_t par = (x0 + x1 + x2 + x3) & 1u;
    for (int d = 0; d < 4; ++d) {
      const uint32_t e = links[(size_t)d * G.V + lpos(G, i, par)];
      const uint32_t dd = (uint32_t)((d + 3) % 4);
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
          for (int part = 0; part < 2; ++part) {
            const double v = mats[18u * e + part + 2 * (c + 3 * r)];
            const size_t k = o + (size_t)G.V * (part + 2 * (r + 3 * (c + 3 * dd)));
            out[k] = __builtin_bswap64(__double_as_longlong(v));
          }
    }
  }
}

template <int D>
__global__ void k_staples(int32_t *st_out, int32_t *re_out, const uint16_t *links, Tables Tb, Geom G) {
  constexpr int NS = 2 * (D - 1), NR = 6 * (D - 1);
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  const uint32_t nl = G.V * (uint32_t)D;
  for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < nl; l += gridDim.x * blockDim.x) {
    const uint32_t i = l / D;
    const int mu = (int)(l - i * D);
    const Site<D> s = decode<D>(G, i);
    uint32_t st[NS], re[NR];
    build_staples<D>(links, G, T, s, mu, st, re);
    for (int j = 0; j < NS; ++j) st_out[(size_t)l * NS + j] = (int32_t)st[j];
    for (int q = 0; q < NR; ++q) re_out[(size_t)l * NR + q] = (int32_t)re[q];
  }
}

__global__ void k_lines(uint16_t *lines, const uint16_t *links, Tables Tb, Geom G, uint32_t nmax_t,
                        uint32_t nmax_x) {
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < G.V; i += gridDim.x * blockDim.x) {
    for (int d = 0; d < G.D; ++d) {
      const uint32_t nmax = d == 0 ? nmax_t : nmax_x;
      if (!nmax) continue;
      uint32_t g = link_at(links, G, i, d);
      uint32_t j = i;
      lines[line_row(d, 1, nmax_t, nmax_x) * G.V + i] = (uint16_t)g;
      for (uint32_t n = 2; n <= nmax; ++n) {
        j = step_site(G, j, d, 1);
        g = T.M(g, link_at(links, G, j, d));
        lines[line_row(d, n, nmax_t, nmax_x) * G.V + i] = (uint16_t)g;
      }
    }
  }
}

__global__ void k_loops(int32_t *out, int kind, int k, int l, const uint16_t *links, const uint16_t *lines,
                        Tables Tb, Geom G, uint32_t nmax_t, uint32_t nmax_x) {
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  const int D = G.D;
  const uint32_t nsite = kind == kLoopPoly ? G.V / G.ext[0] : G.V;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nsite; q += gridDim.x * blockDim.x) {
    if (kind == kLoopPlaq) {
      int p = 0;
      const int np = D * (D - 1) / 2;
      for (int d1 = 0; d1 < D; ++d1)
        for (int d2 = d1 + 1; d2 < D; ++d2) out[(size_t)q * np + p++] = (int32_t)plaq_elem(links, G, T, q, d1, d2);
    } else if (kind == kLoopRect) {
      int p = 0;
      const int np = D * (D - 1);
      for (int d1 = 0; d1 < D; ++d1)
        for (int d2 = 0; d2 < D; ++d2)
          if (d1 != d2) out[(size_t)q * np + p++] = (int32_t)rect_elem(links, G, T, q, d1, d2);
    } else if (kind == kLoopPoly) {
      out[q] = (int32_t)poly_elem(links, G, T, q * G.ext[0]);
    } else {
      for (int d1 = 1; d1 < D; ++d1)
        out[(size_t)q * (D - 1) + (d1 - 1)] =
            (int32_t)wilson_elem(lines, G, T, q, d1, (uint32_t)k + 1, (uint32_t)l + 1, nmax_t, nmax_x);
    }
  }
}

constexpr uint32_t kObsMaxBlocks = 1024;

__global__ __launch_bounds__(kBlock) void k_obs_local(double *partials, const uint16_t *links, Tables Tb, Geom G) {
  __shared__ double s_part[kBlock / 64];
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  const int D = G.D;
  double pre = 0, pim = 0, pl = 0, rc = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < G.V; i += gridDim.x * blockDim.x) {
    if (i % G.ext[0] == 0) {
      const uint32_t p = poly_elem(links, G, T, i);
      pre += Tb.retr[p];
      pim += Tb.imtr[p];
    }
    for (int d1 = 0; d1 < D; ++d1)
      for (int d2 = d1 + 1; d2 < D; ++d2) pl += Tb.retr[plaq_elem(links, G, T, i, d1, d2)];
    for (int d1 = 0; d1 < D; ++d1)
      for (int d2 = 0; d2 < D; ++d2)
        if (d1 != d2) rc += Tb.retr[rect_elem(links, G, T, i, d1, d2)];
  }
  const double a = block_sum(pre, s_part);
  const double b = block_sum(pim, s_part);
  const double c = block_sum(pl, s_part);
  const double d = block_sum(rc, s_part);
  if (threadIdx.x == 0) {
    partials[(size_t)0 * gridDim.x + blockIdx.x] = a;
    partials[(size_t)1 * gridDim.x + blockIdx.x] = b;
    partials[(size_t)2 * gridDim.x + blockIdx.x] = c;
    partials[(size_t)3 * gridDim.x + blockIdx.x] = d;
  }
}

__global__ __launch_bounds__(kBlock) void k_obs_wilson(double *partials, const uint16_t *lines, Tables Tb, Geom G,
                                                       uint32_t wl, uint32_t nmax_t, uint32_t nmax_x) {
  __shared__ double s_part[kBlock / 64];
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  const uint32_t kl = blockIdx.y;
  const uint32_t K = kl / wl + 1, Tt = kl % wl + 1;
  double w = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < G.V; i += gridDim.x * blockDim.x)
    for (int d1 = 1; d1 < G.D; ++d1) w += Tb.retr[wilson_elem(lines, G, T, i, d1, K, Tt, nmax_t, nmax_x)];
  const double r = block_sum(w, s_part);
  if (threadIdx.x == 0) partials[(size_t)(4 + kl) * gridDim.x + blockIdx.x] = r;
}

__global__ void k_obs_final(double *out, const double *partials, Geom G, uint32_t nobs, uint32_t nb) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nobs) return;
  double s = 0;
  for (uint32_t b = 0; b < nb; ++b) s += partials[(size_t)q * nb + b];
  out[q] = s / obs_norm(G, q);
}

// One lane, the reference's summation order at one thread.
__global__ void k_obs_ordered(double *out, const uint16_t *links, const uint16_t *lines, Tables Tb, Geom G,
                              uint32_t wk, uint32_t wl, uint32_t nmax_t, uint32_t nmax_x) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const TabView T{Tb.mult, Tb.inv, Tb.P};
  const int D = G.D;
  double lre = 0, lim = 0;
  for (uint32_t i = 0; i < G.V; i += G.ext[0]) {
    const uint32_t p = poly_elem(links, G, T, i);
    lre += Tb.retr[p];
    lim += Tb.imtr[p];
  }
  out[0] = lre / obs_norm(G, 0);
  out[1] = lim / obs_norm(G, 1);
  double sp = 0;
  for (uint32_t i = 0; i < G.V; ++i)
    for (int d1 = 0; d1 < D; ++d1)
      for (int d2 = d1 + 1; d2 < D; ++d2) sp += Tb.retr[plaq_elem(links, G, T, i, d1, d2)];
  out[2] = sp / obs_norm(G, 2);
  double sr = 0;
  for (uint32_t i = 0; i < G.V; ++i)
    for (int d1 = 0; d1 < D; ++d1)
      for (int d2 = 0; d2 < D; ++d2)
        if (d1 != d2) sr += Tb.retr[rect_elem(links, G, T, i, d1, d2)];
  out[3] = sr / obs_norm(G, 3);
  for (uint32_t k = 0; k < wk; ++k)
    for (uint32_t l = 0; l < wl; ++l) {
      double w = 0;
      for (uint32_t i = 0; i < G.V; ++i)
        for (int d1 = 1; d1 < D; ++d1) w += Tb.retr[wilson_elem(lines, G, T, i, d1, k + 1, l + 1, nmax_t, nmax_x)];
      out[4 + k * wl + l] = w / obs_norm(G, 4);
    }
}

// Batched measurement: block b sums the Polyakov (t = 0 sites), plaquette and
// rectangle loops of its task's sites, in a fixed tree.
__global__ __launch_bounds__(kBlock) void k_obs_batch(double *partials, const BatchChain *chains,
                                                      const ObsTask *tasks, uint32_t ntasks) {
  __shared__ double s_part[kBlock / 64];
  const ObsTask tk = tasks[blockIdx.x];
  const BatchChain &BC = chains[tk.chain];
  const Geom G = BC.G;
  const TabView T{BC.T.mult, BC.T.inv, BC.T.P};
  const uint16_t *links = BC.links;
  const int D = G.D;
  double pre = 0, pim = 0, pl = 0, rc = 0;
  for (uint32_t t = threadIdx.x; t < tk.nsites; t += blockDim.x) {
    const uint32_t i = tk.site0 + t;
    if (i % G.ext[0] == 0) {
      const uint32_t p = poly_elem(links, G, T, i);
      pre += BC.T.retr[p];
      pim += BC.T.imtr[p];
    }
    for (int d1 = 0; d1 < D; ++d1)
      for (int d2 = d1 + 1; d2 < D; ++d2) pl += BC.T.retr[plaq_elem(links, G, T, i, d1, d2)];
    for (int d1 = 0; d1 < D; ++d1)
      for (int d2 = 0; d2 < D; ++d2)
        if (d1 != d2) rc += BC.T.retr[rect_elem(links, G, T, i, d1, d2)];
  }
  const double a = block_sum(pre, s_part);
  const double b = block_sum(pim, s_part);
  const double c = block_sum(pl, s_part);
  const double d = block_sum(rc, s_part);
  if (threadIdx.x == 0) {
    partials[(size_t)0 * ntasks + blockIdx.x] = a;
    partials[(size_t)1 * ntasks + blockIdx.x] = b;
    partials[(size_t)2 * ntasks + blockIdx.x] = c;
    partials[(size_t)3 * ntasks + blockIdx.x] = d;
  }
}

__global__ void k_obs_batch_final(double *out, const double *partials, const BatchChain *chains,
                                  const uint32_t *first, const uint32_t *n, uint32_t nchains, uint32_t ntasks) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nchains * 4) return;
  const uint32_t c = t / 4, q = t % 4;
  double s = 0;
  for (uint32_t k = 0; k < n[c]; ++k) s += partials[(size_t)q * ntasks + first[c] + k];
  out[t] = s / obs_norm(chains[c].G, q);
}

inline uint32_t grid_for(uint64_t n, uint32_t cap = 8192) {
  uint64_t b = (n + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (uint32_t)b;
}

}  // namespace

size_t sweep_lds_bytes(const Tables &T) {
  return (size_t)T.P * sizeof(double) + (size_t)T.P * sizeof(uint16_t) + (size_t)T.npool * sizeof(uint16_t);
}

size_t sweep_cls_lds_bytes(const Tables &T, int mode) {
  if (mode == kClsField)
    return sweep_cls_lds_bytes(T, kClsNarrow) + 2 * pad16(4 * (size_t)T.P) + kGf4Keys * sizeof(uint16_t);
  if (mode == kClsMatrix) return 80 * (size_t)T.P + pad16(2 * (size_t)T.P) + pad16(2 * (size_t)T.npool);
  if (mode == kClsWide)
    return pad16(2 * (size_t)T.P) + pad16(2 * (size_t)T.npool) + pad16(2 * (size_t)T.P) +
           pad16(2 * (size_t)T.npool * T.P);
  return pad16(2 * (size_t)T.P) + pad16(2 * (size_t)T.npool) + pad16(T.P) +
         pad16((size_t)T.npool * T.P) + (T.mdq ? pad16(4 * (size_t)T.ncls) : 0);
}

hipError_t launch_sweep(const SweepArgs &A, int max_blocks, hipStream_t s) {
  if (A.count == 0) return hipSuccess;
  // small passes: the narrow small-pass kernel, with the GF(4) products
  // where the chain runs the field form and its tables fit beside the draws
  const bool fld = A.use_cls == kClsField && A.T.fmat;
  const size_t lds_n = sweep_cls_lds_bytes(A.T, kClsNarrow), lds_f = sweep_cls_lds_bytes(A.T, kClsField);
  const bool fit_f256 = fld && lds_f + cls_small_static_lds(256) <= A.lds_max;
  const bool fit_f128 = fld && lds_f + cls_small_static_lds(128) <= A.lds_max;
  const bool small_f = fit_f128 && A.count < kClsSmallPass ? true : fit_f256;  // field form fits this pass size
  const size_t lds_small = small_f ? lds_f : lds_n;
  const bool sm = A.count < kClsSmallPass ||
                  lds_small + cls_small_static_lds(256) > A.lds_max;  // 128 links where 256 do not fit
  if ((A.use_cls == kClsNarrow || (A.use_cls == kClsField && !A.field_all)) && A.rng == kRngPhilox &&
      A.count < kClsMidPass && lds_small + cls_small_static_lds(sm ? 128 : 256) <= A.lds_max) {
    if (!A.fall || !A.T.cprod) return hipErrorInvalidValue;
    const uint32_t nl = sm ? 128 : 256, nb = (A.count + nl - 1) / nl;
    const size_t lds = lds_small;
    const bool bnd = A.bound && A.T.mdq;
#define DSA_SMALL_F(d, FF)                                                                                     \
    if (sm && bnd)                                                                                           \
      hipLaunchKernelGGL((k_sweep_cls_small<d, 128, true, FF>), dim3(nb), dim3(kClsBlockNarrow), lds, s, A);    \
    else if (sm)                                                                                             \
      hipLaunchKernelGGL((k_sweep_cls_small<d, 128, false, FF>), dim3(nb), dim3(kClsBlockNarrow), lds, s, A);   \
    else if (bnd)                                                                                            \
      hipLaunchKernelGGL((k_sweep_cls_small<d, 256, true, FF>), dim3(nb), dim3(kClsBlockNarrow), lds, s, A);    \
    else                                                                                                     \
      hipLaunchKernelGGL((k_sweep_cls_small<d, 256, false, FF>), dim3(nb), dim3(kClsBlockNarrow), lds, s, A);
#define DSA_SMALL(d)                                                                                       \
  case d:                                                                                                \
    if (small_f) { DSA_SMALL_F(d, true) } else { DSA_SMALL_F(d, false) }                                   \
    break;
    switch (A.G.D) {
      DSA_SMALL(2)
      DSA_SMALL(3)
      DSA_SMALL(4)
      DSA_SMALL(5)
      default: return hipErrorInvalidValue;
    }
#undef DSA_SMALL
#undef DSA_SMALL_F
    return hipGetLastError();
  }
  if (A.use_cls == kClsMatrix) {
    if (!A.fall || !A.T.mat) return hipErrorInvalidValue;
    uint32_t nb = (A.count + kMatBlock - 1) / kMatBlock;
    if (max_blocks > 0 && nb > (uint32_t)max_blocks) nb = (uint32_t)max_blocks;
    const size_t lds = sweep_cls_lds_bytes(A.T, kClsMatrix);
    switch (A.G.D) {
      case 2: hipLaunchKernelGGL(k_sweep_mat<2>, dim3(nb), dim3(kMatBlock), lds, s, A); break;
      case 3: hipLaunchKernelGGL(k_sweep_mat<3>, dim3(nb), dim3(kMatBlock), lds, s, A); break;
      case 4: hipLaunchKernelGGL(k_sweep_mat<4>, dim3(nb), dim3(kMatBlock), lds, s, A); break;
      case 5: hipLaunchKernelGGL(k_sweep_mat<5>, dim3(nb), dim3(kMatBlock), lds, s, A); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  if (A.use_cls == kClsField) {
    if (!A.fall || !A.T.cprod || !A.T.fmat) return hipErrorInvalidValue;
    // a link per lane up to one block per CU, then at least two per lane
    const uint32_t ncu = A.ncu ? A.ncu : 256u;
    uint32_t nb = (A.count + kFieldBlock - 1) / kFieldBlock;
    if (nb > ncu) nb = std::max(ncu, (A.count + 2 * kFieldBlock - 1) / (2 * kFieldBlock));
    if (max_blocks > 0 && nb > (uint32_t)max_blocks) nb = (uint32_t)max_blocks;
    const size_t lds = sweep_cls_lds_bytes(A.T, kClsField);
    const bool bnd = A.bound && A.T.mdq;
#define DSA_FIELD_CASE(DD)                                                                                    \
  case DD:                                                                                                    \
    if (bnd) hipLaunchKernelGGL((k_sweep_cls<DD, false, true, true>), dim3(nb), dim3(kFieldBlock), lds, s, A); \
    else hipLaunchKernelGGL((k_sweep_cls<DD, false, false, true>), dim3(nb), dim3(kFieldBlock), lds, s, A);   \
    break;
    switch (A.G.D) {
      DSA_FIELD_CASE(2)
      DSA_FIELD_CASE(3)
      DSA_FIELD_CASE(4)
      DSA_FIELD_CASE(5)
      default: return hipErrorInvalidValue;
    }
#undef DSA_FIELD_CASE
    return hipGetLastError();
  }
  if (A.use_cls == kClsNarrow || A.use_cls == kClsWide) {
    const bool w = A.use_cls == kClsWide;
    if (!A.fall || (w ? !A.T.cval : !A.T.cprod)) return hipErrorInvalidValue;
    const uint32_t bs = w ? kClsBlockWide : kClsBlockNarrow;
    uint32_t nb = (A.count + bs - 1) / bs;
    if (max_blocks > 0 && nb > (uint32_t)max_blocks) nb = (uint32_t)max_blocks;
    const size_t lds = sweep_cls_lds_bytes(A.T, A.use_cls);
    const bool bnd = A.bound && A.T.mdq;
#define DSA_CLS_CASE(DD)                                                                        \
  case DD:                                                                                      \
    if (w) hipLaunchKernelGGL((k_sweep_cls<DD, true, false>), dim3(nb), dim3(bs), lds, s, A);    \
    else if (bnd) hipLaunchKernelGGL((k_sweep_cls<DD, false, true>), dim3(nb), dim3(bs), lds, s, A); \
    else hipLaunchKernelGGL((k_sweep_cls<DD, false, false>), dim3(nb), dim3(bs), lds, s, A);     \
    break;
    switch (A.G.D) {
      DSA_CLS_CASE(2)
      DSA_CLS_CASE(3)
      DSA_CLS_CASE(4)
      DSA_CLS_CASE(5)
      default: return hipErrorInvalidValue;
    }
#undef DSA_CLS_CASE
    return hipGetLastError();
  }
  uint32_t nb = (A.count + kBlock - 1) / kBlock;
  if (max_blocks > 0 && nb > (uint32_t)max_blocks) nb = (uint32_t)max_blocks;
  const size_t lds = sweep_lds_bytes(A.T);
  switch (A.G.D) {
    case 2: hipLaunchKernelGGL(k_sweep<2>, dim3(nb), dim3(kBlock), lds, s, A); break;
    case 3: hipLaunchKernelGGL(k_sweep<3>, dim3(nb), dim3(kBlock), lds, s, A); break;
    case 4: hipLaunchKernelGGL(k_sweep<4>, dim3(nb), dim3(kBlock), lds, s, A); break;
    case 5: hipLaunchKernelGGL(k_sweep<5>, dim3(nb), dim3(kBlock), lds, s, A); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_kentucky(double *out, const uint16_t *links, const double *mats, const Geom &G, hipStream_t s) {
  if (G.D != 4) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_kentucky, dim3(grid_for(G.V)), dim3(kBlock), 0, s, reinterpret_cast<unsigned long long *>(out),
                     links, mats, G);
  return hipGetLastError();
}

hipError_t launch_sweep_batch(int D, const BatchChain *chains, const BatchTask *tasks, uint32_t ntasks,
                              uint32_t block, size_t lds_bytes, const unsigned long long *sweep_ctr,
                              unsigned long long *acc, hipStream_t s) {
  if (!ntasks) return hipSuccess;
  if (block == 0 || block > (uint32_t)kBatchBlockMax || block % 64) return hipErrorInvalidValue;
  const dim3 g(ntasks), b(block);
  switch (D) {
    case 2: hipLaunchKernelGGL(k_sweep_batch<2>, g, b, lds_bytes, s, chains, tasks, sweep_ctr, acc); break;
    case 3: hipLaunchKernelGGL(k_sweep_batch<3>, g, b, lds_bytes, s, chains, tasks, sweep_ctr, acc); break;
    case 4: hipLaunchKernelGGL(k_sweep_batch<4>, g, b, lds_bytes, s, chains, tasks, sweep_ctr, acc); break;
    case 5: hipLaunchKernelGGL(k_sweep_batch<5>, g, b, lds_bytes, s, chains, tasks, sweep_ctr, acc); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_count_sweep(unsigned long long *sweep_ctr, hipStream_t s) {
  hipLaunchKernelGGL(k_count_sweep, dim3(1), dim3(64), 0, s, sweep_ctr);
  return hipGetLastError();
}

hipError_t launch_measure_batch(double *out, double *partials, const BatchChain *chains, const ObsTask *tasks,
                                uint32_t ntasks, const uint32_t *first, const uint32_t *n, uint32_t nchains,
                                hipStream_t s) {
  if (!ntasks) return hipSuccess;
  hipLaunchKernelGGL(k_obs_batch, dim3(ntasks), dim3(kBlock), 0, s, partials, chains, tasks, ntasks);
  hipLaunchKernelGGL(k_obs_batch_final, dim3((nchains * 4 + 63) / 64), dim3(64), 0, s, out, partials, chains, first,
                     n, nchains, ntasks);
  return hipGetLastError();
}

hipError_t launch_init_philox(uint16_t *links, const Geom &G, uint32_t P, uint32_t seed, hipStream_t s) {
  hipLaunchKernelGGL(k_init_philox, dim3(grid_for((uint64_t)G.V * G.D)), dim3(kBlock), 0, s, links, G, P, seed);
  return hipGetLastError();
}

hipError_t launch_fill(uint16_t *links, const Geom &G, uint16_t v, hipStream_t s) {
  const uint32_t n = G.V * (uint32_t)G.D;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, s, links, n, v);
  return hipGetLastError();
}

hipError_t launch_pack(uint16_t *links, const int32_t *aos, const Geom &G, hipStream_t s) {
  hipLaunchKernelGGL(k_pack, dim3(grid_for((uint64_t)G.V * G.D)), dim3(kBlock), 0, s, links, aos, G);
  return hipGetLastError();
}

hipError_t launch_unpack(int32_t *aos, const uint16_t *links, const Geom &G, hipStream_t s) {
  hipLaunchKernelGGL(k_unpack, dim3(grid_for((uint64_t)G.V * G.D)), dim3(kBlock), 0, s, aos, links, G);
  return hipGetLastError();
}

hipError_t launch_staples(int32_t *st, int32_t *re, const uint16_t *links, const Tables &T, const Geom &G,
                          hipStream_t s) {
  const dim3 g(grid_for((uint64_t)G.V * G.D)), b(kBlock);
  switch (G.D) {
    case 2: hipLaunchKernelGGL(k_staples<2>, g, b, 0, s, st, re, links, T, G); break;
    case 3: hipLaunchKernelGGL(k_staples<3>, g, b, 0, s, st, re, links, T, G); break;
    case 4: hipLaunchKernelGGL(k_staples<4>, g, b, 0, s, st, re, links, T, G); break;
    case 5: hipLaunchKernelGGL(k_staples<5>, g, b, 0, s, st, re, links, T, G); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_loops(int32_t *out, int kind, int k, int l, const uint16_t *links, const uint16_t *lines,
                        const Tables &T, const Geom &G, uint32_t nmax_t, uint32_t nmax_x, hipStream_t s) {
  hipLaunchKernelGGL(k_loops, dim3(grid_for(G.V)), dim3(kBlock), 0, s, out, kind, k, l, links, lines, T, G,
                     nmax_t, nmax_x);
  return hipGetLastError();
}

hipError_t launch_lines(uint16_t *lines, const uint16_t *links, const Tables &T, const Geom &G, uint32_t nmax_t,
                        uint32_t nmax_x, hipStream_t s) {
  hipLaunchKernelGGL(k_lines, dim3(grid_for(G.V)), dim3(kBlock), 0, s, lines, links, T, G, nmax_t, nmax_x);
  return hipGetLastError();
}

uint32_t obs_blocks(const Geom &G) { return grid_for(G.V, kObsMaxBlocks); }

hipError_t launch_measure_tree(double *out, double *partials, const uint16_t *links, const uint16_t *lines,
                               const Tables &T, const Geom &G, uint32_t wk, uint32_t wl, uint32_t nmax_t,
                               uint32_t nmax_x, hipStream_t s) {
  const uint32_t nb = obs_blocks(G);
  hipLaunchKernelGGL(k_obs_local, dim3(nb), dim3(kBlock), 0, s, partials, links, T, G);
  if (wk && wl) hipLaunchKernelGGL(k_obs_wilson, dim3(nb, wk * wl), dim3(kBlock), 0, s, partials, lines, T, G, wl,
                                   nmax_t, nmax_x);
  const uint32_t nobs = 4 + wk * wl;
  hipLaunchKernelGGL(k_obs_final, dim3((nobs + 63) / 64), dim3(64), 0, s, out, partials, G, nobs, nb);
  return hipGetLastError();
}

hipError_t launch_measure_ordered(double *out, const uint16_t *links, const uint16_t *lines, const Tables &T,
                                  const Geom &G, uint32_t wk, uint32_t wl, uint32_t nmax_t, uint32_t nmax_x,
                                  hipStream_t s) {
  hipLaunchKernelGGL(k_obs_ordered, dim3(1), dim3(64), 0, s, out, links, lines, T, G, wk, wl, nmax_t, nmax_x);
  return hipGetLastError();
}

// The device work space of the entry points that allocate nothing
// (dsa_chain_measure_noalloc): one array per GPU, loaded with the code object
// and held until the process ends.  Host code never passes `scratch` itself
// to a kernel (on the host it names a stand-in, not device memory): it passes
// the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

hipError_t scratch_region(void **ptr) {
  static std::mutex m;
  static std::map<int, void *> addr;  // device -> address of its copy of `scratch`
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(m);
  auto it = addr.find(dev);
  if (it == addr.end()) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    it = addr.emplace(dev, p).first;
  }
  *ptr = it->second;
  return hipSuccess;
}

}  // namespace dsa
