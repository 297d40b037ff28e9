// llr.hip -- energy-window constrained Metropolis and the Robbins-Monro
// update of the LLR method (llr.cpp:178-250, llr2d.cpp:178-259), on the
// device.
//
// The reference visits links serially; every hit is tested against the
// window on a running energy.  Here the links of one colour pass (which share
// no plaquette and no rectangle) first make their NHIT tilted hits
// (weight exp(-a dE_w - bg dE_bg)) in parallel, each producing a proposal and
// its energy change; then, per chain, the proposals are accepted in ascending
// site order as long as the running energy stays in [E_lo, E_hi] (drive-in:
// as long as it moves closer to the window centre), and refused links are
// restored.  Each link's step is a reversible kernel restricted to the window
// by rejection, so the sweep keeps the constrained measure.  Energies are
// integers in units of 2^-32 so that sums do not depend on their order.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "dsa_math.hpp"
#include "kernels.hpp"

namespace dsa {

namespace {

__device__ inline long long wave_sum_i64(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}


// a value that is the same in every lane, made known to the compiler as
// such (it then lives in scalar registers and its branches are scalar)
__device__ inline long long uniform64(long long v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((unsigned long long)v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// The sum over each quad of lanes (4l .. 4l + 3), in every lane of the quad:
// DPP quad permutes (lane ^ 1, then lane ^ 2), no LDS round trip.  Every
// lane of a quad must be active.
__device__ inline int32_t quad_sum(int32_t v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xf, 0xf, false);  // quad_perm [1, 0, 3, 2]
  v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xf, 0xf, false);  // quad_perm [2, 3, 0, 1]
  return v;
}

template <int CTRL>
__device__ inline long long dpp64(long long x) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)x, CTRL, 0xf, 0xf, false);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)((unsigned long long)x >> 32), CTRL, 0xf, 0xf, false);
  return (long long)(((unsigned long long)hi << 32) | lo);
}

__device__ inline long long quad_sum64(long long v) {
  v += dpp64<0xB1>(v);
  v += dpp64<0x4E>(v);
  return v;
}

// lane k's value (k uniform over the wave)
__device__ inline long long readlane64(long long v, uint32_t k) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)k);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((unsigned long long)v >> 32), (int)k);
  return (long long)(((unsigned long long)hi << 32) | lo);
}

__device__ inline bool llr_active(uint32_t phase) { return phase == kLlrSeed || phase == kLlrRM; }

constexpr int kChanged = 0x80;  // flag in a proposal's hit count: it moved its link

__device__ inline bool llr_inside(const LlrChain &LC, long long E1, long long E2) {
  if (LC.axis == 3) return E1 >= LC.lo && E1 <= LC.hi && E2 >= LC.lo2 && E2 <= LC.hi2;
  const long long Ew = LC.axis == 1 ? E1 : E2;
  return Ew >= LC.lo && Ew <= LC.hi;
}

// drive-in: is the energy outside the window (box) at the start of a pass?
__device__ inline bool llr_drive(const LlrChain &LC, const LlrState &S) {
  return S.phase == kLlrSeed && !llr_inside(LC, S.S1, S.S2);
}

__device__ inline long long llabs_(long long v) { return v < 0 ? -v : v; }

// The acceptance rule of one pass: windowed energies (E1, E2) before (b) and
// after (a) a link's proposal.  Axes 1, 2: inside the window, or during
// drive-in closer to its centre (llr.cpp:203-212).  Axis 3: inside the box,
// or during drive-in closer to the cell centre in units of the half widths
// (squared distances; llr2d.cpp:200-213 compares hypot).
struct LlrRule {
  int axis;
  bool drive;
  long long lo, hi, lo2, hi2, mid, mid2;
  double hw, hw2;
  __device__ double dist2(long long E1, long long E2) const {
    const double x = from_fix(E1 - mid) / hw, y = from_fix(E2 - mid2) / hw2;
    return x * x + y * y;
  }
  __device__ bool operator()(long long E1b, long long E2b, long long E1a, long long E2a) const {
    if (axis == 3) {
      if (drive) return dist2(E1a, E2a) < dist2(E1b, E2b);
      return E1a >= lo && E1a <= hi && E2a >= lo2 && E2a <= hi2;
    }
    const long long Eb = axis == 1 ? E1b : E2b, Ea = axis == 1 ? E1a : E2a;
    if (drive) return llabs_(Ea - mid) < llabs_(Eb - mid);
    return Ea >= lo && Ea <= hi;
  }
};

__device__ inline LlrRule llr_rule_of(const LlrChain &LC, const LlrState &S) {
  LlrRule R;
  R.axis = LC.axis;
  R.drive = llr_drive(LC, S);
  R.lo = LC.lo;
  R.hi = LC.hi;
  R.lo2 = LC.lo2;
  R.hi2 = LC.hi2;
  R.mid = LC.lo + (LC.hi - LC.lo) / 2;
  R.mid2 = LC.lo2 + (LC.hi2 - LC.lo2) / 2;
  R.hw = 0.5 * LC.W;
  R.hw2 = 0.5 * LC.W2;
  return R;
}

// Bound on |exact - class| for A1 dS1 + A2 dS2 computed from class sums
// instead of the element-value sums of update (the within-class spread of the
// table's ReTr values, the quantisation of the class values, and the rounding
// of the element sums and of the combination, the latter with a factor 4).
__device__ inline double llr_margin(const LlrChain &LC, double A1, double A2) {
  A1 = fabs(A1);
  A2 = fabs(A2);
  const double vals = A1 * LC.ce1 + A2 * LC.ce2;
  const double rounding = A1 * LC.cf1 + A2 * LC.cf2 + 8.0 * 0x1p-53 * (A1 * LC.cr1 + A2 * LC.cr2);
  return LC.cmscale * (2.0 * vals * (1.0 + 1e-6) + 8.0 * rounding);
}

// Certified comparison exp(-x) > u (u in [0, 1)) from x_c with |x - x_c| <= M:
// 1 true, 0 false, -1 undecided.  Below the band exp(-x) >= 1 > u.  Above
// it, v_exp_f32 (2^y, ~1 ulp) on the float-rounded argument is within 5e-6
// relative of exp(-x_c) for x_c < 60, so it decides with 1.5e-5 of slack,
// exp(-x) < exp(-37) < 1e-15 <= u is false, and what is left is decided
// with dsa_exp (within one ulp).  Inside the band exp(-x) >= 1 - 2M.
__device__ inline int exp_gt_certified(double xc, double M, double u) {
  if (xc + M <= 0.0) return 1;
  if (xc - M > 0.0) {
    if (xc - M > 37.0 && u >= 1e-15) return 0;
    if (xc < 60.0) {
      const double e = (double)__builtin_amdgcn_exp2f((float)(-xc * 1.4426950408889634)), w = 1.02 * M + 0x1p-16;
      if (u < e * (1.0 - w)) return 1;
      if (u >= e * (1.0 + w)) return 0;
    }
    const double ed = dsa_exp(-xc), wd = 1.02 * M + 0x1p-46;
    if (wd >= 0.5) return -1;
    return u < ed * (1.0 - wd) ? 1 : (u >= ed * (1.0 + wd) ? 0 : -1);
  }
  return u < 1.0 - (2.02 * M + 0x1p-46) ? 1 : -1;
}

// Certified sign test y > 0 from y_c with |y - y_c| <= M.
__device__ inline int gt0_certified(double yc, double M) { return yc - M > 0.0 ? 1 : (yc + M <= 0.0 ? 0 : -1); }

template <int D>
__global__ __launch_bounds__(kBlock) void k_llr_pass(const LlrChain *chains, LlrState *st, const BatchTask *tasks) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int32_t s_rcq[64];  // class values (2^-24), indexed by the byte tables
  // class path: the hits' proposal indices and acceptance uniforms, drawn by
  // the whole block before the links are updated (a block of fewer links than
  // lanes draws them several lanes per link)
  __shared__ double s_u2[kNHit][kBlock];
  __shared__ uint16_t s_jp[kNHit][kBlock];
  const BatchTask tk = tasks[blockIdx.x];
  const LlrChain &LC = chains[tk.chain];
  const LlrState &S = st[tk.chain];
  if (!llr_active(S.phase)) return;  // uniform over the block
  const BatchChain &BC = LC.base;
  const uint32_t P = BC.T.P, np = BC.T.npool;
  double *s_retr = reinterpret_cast<double *>(smem);
  uint16_t *s_inv = reinterpret_cast<uint16_t *>(s_retr + P);
  uint16_t *s_pool = s_inv + P;
  // class path: class of each element and of pool[j] * t, after the element tables
  const bool cls = LC.cls != 0, fld = cls && LC.field != 0;
  uint8_t *s_cls = smem + pad16(llr_lds_bytes(BC.T, false));
  uint8_t *s_cprod = s_cls + pad16(P);
  // field: the GF(4) tables after the class tables
  uint32_t *s_fm = reinterpret_cast<uint32_t *>(s_cprod + pad16((size_t)np * P));
  uint32_t *s_fmi = reinterpret_cast<uint32_t *>(reinterpret_cast<uint8_t *>(s_fm) + pad16(4 * (size_t)P));
  uint16_t *s_fkey = reinterpret_cast<uint16_t *>(reinterpret_cast<uint8_t *>(s_fmi) + pad16(4 * (size_t)P));
  if (fld) {  // uniform over the block
    constexpr uint32_t nk = kGf4Keys * sizeof(uint16_t) / 16;
    const u32x4 *ks = reinterpret_cast<const u32x4 *>(BC.T.fkey);
    Staged<5, uint32_t> fm, fmi;
    Staged<2, u32x4> fk;
    fm.load(BC.T.fmat, P);
    fmi.load(BC.T.finv, P);
    fk.load(ks, nk);
    fm.store(s_fm, BC.T.fmat, P);
    fmi.store(s_fmi, BC.T.finv, P);
    fk.store(reinterpret_cast<u32x4 *>(s_fkey), ks, nk);
  }
  {
    // every load of the block's staging issued before the first LDS store
    Staged<5, double> g_retr;
    Staged<5, uint16_t> g_inv;
    Staged<1, uint16_t> g_pool;
    Staged<5, uint8_t> g_cls;
    Staged<1, int32_t> g_rcq;
    Staged<10, u32x4> g_tab;
    const uint32_t n16 = cls ? (uint32_t)(pad16((size_t)np * P) / 16) : 0u;
    const uint32_t ncls = cls ? BC.T.ncls : 0u, pcls = cls ? P : 0u;
    const u32x4 *src = reinterpret_cast<const u32x4 *>(BC.T.cprod);
    g_retr.load(BC.T.retr, P);
    g_inv.load(BC.T.inv, P);
    g_pool.load(BC.T.pool, np);
    g_cls.load(BC.T.cls, pcls);
    g_rcq.load(BC.T.rcq, ncls);
    g_tab.load(src, n16);
    g_retr.store(s_retr, BC.T.retr, P);
    g_inv.store(s_inv, BC.T.inv, P);
    g_pool.store(s_pool, BC.T.pool, np);
    g_cls.store(s_cls, BC.T.cls, pcls);
    g_rcq.store(s_rcq, BC.T.rcq, ncls);
    g_tab.store(reinterpret_cast<u32x4 *>(s_cprod), src, n16);
  }
  if (cls) {
    const uint32_t sl = (uint32_t)S.sweep, sh = (uint32_t)(S.sweep >> 32), n = tk.count;
    const double npd = (double)np;
    for (uint32_t j = threadIdx.x; j < n * (uint32_t)kNHit; j += blockDim.x) {
      const uint32_t l = j % n, h = j / n;
      const uint32_t lk = gmem(BC.sites)[tk.offset + l] * (uint32_t)D + (uint32_t)tk.mu;
      const U4 w = philox4x32_10(U4{lk, h, sl, sh}, BC.seed, kKeyLLR);
      s_jp[h][l] = (uint16_t)(npd * u53(w.x, w.y));
      s_u2[h][l] = u53(w.z, w.w);
    }
  }
  __syncthreads();

  constexpr int NS = 2 * (D - 1), NR = 6 * (D - 1), NL = NS + NR;
  const TabView T{BC.T.mult, s_inv, P};
  const ClsNarrow Lk{s_cprod, s_cls, s_rcq};
  const Geom &G = BC.G;
  const int axis = LC.axis;
  const LlrRule R = llr_rule_of(LC, S);
  const bool drive = R.drive;
  // drive-in direction: axes 1, 2 toward the window; axis 3 down the gradient
  // of the squared distance to the cell centre at the start of the pass
  const double sign = from_fix(axis == 1 ? S.S1 : S.S2) < LC.Ewin ? 1.0 : -1.0;
  const double gx = from_fix(S.S1 - R.mid) / R.hw / R.hw, gy = from_fix(S.S2 - R.mid2) / R.hw2 / R.hw2;
  const double a = S.a, a2 = S.a2, bg = LC.bg;
  // the class path's test: coefficients of (dS1, dS2) and their bound
  double A1, A2;
  if (drive) {
    A1 = axis == 3 ? gx : axis == 1 ? sign : 0.0;
    A2 = axis == 3 ? gy : axis == 2 ? sign : 0.0;
  } else {
    A1 = axis == 2 ? bg : a;
    A2 = axis == 3 ? a2 : axis == 1 ? bg : a;
  }
  const double margin = cls ? llr_margin(LC, A1, A2) : 0.0;
  const double A1q = A1 * ClsNarrow::kScale, A2q = A2 * ClsNarrow::kScale;
  const uint32_t sw_lo = (uint32_t)S.sweep, sw_hi = (uint32_t)(S.sweep >> 32);
  const double npool_d = (double)np;
  if (fld && blockDim.x >= 4u * tk.count) {
    // A small task (a quarter of the block's lanes or fewer): four lanes per
    // link.  Lane q of a link builds transverse direction q's eight products
    // (build_t_field_dir), and each hit's class sums are split over the
    // lanes and summed across them; every lane of the link then holds the
    // same sums and takes the same decision.  Integer sums do not depend on
    // the order, and the element-value re-decision gathers all products in
    // the reference's order first, so the results are those of one lane per
    // link.
    const uint32_t t = threadIdx.x >> 2, q = threadIdx.x & 3u;
    if (t < tk.count) {  // a link's four lanes are all in or all out
      const uint32_t idx = tk.offset + t;
      const uint32_t i = gmem(BC.sites)[idx];
      const int mu = tk.mu;
      const Site<D> s = decode<D>(G, i);
      const size_t lp = (size_t)mu * G.V + lpos(G, i, s.par);
      uint32_t b = gmem(BC.links)[lp];
      const uint32_t b0 = b;
      const bool own = q < (uint32_t)(D - 1);  // holds direction q's products
      const uint32_t lane0 = (threadIdx.x & 63u) & ~3u;
      const FieldView FV{s_fm, s_fmi, s_fkey};
      uint32_t t8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (own) {
        const Gf4R Rb = FV.R(b);
        const Gf4L umu_p = FV.L(U<D>(BC.links, G, s, mu, mu, 1, mu, 0));
        const Gf4R Cm = gf4_right(gf4_mul(FV.L(U<D>(BC.links, G, s, mu, mu, -1, mu, 0)), Rb));
        build_t_field_dir<D>(BC.links, G, FV, s, mu, (int)q, Rb, umu_p, Cm, t8);
      }
      // the link's class sums through `row`: this lane's terms (2 plaquette,
      // 6 rectangle), then summed over the four lanes
      auto sums4 = [&](ClsNarrow::Row row, int32_t &rc, int32_t &mc) {
        int32_t x = 0, y = 0;
        if (own) {
          uint32_t off[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) off[k] = Lk.idx(row, t8[k]);
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const int32_t v = Lk.at(off[k]);
            if (k < 2) x += v; else y += v;
          }
        }
        rc = quad_sum(x);
        mc = quad_sum(y);
      };
      uint32_t nacc = 0;
      long long d1 = 0, d2 = 0;
      int32_t r1c, m1c;
      sums4(Lk.row0(), r1c, m1c);
      for (int h = 0; h < kNHit; ++h) {
        const double u2 = s_u2[h][t];
        const uint32_t jp = s_jp[h][t];
        int32_t r1nc, m1nc;
        sums4(Lk.row(jp, P), r1nc, m1nc);
        const double xc = A1q * ((double)r1nc - (double)r1c) + A2q * ((double)m1nc - (double)m1c);
        const double M = margin + 0x1p-50 * fabs(xc);
        int dec = drive ? (axis == 3 ? gt0_certified(-xc, M) : gt0_certified(xc, M)) : exp_gt_certified(xc, M, u2);
        if (dec < 0) {  // decide exactly as the element path does: every product, in order
          uint32_t tt[NL];
#pragma unroll
          for (int jj = 0; jj < D - 1; ++jj) {
            const int src = (int)lane0 + jj;
            tt[2 * jj] = (uint32_t)__shfl((int)t8[0], src);
            tt[2 * jj + 1] = (uint32_t)__shfl((int)t8[1], src);
#pragma unroll
            for (int k = 0; k < 6; ++k) tt[NS + 6 * jj + k] = (uint32_t)__shfl((int)t8[2 + k], src);
          }
          const uint32_t bnew = T.M(b, s_pool[jp]), bi = T.I(b);
          double e1 = 0, e2 = 0, e1n = 0, e2n = 0;
#pragma unroll
          for (int k = 0; k < NL; ++k) {
            const double v = s_retr[T.M(T.M(b, tt[k]), bi)], vn = s_retr[T.M(T.M(bnew, tt[k]), bi)];
            if (k < NS) {
              e1 += v;
              e1n += vn;
            } else {
              e2 += v;
              e2n += vn;
            }
          }
          const double dS1 = e1n - e1, dS2 = e2n - e2;
          bool accept;
          if (axis == 3) {
            accept = drive ? gx * dS1 + gy * dS2 < 0.0 : dsa_exp(-a * dS1 - a2 * dS2) > u2;
          } else {
            const double dEw = axis == 1 ? dS1 : dS2, dEbg = axis == 1 ? dS2 : dS1;
            accept = drive ? sign * dEw > 0.0 : dsa_exp(-a * dEw - bg * dEbg) > u2;
          }
          dec = accept ? 1 : 0;
        }
        if (dec) {
          const uint32_t p = s_pool[jp];
          b = T.M(b, p);
          if (own) {
#pragma unroll
            for (int k = 0; k < 8; ++k) t8[k] = T.M(t8[k], p);
          }
          r1c = r1nc;
          m1c = m1nc;
          ++nacc;
        }
      }
      if (b != b0) {  // energy change in fixed point: s_k = t_k b^-1, this lane's terms, then summed
        const uint32_t bi = T.I(b);
        long long p1 = 0, p2 = 0, q1 = 0, q2 = 0;
        if (own) {
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const uint32_t sk = T.M(t8[k], bi);
            const long long vp = to_fix(s_retr[T.M(b, sk)]), vq = to_fix(s_retr[T.M(b0, sk)]);
            if (k < 2) {
              p1 += vp;
              q1 += vq;
            } else {
              p2 += vp;
              q2 += vq;
            }
          }
        }
        d1 = quad_sum64(p1 - q1);
        d2 = quad_sum64(p2 - q2);
      }
      if (q == 0) {
        if (b != b0) gmem(BC.links)[lp] = (uint16_t)b;
        gmem(LC.old)[idx] = (uint16_t)b0;
        gmem(LC.dE1)[idx] = d1;
        gmem(LC.dE2)[idx] = d2;
        gmem(LC.nacc)[idx] = (uint8_t)(nacc | (b != b0 ? (uint32_t)kChanged : 0u));
      }
    }
    return;
  }
  for (uint32_t t = threadIdx.x; t < tk.count; t += blockDim.x) {
    const uint32_t idx = tk.offset + t;
    const uint32_t i = gmem(BC.sites)[idx];
    const int mu = tk.mu;
    const Site<D> s = decode<D>(G, i);
    const size_t lp = (size_t)mu * G.V + lpos(G, i, s.par);
    uint32_t b = gmem(BC.links)[lp];
    const uint32_t b0 = b;
    const uint32_t link = i * (uint32_t)D + (uint32_t)mu;
    uint32_t nacc = 0;
    long long d1 = 0, d2 = 0;
    if (cls) {
      // t_k = s_k b, built directly: ReTr(b p s_k) = ReTr(t_k p), of the class
      // of pool[j] t_k; the element value ReTr(b s_k) is retr[b t_k b^-1]
      uint32_t tt[NL];
      if (fld)
        build_t_field<D>(BC.links, G, FieldView{s_fm, s_fmi, s_fkey}, s, mu, b, tt);
      else
        build_t<D>(BC.links, G, T, s, mu, b, tt);
      auto sums = [&](ClsNarrow::Row row, int32_t &rc, int32_t &mc) {
        int32_t x = 0, y = 0;
#pragma unroll
        for (int k = 0; k < NL; ++k) {
          const int32_t v = Lk.val(row, tt[k]);
          if (k < NS) x += v; else y += v;
        }
        rc = x;
        mc = y;
      };
      int32_t r1c, m1c;
      sums(Lk.row0(), r1c, m1c);
      for (int h = 0; h < kNHit; ++h) {
        const double u2 = s_u2[h][t];
        const uint32_t jp = s_jp[h][t];
        int32_t r1nc, m1nc;
        sums(Lk.row(jp, P), r1nc, m1nc);
        const double xc = A1q * ((double)r1nc - (double)r1c) + A2q * ((double)m1nc - (double)m1c);
        const double M = margin + 0x1p-50 * fabs(xc);
        // drive-in: axes 1, 2 accept sign dEw > 0; axis 3 accepts gx dS1 + gy dS2 < 0
        int dec = drive ? (axis == 3 ? gt0_certified(-xc, M) : gt0_certified(xc, M)) : exp_gt_certified(xc, M, u2);
        if (dec < 0) {  // decide exactly as the element path does, from the same sums
          const uint32_t bnew = T.M(b, s_pool[jp]), bi = T.I(b);
          double e1 = 0, e2 = 0, e1n = 0, e2n = 0;
#pragma unroll
          for (int k = 0; k < NL; ++k) {
            const double v = s_retr[T.M(T.M(b, tt[k]), bi)], vn = s_retr[T.M(T.M(bnew, tt[k]), bi)];
            if (k < NS) {
              e1 += v;
              e1n += vn;
            } else {
              e2 += v;
              e2n += vn;
            }
          }
          const double dS1 = e1n - e1, dS2 = e2n - e2;
          bool accept;
          if (axis == 3) {
            accept = drive ? gx * dS1 + gy * dS2 < 0.0 : dsa_exp(-a * dS1 - a2 * dS2) > u2;
          } else {
            const double dEw = axis == 1 ? dS1 : dS2, dEbg = axis == 1 ? dS2 : dS1;
            accept = drive ? sign * dEw > 0.0 : dsa_exp(-a * dEw - bg * dEbg) > u2;
          }
          dec = accept ? 1 : 0;
        }
        if (dec) {
          const uint32_t p = s_pool[jp];
          b = T.M(b, p);
#pragma unroll
          for (int k = 0; k < NL; ++k) tt[k] = T.M(tt[k], p);
          r1c = r1nc;
          m1c = m1nc;
          ++nacc;
        }
      }
      if (b != b0) {  // energy change in fixed point: s_k = t_k b^-1
        gmem(BC.links)[lp] = (uint16_t)b;
        const uint32_t bi = T.I(b);
        long long p1 = 0, p2 = 0, q1 = 0, q2 = 0;
#pragma unroll
        for (int k = 0; k < NL; ++k) {
          const uint32_t sk = T.M(tt[k], bi);
          const long long vp = to_fix(s_retr[T.M(b, sk)]), vq = to_fix(s_retr[T.M(b0, sk)]);
          if (k < NS) {
            p1 += vp;
            q1 += vq;
          } else {
            p2 += vp;
            q2 += vq;
          }
        }
        d1 = p1 - q1;
        d2 = p2 - q2;
      }
    } else {
      uint32_t stp[NS], re[NR];
      build_staples<D>(BC.links, G, T, s, mu, stp, re);
      // element sums (reference order) in double and in fixed point
      double r1 = 0, m1 = 0;
      long long q1 = 0, q2 = 0;
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const double v = s_retr[T.M(b, stp[j])];
        r1 += v;
        q1 += to_fix(v);
      }
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        const double v = s_retr[T.M(b, re[q])];
        m1 += v;
        q2 += to_fix(v);
      }
      for (int h = 0; h < kNHit; ++h) {
        const U4 w = philox4x32_10(U4{link, (uint32_t)h, sw_lo, sw_hi}, BC.seed, kKeyLLR);
        const double u1 = u53(w.x, w.y), u2 = u53(w.z, w.w);
        const uint32_t bnew = T.M(b, s_pool[(uint32_t)(npool_d * u1)]);
        double r1n = 0, m1n = 0;
#pragma unroll
        for (int j = 0; j < NS; ++j) r1n += s_retr[T.M(bnew, stp[j])];
#pragma unroll
        for (int q = 0; q < NR; ++q) m1n += s_retr[T.M(bnew, re[q])];
        const double dS1 = r1n - r1, dS2 = m1n - m1;
        bool accept;
        if (axis == 3) {
          if (drive)
            accept = gx * dS1 + gy * dS2 < 0.0;  // greedy move toward the cell centre
          else
            accept = dsa_exp(-a * dS1 - a2 * dS2) > u2;  // llr2d.cpp:211-212
        } else {
          const double dEw = axis == 1 ? dS1 : dS2;
          const double dEbg = axis == 1 ? dS2 : dS1;
          if (drive)
            accept = sign * dEw > 0.0;  // greedy move toward the window (llr.cpp:208-209)
          else
            accept = dsa_exp(-a * dEw - bg * dEbg) > u2;  // llr.cpp:211-212
        }
        if (accept) {
          b = bnew;
          r1 = r1n;
          m1 = m1n;
          ++nacc;
        }
      }
      if (b != b0) {
        gmem(BC.links)[lp] = (uint16_t)b;
        long long p1 = 0, p2 = 0;
#pragma unroll
        for (int j = 0; j < NS; ++j) p1 += to_fix(s_retr[T.M(b, stp[j])]);
#pragma unroll
        for (int q = 0; q < NR; ++q) p2 += to_fix(s_retr[T.M(b, re[q])]);
        d1 = p1 - q1;
        d2 = p2 - q2;
      }
    }
    gmem(LC.old)[idx] = (uint16_t)b0;
    gmem(LC.dE1)[idx] = d1;
    gmem(LC.dE2)[idx] = d2;
    gmem(LC.nacc)[idx] = (uint8_t)(nacc | (b != b0 ? (uint32_t)kChanged : 0u));
  }
}

// A proposal that left its link as it was changes no energy, and its rule
// does not depend on where it falls: with the window rule the energy is
// inside the window when it is met (accepted), with the drive-in rule it
// does not move closer (refused).  Such proposals are settled apart and do
// not break the runs below.
//
// One block of kDecide lanes per chain, a segment of kDecideSeg proposals at
// a time, kDecidePer consecutive ones per lane.  Block-wide inclusive prefix
// sums of the energy changes give every proposal the running energy if every
// proposal from the segment's start up to it is accepted: the first one whose
// rule fails ends the run of acceptances, found with one block-wide minimum.
// From there one wave walks the rest of the segment by runs of acceptances
// and refusals (a refused proposal leaves the energy unchanged), 64
// proposals at a time, reading the prefix sums from LDS.  This is the
// sequential rule exactly: each decision sees the energy after all earlier
// accepted proposals.
constexpr int kDecide = 512;
constexpr int kDecideWaves = kDecide / 64;
constexpr int kDecidePer = 8;
constexpr int kDecideSeg = kDecide * kDecidePer;

// Inclusive prefix sums over the wave: DPP row shifts (1, 2, 4, 8) scan each
// row of 16 lanes, then the three lower rows' totals are added by row.
// Lanes shifted in from outside a row read 0.
template <int CTRL>
__device__ inline long long dpp_shift64(long long v) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)((unsigned long long)v >> 32), CTRL, 0xf, 0xf, false);
  return (long long)(((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo);
}

__device__ inline long long scan_i64(long long v, int lane) {
  v += dpp_shift64<0x111>(v);  // row_shr:1
  v += dpp_shift64<0x112>(v);  // row_shr:2
  v += dpp_shift64<0x114>(v);  // row_shr:4
  v += dpp_shift64<0x118>(v);  // row_shr:8
  const long long t0 = readlane64(v, 15), t1 = readlane64(v, 31), t2 = readlane64(v, 47);
  const int row = lane >> 4;
  return v + (row == 0 ? 0 : row == 1 ? t0 : row == 2 ? t0 + t1 : t0 + t1 + t2);
}

// Inclusive prefix sums within each row of 16 lanes (lanes shifted in from
// outside a row read 0).
__device__ inline long long row_scan_i64(long long v) {
  v += dpp_shift64<0x111>(v);
  v += dpp_shift64<0x112>(v);
  v += dpp_shift64<0x114>(v);
  v += dpp_shift64<0x118>(v);
  return v;
}

__device__ inline int row_scan_i32(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
  return v;
}

__device__ inline int scan_i32(int v, int lane) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
  const int t0 = __builtin_amdgcn_readlane(v, 15), t1 = __builtin_amdgcn_readlane(v, 31),
            t2 = __builtin_amdgcn_readlane(v, 47);
  const int row = lane >> 4;
  return v + (row == 0 ? 0 : row == 1 ? t0 : row == 2 ? t0 + t1 : t0 + t1 + t2);
}

// The smallest idx over the lanes whose pred holds (or n), for idx
// increasing with the lane index.  Three slots rotate: the slot of call k is
// reset after call k-2's readers have all passed call k-1's barrier, and
// before call k's writers (they pass call k-1's barrier after the reset in
// call k-1).
__device__ inline uint32_t block_first(bool pred, uint32_t idx, uint32_t n, unsigned *slots, uint32_t &k) {
  const unsigned long long bm = __ballot(pred);
  const unsigned cand = bm ? (unsigned)__builtin_amdgcn_readlane((int)idx, (int)__builtin_ctzll(bm)) : 0u;
  if ((threadIdx.x & 63u) == 0 && bm) atomicMin(&slots[k], cand);
  __syncthreads();
  const unsigned r = (unsigned)__builtin_amdgcn_readfirstlane((int)slots[k]);
  const uint32_t nx = k == 2 ? 0 : k + 1;
  if (threadIdx.x == 0) slots[nx == 2 ? 0 : nx + 1] = 0xffffffffu;
  k = nx;
  return r < n ? r : n;
}

// The contested walk under the window rule (every pass that is not a
// drive-in pass): a changed proposal k is accepted iff the energy after it is
// inside the window (axes 1, 2: the ruling energy; axis 3: both, the box),
// i.e. (unsigned) (E + e_k - lo) <= hi - lo.  With q the inclusive and x the
// exclusive prefix sums of the changes, a run of acceptances from lane s, met
// with energy E_s, has energy lo + o + q_k after lane k (o = E_s - lo - x_s),
// so it ends at the first changed lane whose o + q_k leaves [0, hi - lo]: one
// add and one compare per lane against uniform values.  (Unchanged lanes
// never end it: the energy after them is the energy before, inside.)  A run
// of refusals from a failing lane f leaves E unchanged and ends at the first
// changed lane after f whose own change keeps E inside.  The loop keeps only
// what the next run needs (the ruling energies, the lanes still undecided)
// and marks each run's first lane in T, so that a lane's outcome is the
// parity of the marks at or below it: even for an acceptance run, odd for a
// refusal run.  The other energy and the hits of the accepted proposals are
// summed per lane over the chunks and across the wave once at the end.
template <int AX>
__device__ void walk_window(const LlrRule &R, uint32_t n, uint32_t f0, const long long *sP1, const long long *sP2,
                            const int *sPN, const uint8_t *sChg, const uint32_t *sPos, const uint16_t *sOld,
                            uint16_t *links, long long &E1, long long &E2, unsigned long long &acc) {
  const int lane = (int)(threadIdx.x & 63u);
  n = (uint32_t)__builtin_amdgcn_readfirstlane((int)n);
  f0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)f0);
  E1 = uniform64(E1);
  E2 = uniform64(E2);
  const unsigned long long Wa = (unsigned long long)(R.hi - R.lo), Wb = (unsigned long long)(R.hi2 - R.lo2);
  const unsigned long long upto = ~0ull >> (63 - lane);  // lanes 0..lane
  // the ruling energies minus lo, uniform over the wave
  long long xa = (AX == 2 ? E2 : E1) - R.lo, xb = AX == 3 ? E2 - R.lo2 : 0;
  long long so = 0;  // per lane: the other energy's gain by the accepted proposals
  int sn = 0;        // per lane: hits of the accepted proposals
  // a chunk's values, loaded while the chunk before it is walked: inclusive
  // (q) and exclusive (x) prefix sums (lane 0's exclusive ones being the
  // sums before the chunk), the changed flag and the link to restore.  Every
  // load is in bounds and unconditional, so that they wait together; lanes
  // past the segment read values that are never used.
  struct Chunk {
    long long q1, q2, x1, x2;
    int qn, xn;
    uint32_t ch, pos, old;
  };
  auto load_chunk = [&](uint32_t b) {
    Chunk c;
    const uint32_t k = b + (uint32_t)lane < (uint32_t)kDecideSeg ? b + (uint32_t)lane : (uint32_t)kDecideSeg - 1;
    const uint32_t kp = k ? k - 1 : 0;
    c.q1 = sP1[k];
    c.q2 = sP2[k];
    c.x1 = sP1[kp];
    c.x2 = sP2[kp];
    c.qn = sPN[k];
    c.xn = sPN[kp];
    c.ch = sChg[k];
    c.pos = sPos[k];
    c.old = sOld[k];
    if (!k) {
      c.x1 = 0;
      c.x2 = 0;
      c.xn = 0;
    }
    return c;
  };
  Chunk nxt = load_chunk(f0 & ~63u);
  for (uint32_t base = f0 & ~63u; base < n; base += 64) {  // n <= kDecideSeg: base + 63 < kDecideSeg
    const uint32_t cn = n - base < 64 ? n - base : 64;
    const bool valid = (uint32_t)lane < cn;
    const Chunk cur = nxt;
    nxt = load_chunk(base + 64);
    const long long q1 = cur.q1, q2 = cur.q2, x1 = cur.x1, x2 = cur.x2;
    const int qn = cur.qn, xn = cur.xn;
    const uint32_t ch = cur.ch;
    const uint32_t s_init = base < f0 ? f0 - base : 0;  // lanes before it are settled
    // the ruling energy: S1 on axes 1 and 3, S2 on axis 2; the box adds S2;
    // the other one (axes 1, 2) is summed per lane
    const long long qa = AX == 2 ? q2 : q1, xa_ = AX == 2 ? x2 : x1;
    const long long ea = qa - xa_, eb = q2 - x2;
    // changed lanes not yet decided, uniform over the wave
    unsigned long long rem = __ballot(valid & (ch != 0)) & (~0ull << s_init);
    unsigned long long T = 0;
    long long o = xa - readlane64(xa_, s_init);
    long long ob = AX == 3 ? xb - readlane64(x2, s_init) : 0;
    bool tail_refused = false;
    for (;;) {
      bool out = (unsigned long long)(qa + o) > Wa;
      if (AX == 3) out = out | ((unsigned long long)(q2 + ob) > Wb);
      const unsigned long long fm = __ballot(out) & rem;
      if (!fm) break;  // the rest of the chunk is accepted
      const uint32_t f = (uint32_t)__builtin_ctzll(fm);
      T |= 1ull << f;
      xa = readlane64(o + xa_, f);  // the energy met by lane f
      if (AX == 3) xb = readlane64(ob + x2, f);
      bool in = (unsigned long long)(ea + xa) <= Wa;
      if (AX == 3) in = in & ((unsigned long long)(eb + xb) <= Wb);
      rem &= ~1ull << f;  // the lanes after f (none for f = 63)
      const unsigned long long gm = __ballot(in) & rem;
      if (!gm) {  // the rest of the chunk is refused: the energy stays
        tail_refused = true;
        break;
      }
      const uint32_t g = (uint32_t)__builtin_ctzll(gm);
      T |= 1ull << g;
      rem &= ~0ull << g;
      o = xa - readlane64(xa_, g);
      if (AX == 3) ob = xb - readlane64(x2, g);
    }
    if (!tail_refused) {
      xa = o + readlane64(qa, cn - 1);
      if (AX == 3) xb = ob + readlane64(q2, cn - 1);
    }
    const bool refused = (__popcll(T & upto) & 1) != 0;
    const bool mine = valid && (uint32_t)lane >= s_init;
    if (mine && ch && refused) gmem(links)[cur.pos] = (uint16_t)cur.old;
    if (mine && !refused) {  // unchanged lanes add nothing
      if (AX == 1) so += q2 - x2;
      if (AX == 2) so += q1 - x1;
      sn += qn - xn;
    }
  }
  if (AX != 3) so = uniform64(wave_sum_i64(so));
  const long long hits = uniform64(wave_sum_i64(sn));
  if (AX == 1) E2 += so;
  if (AX == 2) E1 += so;
  if (AX != 2) E1 = xa + R.lo;
  if (AX == 2) E2 = xa + R.lo;
  if (AX == 3) E2 = xb + R.lo2;
  acc += (unsigned long long)hits;
}

__global__ __launch_bounds__(kDecide) void k_llr_decide(const LlrChain *chains, LlrState *st, const uint32_t *idx,
                                                        uint32_t pass) {
  __shared__ long long sP1[kDecideSeg], sP2[kDecideSeg];
  __shared__ int sPN[kDecideSeg];
  __shared__ uint32_t sPos[kDecideSeg];
  __shared__ uint16_t sOld[kDecideSeg];
  __shared__ uint8_t sChg[kDecideSeg];
  __shared__ long long sW1[16], sW2[16];  // per wave; a row of 16 lanes reads them all
  __shared__ int sWN[16], sWT[16];
  __shared__ unsigned slots[3];
  __shared__ long long sE[2];
  __shared__ unsigned long long sAcc;
  static_assert(kDecideWaves <= 16, "the waves' totals fit one row of 16 lanes");
  const uint32_t c = idx ? gmem(idx)[blockIdx.x] : blockIdx.x;
  const uint32_t tid = threadIdx.x;
  const int lane = (int)(tid & 63u), wv = (int)(tid >> 6);
  const LlrChain &LC = chains[c];
  LlrState &S = st[c];
  if (!llr_active(S.phase) || pass >= LC.npass) return;  // uniform over the block
  const PassRange pr = LC.passes[pass];
  const LlrRule R = llr_rule_of(LC, S);
  long long E1 = S.S1, E2 = S.S2;  // uniform over the block
  unsigned long long acc = 0;
  uint16_t *links = LC.base.links;
  if (tid < 3) slots[tid] = 0xffffffffu;
  if (tid < 16) {  // the slots past the last wave stay 0
    sW1[tid] = 0;
    sW2[tid] = 0;
    sWN[tid] = 0;
    sWT[tid] = 0;
  }
  uint32_t k = 0;
  // a lane's kDecidePer consecutive proposals of a segment, loaded one
  // segment ahead so that the next segment's loads are in flight while this
  // one is decided; past the pass's end they are unchanged proposals with no
  // hits
  struct In {
    long long d1[kDecidePer], d2[kDecidePer];
    uint32_t pos[kDecidePer];
    uint16_t old[kDecidePer];
    uint8_t raw[kDecidePer];
  };
  auto load_in = [&](uint32_t seg, In &v) {
#pragma unroll
    for (int i = 0; i < kDecidePer; ++i) {
      const uint32_t j = seg + tid * (uint32_t)kDecidePer + (uint32_t)i;
      const bool ok = j < pr.count;
      const uint32_t idx = pr.offset + (ok ? j : 0u);
      v.d1[i] = ok ? gmem(LC.dE1)[idx] : 0;
      v.d2[i] = ok ? gmem(LC.dE2)[idx] : 0;
      v.raw[i] = ok ? gmem(LC.nacc)[idx] : (uint8_t)0;
      v.pos[i] = ok ? gmem(LC.pos)[idx] : 0u;
      v.old[i] = ok ? gmem(LC.old)[idx] : (uint16_t)0;
    }
  };
  In nxt;
  load_in(0, nxt);
  for (uint32_t seg = 0; seg < pr.count; seg += kDecideSeg) {
    const uint32_t n = pr.count - seg < (uint32_t)kDecideSeg ? pr.count - seg : (uint32_t)kDecideSeg;
    const In cur = nxt;
    if (seg + kDecideSeg < pr.count) load_in(seg + kDecideSeg, nxt);
    bool chg[kDecidePer];
    int dn[kDecidePer];  // hits of proposals that moved their link
    long long l1 = 0, l2 = 0;
    int ln = 0, lt = 0;  // lane totals; lt: hits of the unchanged proposals
#pragma unroll
    for (int i = 0; i < kDecidePer; ++i) {
      chg[i] = (cur.raw[i] & kChanged) != 0;
      dn[i] = chg[i] ? cur.raw[i] & ~kChanged : 0;
      l1 += cur.d1[i];
      l2 += cur.d2[i];
      ln += dn[i];
      lt += chg[i] ? 0 : cur.raw[i];
    }
    if (R.drive && R.axis != 3) {
      // Drive-in on one window (axes 1, 2) with the energy E on one side of
      // the centre: a proposal gets strictly closer iff it moves toward the
      // centre and lands no further beyond it than E was before it.  If
      // accepting every changed proposal that moves toward the centre leaves
      // the energy on the same side (or at the centre), each of them met an
      // energy at least its own change away from the centre: the sequential
      // rule accepts exactly those and refuses the others, and the segment
      // is settled with one block-wide sum instead of the walk below.
      const long long E = R.axis == 1 ? E1 : E2;
      const int sgn = E < R.mid ? 1 : (E > R.mid ? -1 : 0);  // uniform over the block
      bool toward[kDecidePer];
      long long a1 = 0, a2 = 0, an = 0;
#pragma unroll
      for (int i = 0; i < kDecidePer; ++i) {
        const long long dw = R.axis == 1 ? cur.d1[i] : cur.d2[i];
        toward[i] = chg[i] && (sgn > 0 ? dw > 0 : (sgn < 0 && dw < 0));
        a1 += toward[i] ? cur.d1[i] : 0;
        a2 += toward[i] ? cur.d2[i] : 0;
        an += toward[i] ? dn[i] : 0;
      }
      a1 = wave_sum_i64(a1);
      a2 = wave_sum_i64(a2);
      an = wave_sum_i64(an);
      if (lane == 0) {
        sW1[wv] = a1;
        sW2[wv] = a2;
        sWN[wv] = (int)an;
      }
      __syncthreads();
      a1 = 0;
      a2 = 0;
      long long nt = 0;
      for (int w = 0; w < kDecideWaves; ++w) {
        a1 += sW1[w];
        a2 += sW2[w];
        nt += sWN[w];
      }
      __syncthreads();  // the wave slots are written again below
      const long long Eend = E + (R.axis == 1 ? a1 : a2);
      if (sgn != 0 && (sgn > 0 ? Eend <= R.mid : Eend >= R.mid)) {
#pragma unroll
        for (int i = 0; i < kDecidePer; ++i)
          if (chg[i] && !toward[i]) gmem(links)[cur.pos[i]] = cur.old[i];
        E1 += a1;
        E2 += a2;
        acc += (unsigned long long)nt;
        continue;
      }
    }
    // block-wide prefix sums: the lane's own inclusive ones plus the totals
    // of the lanes below it (in the wave, then of the waves below)
    const long long w1 = scan_i64(l1, lane), w2 = scan_i64(l2, lane);
    const int wn = scan_i32(ln, lane), wt = scan_i32(lt, lane);
    if (lane == 63) {
      sW1[wv] = w1;
      sW2[wv] = w2;
      sWN[wv] = wn;
      sWT[wv] = wt;
    }
    __syncthreads();
    long long b1 = w1 - l1, b2 = w2 - l2;
    int bn = wn - ln;
    {
      const int wl = lane & 15;
      const long long t1 = row_scan_i64(sW1[wl]), t2 = row_scan_i64(sW2[wl]);
      const int tn = row_scan_i32(sWN[wl]), tt = row_scan_i32(sWT[wl]);
      const uint32_t wu = (uint32_t)__builtin_amdgcn_readfirstlane(wv);
      if (wu > 0) {
        b1 += readlane64(t1, wu - 1);
        b2 += readlane64(t2, wu - 1);
        bn += __builtin_amdgcn_readlane(tn, (int)(wu - 1));
      }
      // unchanged proposals: all accepted under the window rule
      if (!R.drive) acc += (unsigned long long)(unsigned)__builtin_amdgcn_readlane(tt, 15);
    }
    // every proposal's prefix sums into LDS, and the lane's first proposal
    // whose rule fails if all from the segment's start up to it are accepted
    uint32_t fl = kDecidePer;
#pragma unroll
    for (int i = 0; i < kDecidePer; ++i) {
      b1 += cur.d1[i];
      b2 += cur.d2[i];
      bn += dn[i];
      const uint32_t j = tid * (uint32_t)kDecidePer + (uint32_t)i;
      sP1[j] = b1;
      sP2[j] = b2;
      sPN[j] = bn;
      sPos[j] = cur.pos[i];
      sOld[j] = cur.old[i];
      sChg[j] = chg[i] ? 1 : 0;
      if (fl == kDecidePer && chg[i] && !R(E1 + b1 - cur.d1[i], E2 + b2 - cur.d2[i], E1 + b1, E2 + b2))
        fl = (uint32_t)i;
    }
    // the whole segment at once: the first proposal whose rule fails (the
    // barrier inside also publishes the stores above)
    const uint32_t f0 = block_first(fl < (uint32_t)kDecidePer, tid * (uint32_t)kDecidePer + fl, n, slots, k);
    if (f0 > 0) {
      E1 += uniform64(sP1[f0 - 1]);
      E2 += uniform64(sP2[f0 - 1]);
      acc += (unsigned long long)__builtin_amdgcn_readfirstlane(sPN[f0 - 1]);
    }
    if (f0 < n) {
      // contested: one wave walks the rest by runs of acceptances and
      // refusals, 64 proposals at a time, with no block barriers
      if (wv == 0 && !R.drive) {
        if (R.axis == 1)
          walk_window<1>(R, n, f0, sP1, sP2, sPN, sChg, sPos, sOld, links, E1, E2, acc);
        else if (R.axis == 2)
          walk_window<2>(R, n, f0, sP1, sP2, sPN, sChg, sPos, sOld, links, E1, E2, acc);
        else
          walk_window<3>(R, n, f0, sP1, sP2, sPN, sChg, sPos, sOld, links, E1, E2, acc);
        if (lane == 0) {
          sE[0] = E1;
          sE[1] = E2;
          sAcc = acc;
        }
      } else if (wv == 0) {
        for (uint32_t base = f0 & ~63u; base < n; base += 64) {
          const uint32_t cn = n - base < 64 ? n - base : 64;
          const uint32_t kk = base + (uint32_t)lane;
          const bool valid = (uint32_t)lane < cn;
          // the chunk's prefix sums stay in registers; the walk reads them
          // across lanes (readlane) instead of from LDS
          const long long c1 = base ? sP1[base - 1] : 0, c2 = base ? sP2[base - 1] : 0;  // before the chunk
          const int cN = base ? sPN[base - 1] : 0;
          const long long q1 = valid ? sP1[kk] : c1, q2 = valid ? sP2[kk] : c2;
          const int qn = valid ? sPN[kk] : cN;
          const long long e1 = valid ? q1 - (kk ? sP1[kk - 1] : 0) : 0, e2 = valid ? q2 - (kk ? sP2[kk - 1] : 0) : 0;
          const bool chg = valid && sChg[kk];
          uint32_t s0 = base < f0 ? f0 - base : 0;  // uniform over the wave
          while (s0 < cn) {
            const long long B1 = s0 ? readlane64(q1, s0 - 1) : c1, B2 = s0 ? readlane64(q2, s0 - 1) : c2;
            const long long N1 = E1 + q1 - B1, N2 = E2 + q2 - B2;
            const unsigned long long bm = __ballot(chg && (uint32_t)lane >= s0 && !R(N1 - e1, N2 - e2, N1, N2));
            const uint32_t f = bm ? (uint32_t)__builtin_ctzll(bm) : cn;
            if (f > s0) {  // accept [s0, f)
              E1 += readlane64(q1, f - 1) - B1;
              E2 += readlane64(q2, f - 1) - B2;
              acc += (unsigned long long)(__builtin_amdgcn_readlane(qn, (int)(f - 1)) -
                                          (s0 ? __builtin_amdgcn_readlane(qn, (int)(s0 - 1)) : cN));
            }
            if (f >= cn) break;
            const unsigned long long gm = __ballot(chg && (uint32_t)lane > f && R(E1, E2, E1 + e1, E2 + e2));
            const uint32_t g = gm ? (uint32_t)__builtin_ctzll(gm) : cn;
            if (chg && (uint32_t)lane >= f && (uint32_t)lane < g) gmem(links)[sPos[kk]] = sOld[kk];
            s0 = g;
          }
        }
        if (lane == 0) {
          sE[0] = E1;
          sE[1] = E2;
          sAcc = acc;
        }
      }
      __syncthreads();
      E1 = uniform64(sE[0]);
      E2 = uniform64(sE[1]);
      acc = (unsigned long long)uniform64((long long)sAcc);
    }
    __syncthreads();  // the next segment overwrites the shared arrays
  }
  if (tid == 0) {
    S.S1 = E1;
    S.S2 = E2;
    S.acc += acc;
  }
}

// End of a sweep: drive-in bookkeeping (llr.cpp:223-231) and the
// Robbins-Monro step a += 12/(W^2 (m+1)) (<E> - Ewin) (llr.cpp:237-249).
// idx: the chains to end (idx[0..n)), or every chain 0..n-1 when null.
__global__ void k_llr_sweep_end(const LlrChain *chains, LlrState *st, const uint32_t *idx, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = idx ? idx[i] : i;
  const LlrChain &LC = chains[c];
  LlrState &S = st[c];
  if (!llr_active(S.phase)) return;
  S.sweep += 1;
  S.hit += (unsigned long long)LC.base.G.V * (unsigned long long)LC.base.G.D * kNHit;
  const long long Ew = LC.axis == 2 ? S.S2 : S.S1;  // the S1 window for axes 1 and 3
  if (S.phase == kLlrSeed) {
    S.seed_sweeps += 1;
    if (llr_inside(LC, S.S1, S.S2)) {
      S.phase = LC.NRM ? kLlrRM : kLlrDone;  // NRM = 0: drive-in only (llr2d.cpp:192-202)
      S.reached = 1;
      S.k = 0;
      S.m = 0;
      S.sumE = 0;
      S.sumE2 = 0;
      S.resync = 1;
    } else if (S.seed_sweeps >= LC.cap) {
      S.phase = kLlrDone;
      S.reached = 0;
    }
    return;
  }
  S.sumE += from_fix(Ew);
  if (LC.axis == 3) S.sumE2 += from_fix(S.S2);
  S.k += 1;
  if (S.k == LC.K) {
    const double a_old = S.a, a2_old = S.a2;
    const double g = S.sumE / LC.K - LC.Ewin;
    S.a += (12.0 / (LC.W * LC.W * (S.m + 1))) * g;
    S.last_da = fabs(S.a - a_old);
    if (LC.axis == 3) {  // llr2d.cpp:247-248
      const double g2 = S.sumE2 / LC.K - LC.Ewin2;
      S.a2 += (12.0 / (LC.W2 * LC.W2 * (S.m + 1))) * g2;
      S.last_da = fmax(S.last_da, fabs(S.a2 - a2_old));
      S.sumE2 = 0;
    }
    S.m += 1;
    S.k = 0;
    S.sumE = 0;
    S.resync = 1;
    if ((LC.tol > 0.0 && S.last_da < LC.tol) || S.m >= LC.NRM) S.phase = kLlrDone;
  }
}

// Recompute S1 = sum of plaquette ReTr and S2 = sum of 2x1 rectangle ReTr
// (llr.cpp:46-55,139-151, each rectangle once) for the chains that asked
// (resync 1: measure the drift first; 2: initial, no drift); idx as for
// k_llr_sweep_end, the sum's tasks covering the same chains.
__global__ void k_llr_resync_zero(LlrState *st, const uint32_t *idx, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = idx ? idx[i] : i;
  if (st[c].resync) {
    st[c].r1 = 0;
    st[c].r2 = 0;
  }
}

__global__ __launch_bounds__(kBlock) void k_llr_resync_sum(const LlrChain *chains, LlrState *st,
                                                           const ObsTask *tasks) {
  const ObsTask tk = tasks[blockIdx.x];
  LlrState &S = st[tk.chain];
  if (!S.resync) return;
  const BatchChain &BC = chains[tk.chain].base;
  const Geom G = BC.G;
  const TabView T{BC.T.mult, BC.T.inv, BC.T.P};
  const int D = G.D;
  long long p1 = 0, p2 = 0;
  for (uint32_t t = threadIdx.x; t < tk.nsites; t += blockDim.x) {
    const uint32_t i = tk.site0 + t;
    for (int d1 = 0; d1 < D; ++d1)
      for (int d2 = d1 + 1; d2 < D; ++d2) p1 += to_fix(BC.T.retr[plaq_elem(BC.links, G, T, i, d1, d2)]);
    for (int d1 = 0; d1 < D; ++d1)
      for (int d2 = 0; d2 < D; ++d2)
        if (d1 != d2) p2 += to_fix(BC.T.retr[rect_elem(BC.links, G, T, i, d1, d2)]);
  }
  p1 = wave_sum_i64(p1);
  p2 = wave_sum_i64(p2);
  if ((threadIdx.x & 63u) == 0) {
    atomicAdd(reinterpret_cast<unsigned long long *>(&S.r1), (unsigned long long)p1);
    atomicAdd(reinterpret_cast<unsigned long long *>(&S.r2), (unsigned long long)p2);
  }
}

__global__ void k_llr_resync_apply(const LlrChain *chains, LlrState *st, const uint32_t *idx, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = idx ? idx[i] : i;
  if (!st[c].resync) return;
  LlrState &S = st[c];
  if (S.resync == 1) {
    const int axis = chains[c].axis;
    const double d1 = fabs(from_fix(S.S1) - from_fix(S.r1)), d2 = fabs(from_fix(S.S2) - from_fix(S.r2));
    const double drift = axis == 1 ? d1 : axis == 2 ? d2 : fmax(d1, d2);
    if (drift > S.maxdrift) S.maxdrift = drift;
    if (d1 > S.drift1) S.drift1 = d1;
    if (d2 > S.drift2) S.drift2 = d2;
  }
  S.S1 = S.r1;
  S.S2 = S.r2;
  S.resync = 0;
}

inline dim3 lanes(uint32_t n) { return dim3((n + 63) / 64); }

}  // namespace

size_t llr_pass_static_lds() { return kNHit * kBlock * (sizeof(double) + sizeof(uint16_t)) + 64 * sizeof(int32_t); }

hipError_t launch_llr_pass(int D, const LlrChain *chains, LlrState *st, const BatchTask *tasks, uint32_t ntasks,
                           size_t lds_bytes, hipStream_t s) {
  if (!ntasks) return hipSuccess;
  const dim3 g(ntasks), b(kBlock);
  switch (D) {
    case 2: hipLaunchKernelGGL(k_llr_pass<2>, g, b, lds_bytes, s, chains, st, tasks); break;
    case 3: hipLaunchKernelGGL(k_llr_pass<3>, g, b, lds_bytes, s, chains, st, tasks); break;
    case 4: hipLaunchKernelGGL(k_llr_pass<4>, g, b, lds_bytes, s, chains, st, tasks); break;
    case 5: hipLaunchKernelGGL(k_llr_pass<5>, g, b, lds_bytes, s, chains, st, tasks); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_llr_decide(const LlrChain *chains, LlrState *st, const uint32_t *idx, uint32_t n, uint32_t pass,
                             hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_llr_decide, dim3(n), dim3(kDecide), 0, s, chains, st, idx, pass);
  return hipGetLastError();
}

hipError_t launch_llr_sweep_end(const LlrChain *chains, LlrState *st, const uint32_t *idx, uint32_t n,
                                hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_llr_sweep_end, lanes(n), dim3(64), 0, s, chains, st, idx, n);
  return hipGetLastError();
}

hipError_t launch_llr_resync(const LlrChain *chains, LlrState *st, const uint32_t *idx, uint32_t n,
                             const ObsTask *tasks, uint32_t ntasks, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_llr_resync_zero, lanes(n), dim3(64), 0, s, st, idx, n);
  if (ntasks) hipLaunchKernelGGL(k_llr_resync_sum, dim3(ntasks), dim3(kBlock), 0, s, chains, st, tasks);
  hipLaunchKernelGGL(k_llr_resync_apply, lanes(n), dim3(64), 0, s, chains, st, idx, n);
  return hipGetLastError();
}

}  // namespace dsa
