// schedule.cpp -- visiting orders for the sweep (see schedule.hpp).
#include "schedule.hpp"

#include <algorithm>

namespace dsa {

Geom make_geom(unsigned D, unsigned Nt, unsigned Nx) {
  Geom G{};
  G.D = (int)D;
  uint64_t V = 1;
  for (unsigned d = 0; d < D && d < (unsigned)kDMax; ++d) {
    G.ext[d] = d == 0 ? Nt : Nx;
    G.stride[d] = (uint32_t)V;
    V *= G.ext[d];
  }
  G.V = (uint32_t)V;
  G.eo = (Nt % 2 == 0) ? 1 : 0;
  G.halfV = G.V / 2;
  return G;
}

namespace {

void coords(const Geom &G, uint32_t i, unsigned *x) {
  for (int d = 0; d < G.D; ++d) {
    x[d] = i % G.ext[d];
    i /= G.ext[d];
  }
}

uint32_t shift(const Geom &G, uint32_t i, const unsigned *x, int d, int a) {
  const int n = (int)G.ext[d];
  int y = ((int)x[d] + a) % n;
  if (y < 0) y += n;
  return i + (uint32_t)(y - (int)x[d]) * G.stride[d];
}

}  // namespace

void conflict_sites(const Geom &G, uint32_t i, int mu, std::vector<uint32_t> &out) {
  out.clear();
  unsigned x[kDMax];
  coords(G, i, x);
  auto add = [&](uint32_t j) {
    if (j != i) out.push_back(j);
  };
  add(shift(G, i, x, mu, 1));
  add(shift(G, i, x, mu, -1));
  for (int nu = 0; nu < G.D; ++nu) {
    if (nu == mu) continue;
    for (int b : {1, -1, 2, -2}) add(shift(G, i, x, nu, b));
    for (int a : {1, -1}) {
      const uint32_t j = shift(G, i, x, mu, a);
      unsigned y[kDMax];
      coords(G, j, y);
      for (int b : {1, -1}) add(shift(G, j, y, nu, b));
    }
  }
}

std::string build_colour_schedule(const Geom &G, Schedule &S) {
  S = Schedule();
  const uint32_t V = G.V;
  std::vector<uint8_t> col(V);
  std::vector<uint32_t> nb;
  for (int mu = 0; mu < G.D; ++mu) {
    bool even = G.ext[mu] % 2 == 0, t3 = true, t4 = true;
    for (int nu = 0; nu < G.D; ++nu) {
      if (nu == mu) continue;
      if (G.ext[nu] % 3) t3 = false;
      if (G.ext[nu] % 4) t4 = false;
    }
    const unsigned m = (even && t3) ? 3 : (even && t4) ? 4 : 0;
    unsigned ncol = 0;
    if (m) {
      unsigned x[kDMax];
      for (uint32_t i = 0; i < V; ++i) {
        coords(G, i, x);
        unsigned s = 0;
        for (int nu = 0; nu < G.D; ++nu)
          if (nu != mu) s += x[nu];
        col[i] = (uint8_t)((x[mu] & 1u) + 2u * (s % m));
      }
      ncol = 2 * m;
    } else {
      for (uint32_t i = 0; i < V; ++i) {
        conflict_sites(G, i, mu, nb);
        uint64_t used = 0;
        for (uint32_t j : nb)
          if (j < i) used |= 1ull << col[j];
        unsigned c = 0;
        while ((used >> c) & 1u) ++c;
        if (c >= 64) return "colouring needs more than 64 colours";
        col[i] = (uint8_t)c;
        ncol = std::max(ncol, c + 1);
      }
    }
    // verify: no two links of one colour share a loop
    for (uint32_t i = 0; i < V; ++i) {
      conflict_sites(G, i, mu, nb);
      for (uint32_t j : nb)
        if (col[j] == col[i])
          return "internal error: colouring conflict at site " + std::to_string(i) + " direction " +
                 std::to_string(mu);
    }
    S.ncol[mu] = ncol;
    for (unsigned c = 0; c < ncol; ++c) {
      Pass p{mu, (uint32_t)S.sites.size(), 0};
      for (uint32_t i = 0; i < V; ++i)
        if (col[i] == c) S.sites.push_back(i);
      p.count = (uint32_t)S.sites.size() - p.offset;
      if (p.count) S.passes.push_back(p);
    }
  }
  return std::string();
}

std::string build_reference_schedule(const Geom &G, Schedule &S) {
  S = Schedule();
  const uint32_t V = G.V;
  std::vector<uint8_t> par(V);
  {
    unsigned x[kDMax];
    for (uint32_t i = 0; i < V; ++i) {
      coords(G, i, x);
      unsigned s = 0;
      for (int d = 0; d < G.D; ++d) s += x[d];
      par[i] = (uint8_t)(s & 1u);
    }
  }
  std::vector<uint32_t> level(V), nb;
  for (int mu = 0; mu < G.D; ++mu) {
    unsigned total = 0;
    for (int p = 0; p < 2; ++p) {
      uint32_t maxl = 0;
      for (uint32_t i = 0; i < V; ++i) {
        if (par[i] != p) continue;
        conflict_sites(G, i, mu, nb);
        uint32_t l = 0;
        for (uint32_t j : nb)
          if (j < i && par[j] == p) l = std::max(l, level[j] + 1);
        level[i] = l;
        maxl = std::max(maxl, l);
      }
      // bucket the sites of this parity by level (ascending i within a level)
      std::vector<uint32_t> cnt(maxl + 2, 0);
      uint32_t nsites = 0;
      for (uint32_t i = 0; i < V; ++i)
        if (par[i] == p) {
          cnt[level[i] + 1]++;
          nsites++;
        }
      if (!nsites) continue;
      for (uint32_t l = 0; l <= maxl; ++l) cnt[l + 1] += cnt[l];
      const uint32_t base = (uint32_t)S.sites.size();
      S.sites.resize(base + nsites);
      std::vector<uint32_t> fillp(cnt.begin(), cnt.end() - 1);
      for (uint32_t i = 0; i < V; ++i)
        if (par[i] == p) S.sites[base + fillp[level[i]]++] = i;
      for (uint32_t l = 0; l <= maxl; ++l)
        if (cnt[l + 1] > cnt[l]) S.passes.push_back(Pass{mu, base + cnt[l], cnt[l + 1] - cnt[l]});
      total += maxl + 1;
    }
    S.ncol[mu] = total;
  }
  return std::string();
}

}  // namespace dsa
