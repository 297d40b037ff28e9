// group.cpp -- table loader (group.c:13-81 semantics, with errors returned
// instead of abort()).
#include "group.hpp"

#include <algorithm>
#include <cmath>

#include <stdlib.h>
#include <string.h>
#include <zlib.h>

namespace dsa {

namespace {

// Reads the whole file; gzopen passes plain files through unchanged.
bool read_all(const char *path, std::string &out) {
  gzFile f = gzopen(path, "rb");
  if (!f) return false;
  char buf[1 << 16];
  for (;;) {
    int r = gzread(f, buf, sizeof buf);
    if (r < 0) {
      gzclose(f);
      return false;
    }
    if (r == 0) break;
    out.append(buf, (size_t)r);
  }
  gzclose(f);
  return true;
}

}  // namespace

std::string load_group(const char *path, Group &g) {
  g = Group();
  g.path = path ? path : "";
  std::string text;
  if (!path || !read_all(path, text)) return std::string("cannot open group file ") + (path ? path : "(null)");
  const char *p = text.c_str();
  char *e = nullptr;
  long Pl = strtol(p, &e, 10);
  if (e == p) return "group file: missing order";
  if (Pl <= 0 || Pl > (long)kMaxOrder) return "group file: order " + std::to_string(Pl) + " outside [1, " + std::to_string(kMaxOrder) + "]";
  p = e;
  const unsigned P = (unsigned)Pl;
  g.P = P;
  g.retr.resize(P);
  g.imtr.resize(P);
  for (unsigned n = 0; n < P; ++n) {
    g.retr[n] = strtod(p, &e);
    if (e == p) return "group file: truncated in the ReTr row at element " + std::to_string(n);
    p = e;
  }
  for (unsigned n = 0; n < P; ++n) {
    g.imtr[n] = strtod(p, &e);
    if (e == p) return "group file: truncated in the ImTr row at element " + std::to_string(n);
    p = e;
  }
  g.mult.resize((size_t)P * P);
  for (size_t n = 0; n < (size_t)P * P; ++n) {
    long v = strtol(p, &e, 10);
    if (e == p) return "group file: truncated multiplication table at entry " + std::to_string(n);
    if (v < 0 || v >= (long)P) return "group file: table entry " + std::to_string(v) + " out of range";
    g.mult[n] = (uint16_t)v;
    p = e;
  }
  // optional: the defining-representation matrices appended after the table,
  // 18 reals per element (dym-mod-metro-savecfg.cpp:249-262)
  {
    std::vector<double> m((size_t)P * 18);
    size_t n = 0;
    for (; n < m.size(); ++n) {
      m[n] = strtod(p, &e);
      if (e == p) break;
      p = e;
    }
    if (n == m.size()) g.su3.swap(m);
  }
  text.clear();
  text.shrink_to_fit();
  // identity: first n with identity row and column
  for (unsigned n = 0; n < P && g.id < 0; ++n) {
    bool ok = true;
    for (unsigned m = 0; m < P && ok; ++m)
      ok = g.mult[(size_t)n * P + m] == m && g.mult[(size_t)m * P + n] == m;
    if (ok) g.id = (int)n;
  }
  if (g.id < 0) return "Group does not have identity element";
  g.inv.resize(P);
  for (unsigned n = 0; n < P; ++n) {
    bool found = false;
    for (unsigned m = 0; m < P; ++m)
      if (g.mult[(size_t)n * P + m] == g.id && g.mult[(size_t)m * P + n] == g.id) {
        g.inv[n] = (uint16_t)m;
        found = true;
        break;
      }
    if (!found) return "Group does not have inverses";
  }
  double min_retr = 0, nn_retr = 10;
  for (unsigned i = 0; i < P; ++i)
    if (g.retr[i] < min_retr) min_retr = g.retr[i];
  for (unsigned i = 0; i < P; ++i)
    if (g.retr[i] > min_retr && g.retr[i] < nn_retr) nn_retr = g.retr[i];
  for (unsigned i = 0; i < P; ++i)
    if (g.retr[i] < nn_retr + 1e-6 && g.retr[i] > min_retr) g.pool.push_back((uint16_t)i);
  g.min_retr = min_retr;
  g.nn_retr = nn_retr;
  if (g.pool.empty()) return "group has no proposal elements (all ReTr equal)";
  build_classes(g);
  build_field(g);
  return std::string();
}

void build_classes(Group &g) {
  g.cls.clear();
  g.cprod.clear();
  g.rc.clear();
  g.rcq.clear();
  g.mdq.clear();
  g.cval.clear();
  g.cval0.clear();
  g.dmax = g.rmax = 0;
  const size_t P = g.P;
  if (P == 0 || P > kMaxClassOrder) return;
  std::vector<int> cls(P, -1);
  int nc = 0;
  for (size_t x = 0; x < P; ++x) {
    if (cls[x] >= 0) continue;
    if (nc == (int)kMaxClasses) {
      g.rc.clear();
      return;
    }
    // the class of x: { y x y^-1 }
    for (size_t y = 0; y < P; ++y) {
      const size_t yx = g.mult[y * P + x];
      cls[g.mult[yx * P + g.inv[y]]] = nc;
    }
    g.rc.push_back(g.retr[x]);
    ++nc;
  }
  g.cls.resize(P);
  for (size_t e = 0; e < P; ++e) {
    g.cls[e] = (uint8_t)(4 * cls[e]);
    g.dmax = std::max(g.dmax, std::fabs(g.retr[e] - g.rc[cls[e]]));
    g.rmax = std::max(g.rmax, std::fabs(g.retr[e]));
  }
  if (g.rmax > 4.0) {  // class sums must fit int32 in units of 2^-24
    g.cls.clear();
    g.rc.clear();
    g.dmax = g.rmax = 0;
    return;
  }
  for (double v : g.rc) g.rcq.push_back((int32_t)std::llround(std::ldexp(v, 24)));
  g.cprod.resize(g.pool.size() * P);
  for (size_t j = 0; j < g.pool.size(); ++j)
    for (size_t t = 0; t < P; ++t) g.cprod[j * P + t] = g.cls[g.mult[(size_t)g.pool[j] * P + t]];
  // per-class bound on one hit's change of a staple term: a trial term is
  // ReTr(b p r) = ReTr(p' (b r)) with p' = b p b^-1, which is in the pool
  // when the pool is closed under conjugation
  std::vector<char> in_pool(P, 0);
  for (uint16_t p : g.pool) in_pool[p] = 1;
  bool closed = true;
  for (size_t j = 0; j < g.pool.size() && closed; ++j)
    for (size_t y = 0; y < P && closed; ++y)
      closed = in_pool[g.mult[g.mult[y * P + g.pool[j]] * P + g.inv[y]]] != 0;
  if (closed) {
    std::vector<double> md(nc, 0.0);
    for (size_t t = 0; t < P; ++t)
      for (uint16_t p : g.pool) {
        const double d = std::fabs(g.retr[g.mult[(size_t)p * P + t]] - g.retr[t]);
        md[cls[t]] = std::max(md[cls[t]], d);
      }
    for (double v : md) g.mdq.push_back((int32_t)std::ceil(std::ldexp(v, 24)) + 1);
  }
  if (g.rmax < 3.999) {
    std::vector<int16_t> v16;
    for (double v : g.rc) v16.push_back((int16_t)std::lround(std::ldexp(v, 13)));
    g.cval0.resize(P);
    for (size_t t = 0; t < P; ++t) g.cval0[t] = v16[g.cls[t] / 4];
    g.cval.resize(g.cprod.size());
    for (size_t q = 0; q < g.cprod.size(); ++q) g.cval[q] = v16[g.cprod[q] / 4];
  }
}

}  // namespace dsa
