// group.cpp -- table loader (group.c:13-81 semantics, with errors returned
// instead of abort()).
#include "group.hpp"

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
  if (Pl <= 0 || Pl > (long)kMaxOrder) return "group file: order " + std::to_string(Pl) + " outside [1, 65535]";
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
  return std::string();
}

}  // namespace dsa
