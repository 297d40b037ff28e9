// field.cpp -- a faithful representation of the group by 3x3 matrices over
// GF(4) (a subgroup of SL(3,4)), found when the table is loaded, so that the
// sweep kernels can form group products with bit operations instead of
// gathers from the P x P table.
//
// S(1080) is the triple cover 3.A6 of A6, which SL(3,4) contains; the search
// below is generic (any group of order dividing |SL(3,4)| = 60480 that is
// generated by two of its elements) and the result is accepted only after
// every product of the table has been checked against it.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "gf4.hpp"
#include "group.hpp"

namespace dsa {

namespace {

// Plain matrix codes: entry (r, c) in bits 2(3r + c), e = 2 e1 + e0.
uint32_t gmul(uint32_t a, uint32_t b) {
  const uint32_t a1 = a >> 1, a0 = a & 1u, b1 = b >> 1, b0 = b & 1u;
  const uint32_t c1 = (a1 & b1) ^ (a1 & b0) ^ (a0 & b1), c0 = (a1 & b1) ^ (a0 & b0);
  return 2 * c1 + c0;
}

uint32_t ent(uint32_t m, int r, int c) { return (m >> (2 * (3 * r + c))) & 3u; }

uint32_t mmul(uint32_t a, uint32_t b) {
  uint32_t out = 0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      uint32_t s = 0;
      for (int k = 0; k < 3; ++k) s ^= gmul(ent(a, r, k), ent(b, k, c));
      out |= s << (2 * (3 * r + c));
    }
  return out;
}

uint32_t mdet(uint32_t m) {
  auto e = [&](int r, int c) { return ent(m, r, c); };
  return gmul(e(0, 0), gmul(e(1, 1), e(2, 2)) ^ gmul(e(1, 2), e(2, 1))) ^
         gmul(e(0, 1), gmul(e(1, 0), e(2, 2)) ^ gmul(e(1, 2), e(2, 0))) ^
         gmul(e(0, 2), gmul(e(1, 0), e(2, 1)) ^ gmul(e(1, 1), e(2, 0)));
}

constexpr uint32_t kIdentity = 1u | 1u << 8 | 1u << 16;

// The inverse of an invertible matrix: the adjugate divided by the
// determinant (characteristic 2: cofactor signs are all +).
uint32_t minv(uint32_t m) {
  const uint32_t d = mdet(m);
  uint32_t dinv = 0;
  for (uint32_t x = 1; x < 4; ++x)
    if (gmul(d, x) == 1) dinv = x;
  uint32_t out = 0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      // inverse (r, c) = cofactor (c, r) / det
      const int r0 = c == 0 ? 1 : 0, r1 = c == 2 ? 1 : 2, c0 = r == 0 ? 1 : 0, c1 = r == 2 ? 1 : 2;
      const uint32_t cof = gmul(ent(m, r0, c0), ent(m, r1, c1)) ^ gmul(ent(m, r0, c1), ent(m, r1, c0));
      out |= gmul(cof, dinv) << (2 * (3 * r + c));
    }
  return out;
}

int morder(uint32_t m) {
  uint32_t x = m;
  for (int n = 1; n <= 64; ++n) {
    if (x == kIdentity) return n;
    x = mmul(x, m);
  }
  return 0;
}

// A row vector (3 entries, 2 bits each) times a matrix.
uint32_t vmul(uint32_t v, uint32_t m) {
  uint32_t out = 0;
  for (int c = 0; c < 3; ++c) {
    uint32_t s = 0;
    for (int k = 0; k < 3; ++k) s ^= gmul((v >> (2 * k)) & 3u, ent(m, k, c));
    out |= s << (2 * c);
  }
  return out;
}

struct SL34 {
  std::vector<uint32_t> el;   // every matrix of determinant 1
  std::vector<uint8_t> ord;   // its order
  std::vector<uint32_t> inv;  // its inverse
};

const SL34 &sl34() {
  static SL34 s;
  static std::once_flag once;
  std::call_once(once, [] {
    for (uint32_t m = 0; m < (1u << 18); ++m)
      if (mdet(m) == 1) s.el.push_back(m);
    for (uint32_t m : s.el) {
      s.ord.push_back((uint8_t)morder(m));
      s.inv.push_back(minv(m));
    }
  });
  return s;
}

int gorder(const Group &g, uint32_t e) {
  uint32_t x = e;
  for (int n = 1; n <= (int)g.P; ++n) {
    if ((int)x == g.id) return n;
    x = g.mult[(size_t)x * g.P + e];
  }
  return 0;
}

size_t closure(const Group &g, uint32_t a, uint32_t b) {
  std::vector<char> seen(g.P, 0);
  std::vector<uint32_t> q{(uint32_t)g.id};
  seen[g.id] = 1;
  for (size_t h = 0; h < q.size(); ++h)
    for (uint32_t s : {a, b}) {
      const uint32_t y = g.mult[(size_t)q[h] * g.P + s];
      if (!seen[y]) {
        seen[y] = 1;
        q.push_back(y);
      }
    }
  return q.size();
}

// The homomorphism a -> A, b -> B extended over the group (breadth first
// from the identity), or false when some relation of the table fails.
bool extend(const Group &g, uint32_t a, uint32_t b, uint32_t A, uint32_t B, std::vector<uint32_t> &img) {
  img.assign(g.P, 0xFFFFFFFFu);
  img[g.id] = kIdentity;
  std::vector<uint32_t> q{(uint32_t)g.id};
  for (size_t h = 0; h < q.size(); ++h) {
    const uint32_t x = q[h];
    for (int k = 0; k < 2; ++k) {
      const uint32_t y = g.mult[(size_t)x * g.P + (k ? b : a)];
      const uint32_t Y = mmul(img[x], k ? B : A);
      if (img[y] == 0xFFFFFFFFu) {
        img[y] = Y;
        q.push_back(y);
      } else if (img[y] != Y) {
        return false;
      }
    }
  }
  return q.size() == g.P;
}

// Representatives of the classes of `cand` under conjugation by GL(3,4)
// (every g in GL(3,4) is s diag(d, 1, 1) with s in SL(3,4), d = det g).
std::vector<uint32_t> gl_class_reps(const SL34 &S, const std::vector<uint32_t> &cand) {
  std::vector<uint32_t> reps;
  std::vector<char> done(1u << 18, 0);
  for (uint32_t A : cand) {
    if (done[A]) continue;
    reps.push_back(A);
    for (uint32_t d = 1; d < 4; ++d) {
      uint32_t Dm = kIdentity & ~3u, Di = Dm;
      uint32_t dinv = 0;
      for (uint32_t x = 1; x < 4; ++x)
        if (gmul(d, x) == 1) dinv = x;
      Dm |= d;
      Di |= dinv;
      const uint32_t A1 = mmul(mmul(Dm, A), Di);
      for (size_t i = 0; i < S.el.size(); ++i) done[mmul(mmul(S.el[i], A1), S.inv[i])] = 1;
    }
  }
  return reps;
}

}  // namespace

void build_field(Group &g) {
  g.fmat.clear();
  g.fkey.clear();
  const unsigned P = g.P;
  if (P < 2 || 60480 % P != 0 || g.id < 0) return;
  const SL34 &S = sl34();
  std::vector<int> ord(P);
  for (unsigned e = 0; e < P; ++e) ord[e] = gorder(g, e);
  // two generators: the first pair (in index order, a of the largest order
  // first) whose closure is the group
  std::vector<uint32_t> by_ord(P);
  for (unsigned e = 0; e < P; ++e) by_ord[e] = e;
  std::stable_sort(by_ord.begin(), by_ord.end(), [&](uint32_t x, uint32_t y) { return ord[x] > ord[y]; });
  uint32_t a = 0, b = 0;
  bool gen = false;
  int tries = 0;
  for (size_t i = 0; i < P && !gen && tries < 4000; ++i)
    for (size_t j = 0; j < P && !gen && tries < 4000; ++j) {
      if (by_ord[i] == by_ord[j]) continue;
      ++tries;
      if (closure(g, by_ord[i], by_ord[j]) == P) {
        a = by_ord[i];
        b = by_ord[j];
        gen = true;
      }
    }
  if (!gen) return;
  const uint32_t ab = g.mult[(size_t)a * P + b], abb = g.mult[(size_t)ab * P + b];
  std::vector<uint32_t> ca, cb;
  for (size_t i = 0; i < S.el.size(); ++i) {
    if (S.ord[i] == ord[a]) ca.push_back(S.el[i]);
    if (S.ord[i] == ord[b]) cb.push_back(S.el[i]);
  }
  std::vector<uint32_t> img;
  bool found = false;
  for (uint32_t A : gl_class_reps(S, ca)) {
    for (uint32_t B : cb) {
      const uint32_t AB = mmul(A, B);
      if (morder(AB) != ord[ab] || morder(mmul(AB, B)) != ord[abb]) continue;
      if (!extend(g, a, b, A, B, img)) continue;
      std::vector<uint32_t> sorted(img);
      std::sort(sorted.begin(), sorted.end());
      if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) continue;  // not faithful
      found = true;
      break;
    }
    if (found) break;
  }
  if (!found) return;
  // change of basis T (rows v0, v1, v2) such that the first two rows of
  // T M T^-1 tell the elements apart: rows v0 M, v1 M of M, times T^-1
  std::vector<uint16_t> seen(kGf4Keys);
  uint32_t T = 0;
  bool basis = false;
  for (uint32_t v0 = 1; v0 < 64 && !basis; ++v0)
    for (uint32_t v1 = 1; v1 < 64 && !basis; ++v1) {
      uint32_t Tm = v0 | v1 << 6;
      uint32_t v2 = 1;
      for (; v2 < 64; ++v2)
        if (mdet(Tm | v2 << 12) != 0) break;
      if (v2 == 64) continue;  // v0, v1 dependent
      std::fill(seen.begin(), seen.end(), 0);
      bool inj = true;
      for (unsigned e = 0; e < P && inj; ++e) {
        const uint32_t k = vmul(v0, img[e]) | vmul(v1, img[e]) << 6;
        inj = !seen[k];
        seen[k] = 1;
      }
      if (inj) {
        T = Tm | v2 << 12;
        basis = true;
      }
    }
  if (!basis) return;
  const uint32_t Ti = minv(T);
  std::vector<uint32_t> fm(P);
  for (unsigned e = 0; e < P; ++e) fm[e] = gf4_pack(mmul(mmul(T, img[e]), Ti));
  std::vector<uint16_t> key(kGf4Keys, 0xFFFF);
  for (unsigned e = 0; e < P; ++e) {
    const uint32_t k = gf4_key(fm[e]);
    if (key[k] != 0xFFFF) return;
    key[k] = (uint16_t)e;
  }
  // every product of the table, through the packed product and the key
  for (size_t x = 0; x < P; ++x)
    for (size_t y = 0; y < P; ++y)
      if (key[gf4_key(gf4_mul3(fm[x], fm[y]))] != g.mult[x * P + y]) return;
  g.fmat.swap(fm);
  g.fkey.swap(key);
}

}  // namespace dsa
