// schedule.hpp -- which links each sweep pass updates.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "kernels.hpp"

namespace dsa {

struct Pass {
  int mu;
  uint32_t offset, count;  // into Schedule::sites
};

struct Schedule {
  std::vector<uint32_t> sites;  // concatenated site lists, ascending within a pass
  std::vector<Pass> passes;     // executed in order, one kernel launch each
  unsigned ncol[kDMax] = {0};   // colours (or levels) per direction
};

Geom make_geom(unsigned D, unsigned Nt, unsigned Nx);

// Sites j != i whose mu-link shares a plaquette or a 1x2/2x1 rectangle with
// link (i, mu): x +- e_mu and, per nu != mu, x +- e_nu, x +- 2 e_nu,
// x +- e_mu +- e_nu (the mu-links the staples of dym-mod-metro.cpp:63-142 read).
void conflict_sites(const Geom &G, uint32_t i, int mu, std::vector<uint32_t> &out);

// Race-free colouring: colour(x) = (x_mu mod 2) + 2*((sum_{nu != mu} x_nu) mod m),
// m = 3 if every transverse extent is a multiple of 3 and the mu extent is
// even, else m = 4 if every transverse extent is a multiple of 4, else greedy
// in ascending site order.  Passes: mu ascending, colour ascending.  Verified
// conflict-free link by link; returns an error message otherwise.
std::string build_colour_schedule(const Geom &G, Schedule &S);

// The reference's one-thread order (mu, then site parity, then ascending i,
// dym-mod-metro.cpp:152-159) split into levels: level(i) = 1 + max level of
// the conflicting same-parity sites before i.  Links of one level share no
// loop, so running a level in parallel gives exactly the sequential result.
std::string build_reference_schedule(const Geom &G, Schedule &S);

}  // namespace dsa
