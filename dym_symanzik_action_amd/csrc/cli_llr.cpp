// cli_llr.cpp -- drop-ins for the reference's LLR drivers, on the device
// LLR engine (dsa_llr_*).  Built twice: as `llr` and as `llr2d`.
//
//   llr   group D Nt Nx axis bg Etop Ebot step hw a0 seed [K] [NRM] [R]
//         (llr.cpp:276-345: R repeats of a hot ladder walking down and a cold
//         ladder walking up the windows, each window seeded from the previous
//         window's configuration and tilt; ANE lines)
//   llr --replay group D Nt Nx ... (the same arguments): the reference's own
//         serial chain replayed on the device (dsa_llr_ref_*), one window
//         after another: every line equals the reference's at one thread
//   llr2d group D Nt Nx E1top E1bot step1 hw1 c0 c1 c2 step2 hw2 nperp a0 seed [K] [NRM] [R]
//   llr2d seed  group D Nt Nx E1top E1bot step1 hw1 c0 c1 c2 step2 hw2 nperp a0 seed cfgdir
//   llr2d solve group D Nt Nx manifest lo hi a0 K NRM seed
//         (llr2d.cpp:283-512: the cell grid around E2 = c0 + c1 E1 + c2 E1^2,
//         cold then hot ridge walks with transverse fans; the seed pass's
//         CELL manifest and raw .cfg files; the solve pass's ANE2 lines)
//
// Arguments, defaults, usage (stderr, exit 1) and stdout lines follow the
// reference.  What differs is how the work runs: every window or cell that
// does not wait for another one (the ladders of different repeats and
// directions; the transverse cells of one ridge cell; every cell of a solve
// pass) is a chain of one engine, advanced together on the device, and the
// constrained update is the device algorithm (Philox streams, colour passes
// with ordered window acceptance, fixed-point energies).  With DSA_TRACE=1 in
// the environment every finished window or cell prints one TASK line on
// stderr: its label, Philox key, start, whether it was seeded from a
// configuration, whether it reached its window, its tilts and sweep count.
#include <libgen.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "../../include/dsa.h"

namespace {

int die(const char *prog, const char *what) {
  fprintf(stderr, "%s: %s: %s\n", prog, what, dsa_last_error());
  return 2;
}

// Philox key of a task: distinct for every (repeat, walk, window, cell).
int32_t task_seed(int32_t iseed, uint32_t a, uint32_t b, uint32_t c) {
  uint32_t h = (uint32_t)iseed * 0x9E3779B1u;
  for (uint32_t v : {a, b, c}) {
    h ^= v + 0x7F4A7C15u + (h << 6) + (h >> 2);
    h *= 0x85EBCA6Bu;
  }
  return (int32_t)(h & 0x7fffffffu);
}

struct Result {
  bool reached = false;
  double a = 0, a2 = 0, maxdrift = 0;
  std::vector<int32_t> links;  // the final configuration
};

struct Task {
  std::string label;  // names the task in the DSA_TRACE lines
  dsa_llr_params p{};
  std::vector<int32_t> init;  // empty: start as p.start says
  std::function<void(const Result &)> done;
};

// Runs tasks on one engine: free slots take queued tasks, all active chains
// are stepped together, finished chains hand their result to the task's
// continuation (which may queue further tasks).
class Runner {
 public:
  Runner(const char *prog, dsa_group *g, uint32_t D, uint32_t slots) : prog_(prog), g_(g), nslots_(slots) {
    ok_ = dsa_llr_create(D, slots, 0, &L_) == 0;
    trace_ = getenv("DSA_TRACE") && *getenv("DSA_TRACE") && strcmp(getenv("DSA_TRACE"), "0");
  }
  ~Runner() {
    if (L_) dsa_llr_destroy(L_);
  }
  bool ok() const { return ok_; }
  void submit(Task t) { queue_.push_back(std::move(t)); }
  double maxdrift() const { return maxdrift_; }
  double drift1() const { return drift1_; }
  double drift2() const { return drift2_; }

  int run() {
    std::vector<Task *> slot_task(nslots_, nullptr);
    std::vector<Task> live(nslots_);
    uint32_t active = 0;
    while (active || !queue_.empty()) {
      for (uint32_t s = 0; s < nslots_ && !queue_.empty(); ++s) {
        if (slot_task[s]) continue;
        Task t = std::move(queue_.front());
        queue_.erase(queue_.begin());
        uint32_t got = 0;
        if (dsa_llr_add(L_, g_, &t.p, t.init.empty() ? nullptr : t.init.data(), &got)) return die(prog_, "dsa_llr_add");
        live[got] = std::move(t);
        slot_task[got] = &live[got];
        ++active;
      }
      if (dsa_llr_step(L_, 5)) return die(prog_, "dsa_llr_step");
      for (uint32_t s = 0; s < nslots_; ++s) {
        if (!slot_task[s]) continue;
        dsa_llr_status st;
        if (dsa_llr_status_get(L_, s, &st)) return die(prog_, "dsa_llr_status_get");
        if (st.phase != DSA_LLR_DONE) continue;
        Result r;
        r.reached = st.reached != 0;
        r.a = st.a;
        r.a2 = st.a2;
        r.maxdrift = st.maxdrift;
        maxdrift_ = std::max(maxdrift_, st.maxdrift);
        drift1_ = std::max(drift1_, st.drift1);
        drift2_ = std::max(drift2_, st.drift2);
        const dsa_llr_params &p = slot_task[s]->p;
        uint64_t V = p.Nt;
        for (uint32_t d = 1; d < p.D; ++d) V *= p.Nx;
        r.links.resize(V * p.D);
        if (dsa_llr_get_links(L_, s, r.links.data())) return die(prog_, "dsa_llr_get_links");
        if (dsa_llr_remove(L_, s)) return die(prog_, "dsa_llr_remove");
        if (trace_)
          fprintf(stderr, "TASK: %s seed=%d start=%d init=%d reached=%d a=%.17g a2=%.17g sweeps=%llu\n",
                  slot_task[s]->label.c_str(), (int)p.seed, (int)p.start, slot_task[s]->init.empty() ? 0 : 1,
                  (int)r.reached, r.a, r.a2, (unsigned long long)st.sweeps);
        auto done = std::move(slot_task[s]->done);
        slot_task[s] = nullptr;
        --active;
        if (done) done(r);
      }
    }
    return 0;
  }

 private:
  const char *prog_;
  dsa_group *g_;
  dsa_llr *L_ = nullptr;
  uint32_t nslots_;
  bool ok_ = false;
  bool trace_ = false;
  std::vector<Task> queue_;
  double maxdrift_ = 0, drift1_ = 0, drift2_ = 0;
};

// The lines the reference's lattice set-up prints (lattice.c:57 and
// calibrate_rect_mult, llr.cpp:155-171 / llr2d.cpp:155-170): the rectangle
// multiplicity is calibrated as the reference does it, on the reference's
// own hot start for this seed (the device replay of its set-up), so the
// value printed is the reference's.  The engine itself counts every 1x2
// rectangle once, which is that value wherever no rectangle wraps onto
// itself (extents >= 3).  Returns the calibrated value, or -1 after an error.
double print_lattice_setup(const char *prog, dsa_group *g, uint32_t D, uint32_t Nt, uint32_t Nx, int32_t iseed) {
  printf("Initializing nearest neighbours [%d:%d]... \n", (int)Nt, (int)Nx);
  dsa_llr_ref *R = nullptr;
  double rm = 0;
  if (dsa_llr_ref_create(g, D, Nt, Nx, 1, 0.0, iseed, &R) || dsa_llr_ref_rect_mult(R, &rm)) {
    dsa_llr_ref_destroy(R);
    die(prog, "RECT_MULT calibration");
    return -1;
  }
  dsa_llr_ref_destroy(R);
  printf("RECT_MULT calibrated: %.6f\n", rm);
  return rm;
}

// llr2d's setup_lattice output (llr2d.cpp:274-290).
bool lattice_setup_2d(const char *prog, dsa_group *g, uint32_t D, uint32_t Nt, uint32_t Nx, int32_t iseed) {
  const double rm = print_lattice_setup(prog, g, D, Nt, Nx, iseed);
  if (rm < 0) return false;
  if (fabs(rm - 6.0) > 1e-3) printf("WARNING: RECT_MULT=%.3f != 6 -- rectangles wrap; need lattice extent >= 3\n", rm);
  fflush(stdout);
  return true;
}

dsa_llr_params base_params(uint32_t D, uint32_t Nt, uint32_t Nx) {
  dsa_llr_params p{};
  p.D = D;
  p.Nt = Nt;
  p.Nx = Nx;
  p.start = DSA_START_HOT;
  return p;
}

// ---------------------------------------------------------------- llr (1-D)

int main_llr(int argc, char **argv) {
  if (argc < 13) {
    fprintf(stderr,
            "usage: %s group D Nt Nx axis bg Etop Ebot step hw a0 seed [K] [NRM] [R]\n"
            "  axis=1: window S1 (beta1-driven), bg=beta2 ; axis=2: window S2 (beta2-driven), bg=beta1\n",
            argv[0]);
    return 1;
  }
  const char *gf = argv[1];
  const uint32_t D = (uint32_t)atoi(argv[2]), Nt = (uint32_t)atoi(argv[3]), Nx = (uint32_t)atoi(argv[4]);
  const int axis = atoi(argv[5]);
  const double bg = atof(argv[6]), Etop = atof(argv[7]), Ebot = atof(argv[8]), stepE = atof(argv[9]);
  const double hw = atof(argv[10]), a0 = atof(argv[11]);
  const int iseed = atoi(argv[12]);
  const unsigned K = argc > 13 ? (unsigned)atoi(argv[13]) : 50;
  const unsigned NRM = argc > 14 ? (unsigned)atoi(argv[14]) : 80;
  const unsigned R = argc > 15 ? (unsigned)atoi(argv[15]) : 1;
  const int M = (int)((Etop - Ebot) / stepE + 0.5) + 1;
  printf("LLR(grp,D,Nt,Nx,axis,bg): %s %d %d %d %d %g | Etop=%g Ebot=%g step=%g hw=%g a0=%g seed=%d K=%u NRM=%u R=%u "
         "| %d windows\n",
         gf, (int)D, (int)Nt, (int)Nx, axis, bg, Etop, Ebot, stepE, hw, a0, iseed, K, NRM, R, M);
  uint32_t V = Nt;
  for (uint32_t d = 1; d < D; ++d) V *= Nx;
  printf("NPLAQ: %u\n", V * D * (D - 1) / 2);
  printf("NRECT: %u\n", V * D * (D - 1));
  printf("AXIS: %d\n", axis);
  fflush(stdout);
  dsa_group *g = nullptr;
  if (dsa_group_load(gf, &g)) return die(argv[0], "dsa_group_load");
  const double rect_mult = print_lattice_setup(argv[0], g, D, Nt, Nx, iseed);
  if (rect_mult < 0) return 2;
  if (axis == 2 && fabs(rect_mult - 6.0) > 1e-3)
    printf("WARNING: RECT_MULT=%.3f != 6 -- the 1x2 rectangles wrap on this lattice\n"
           "  (extent too small in some direction); the S2 action is degenerate and the\n"
           "  S2 window is invalid. Use a lattice with extent >= 3 for axis=2.\n",
           rect_mult);
  fflush(stdout);
  std::vector<std::vector<double>> an(M);
  std::vector<int> reached(2 * R, 0);
  double maxdrift = 0;
  {
    Runner run(argv[0], g, D, std::max<uint32_t>(2 * R, 1));
    if (!run.ok()) return die(argv[0], "dsa_llr_create");
    // one ladder per (repeat, direction): hot walking down, cold walking up
    std::function<void(unsigned, int, int, double, std::vector<int32_t>)> window;
    window = [&](unsigned r, int dir, int s, double a_prev, std::vector<int32_t> init) {
      const int n = dir > 0 ? s : M - 1 - s;
      Task t;
      t.label = std::to_string(r) + (dir > 0 ? " down " : " up ") + std::to_string(n);
      t.p = base_params(D, Nt, Nx);
      t.p.axis = axis;
      t.p.bg = bg;
      t.p.Ewin = Etop - n * stepE;
      t.p.hw = hw;
      t.p.a0 = a_prev;
      t.p.seed = task_seed(iseed, r, dir > 0 ? 0 : 1, (uint32_t)n);
      t.p.start = dir > 0 ? DSA_START_HOT : DSA_START_COLD;
      t.p.K = K;
      t.p.NRM = NRM;
      t.p.cap = 20000;  // llr.cpp:225
      t.init = std::move(init);
      t.done = [&, r, dir, s, n](const Result &res) {
        if (!res.reached) return;  // stalled: the rest of this ladder is unreached (llr.cpp:267)
        an[n].push_back(res.a);
        ++reached[2 * r + (dir > 0 ? 0 : 1)];
        if (s + 1 < M) window(r, dir, s + 1, res.a, res.links);
      };
      run.submit(std::move(t));
    };
    for (unsigned r = 0; r < R; ++r) {
      window(r, +1, 0, a0, {});
      window(r, -1, 0, a0, {});
    }
    if (int rc = run.run()) return rc;
    maxdrift = run.maxdrift();
  }
  for (unsigned r = 0; r < R; ++r)
    printf("repeat %u: down reached %d/%d, up reached %d/%d\n", r, reached[2 * r], M, reached[2 * r + 1], M);
  printf("\n# Ewin   a_n        sd        nrep\n");
  for (int n = 0; n < M; ++n) {
    const double Ewin = Etop - n * stepE;
    if (an[n].empty()) {
      printf("ANE: %8.2f   --(unreached)\n", Ewin);
      continue;
    }
    double mean = 0;
    for (double x : an[n]) mean += x;
    mean /= (double)an[n].size();
    double var = 0;
    for (double x : an[n]) var += (x - mean) * (x - mean);
    const double sd = an[n].size() > 1 ? sqrt(var / (double)(an[n].size() - 1)) : 0.0;
    printf("ANE: %8.2f  %9.5f  %9.5f  %lu\n", Ewin, mean, sd, (unsigned long)an[n].size());
  }
  printf("max |incr - recompute| E(axis%d) : %.3e\n", axis, maxdrift);
  dsa_group_free(g);
  return 0;
}

// llr --replay: the reference's own chain (llr.cpp:276-345 at one thread:
// minstd streams, serial visiting order, per-hit window test, RECT_MULT as
// calibrated), replayed on the device one window after another, so every line
// equals the reference's.  A parity tool for small lattices.
int main_llr_replay(int argc, char **argv) {
  if (argc < 13) {
    fprintf(stderr,
            "usage: %s --replay group D Nt Nx axis bg Etop Ebot step hw a0 seed [K] [NRM] [R]\n"
            "  axis=1: window S1 (beta1-driven), bg=beta2 ; axis=2: window S2 (beta2-driven), bg=beta1\n",
            argv[0]);
    return 1;
  }
  const char *gf = argv[1];
  const uint32_t D = (uint32_t)atoi(argv[2]), Nt = (uint32_t)atoi(argv[3]), Nx = (uint32_t)atoi(argv[4]);
  const int axis = atoi(argv[5]);
  const double bg = atof(argv[6]), Etop = atof(argv[7]), Ebot = atof(argv[8]), stepE = atof(argv[9]);
  const double hw = atof(argv[10]), a0 = atof(argv[11]);
  const int iseed = atoi(argv[12]);
  const unsigned K = argc > 13 ? (unsigned)atoi(argv[13]) : 50;
  const unsigned NRM = argc > 14 ? (unsigned)atoi(argv[14]) : 80;
  const unsigned R = argc > 15 ? (unsigned)atoi(argv[15]) : 1;
  const int M = (int)((Etop - Ebot) / stepE + 0.5) + 1;
  printf("LLR(grp,D,Nt,Nx,axis,bg): %s %d %d %d %d %g | Etop=%g Ebot=%g step=%g hw=%g a0=%g seed=%d K=%u NRM=%u R=%u "
         "| %d windows\n",
         gf, (int)D, (int)Nt, (int)Nx, axis, bg, Etop, Ebot, stepE, hw, a0, iseed, K, NRM, R, M);
  uint32_t V = Nt;
  for (uint32_t d = 1; d < D; ++d) V *= Nx;
  printf("NPLAQ: %u\n", V * D * (D - 1) / 2);
  printf("NRECT: %u\n", V * D * (D - 1));
  printf("AXIS: %d\n", axis);
  fflush(stdout);
  dsa_group *g = nullptr;
  if (dsa_group_load(gf, &g)) return die(argv[0], "dsa_group_load");
  dsa_llr_ref *C = nullptr;
  if (dsa_llr_ref_create(g, D, Nt, Nx, axis, bg, iseed, &C)) return die(argv[0], "dsa_llr_ref_create");
  double rect_mult = 0;
  dsa_llr_ref_rect_mult(C, &rect_mult);
  printf("Initializing nearest neighbours [%d:%d]... \n", (int)Nt, (int)Nx);
  printf("RECT_MULT calibrated: %.6f\n", rect_mult);
  if (axis == 2 && fabs(rect_mult - 6.0) > 1e-3)
    printf("WARNING: RECT_MULT=%.3f != 6 -- the 1x2 rectangles wrap on this lattice\n"
           "  (extent too small in some direction); the S2 action is degenerate and the\n"
           "  S2 window is invalid. Use a lattice with extent >= 3 for axis=2.\n",
           rect_mult);
  fflush(stdout);
  std::vector<std::vector<double>> an(M);
  // run_pass (llr.cpp:256-274)
  auto run_pass = [&](int dir, int *reached) -> int {
    *reached = 0;
    if (dsa_llr_ref_start(C, dir > 0 ? DSA_START_HOT : DSA_START_COLD)) return die(argv[0], "dsa_llr_ref_start");
    double a_prev = a0;
    for (int s = 0; s < M; ++s) {
      const int n = dir > 0 ? s : M - 1 - s;
      const double Ewin = Etop - n * stepE, Elo = Ewin - hw, Ehi = Ewin + hw;
      int32_t ok = 0;
      if (dsa_llr_ref_seed_window(C, Elo, Ehi, a_prev, &ok)) return die(argv[0], "dsa_llr_ref_seed_window");
      if (!ok) break;
      double a = 0;
      if (dsa_llr_ref_rm_solve(C, Ewin, hw, a_prev, K, NRM, &a)) return die(argv[0], "dsa_llr_ref_rm_solve");
      an[n].push_back(a);
      a_prev = a;
      ++*reached;
    }
    return 0;
  };
  for (unsigned r = 0; r < R; ++r) {
    if (dsa_llr_ref_reseed(C, r)) return die(argv[0], "dsa_llr_ref_reseed");
    int rd = 0, ru = 0;
    if (int rc = run_pass(+1, &rd)) return rc;
    if (int rc = run_pass(-1, &ru)) return rc;
    printf("repeat %u: down reached %d/%d, up reached %d/%d\n", r, rd, M, ru, M);
    fflush(stdout);
  }
  printf("\n# Ewin   a_n        sd        nrep\n");
  for (int n = 0; n < M; ++n) {
    const double Ewin = Etop - n * stepE;
    if (an[n].empty()) {
      printf("ANE: %8.2f   --(unreached)\n", Ewin);
      continue;
    }
    double mean = 0;
    for (double x : an[n]) mean += x;
    mean /= an[n].size();
    double var = 0;
    for (double x : an[n]) var += (x - mean) * (x - mean);
    const double sd = an[n].size() > 1 ? sqrt(var / (an[n].size() - 1)) : 0.0;
    printf("ANE: %8.2f  %9.5f  %9.5f  %lu\n", Ewin, mean, sd, (unsigned long)an[n].size());
  }
  double drift = 0;
  dsa_llr_ref_state(C, nullptr, nullptr, &drift, nullptr, nullptr);
  printf("max |incr - recompute| E(axis%d) : %.3e\n", axis, drift);
  dsa_llr_ref_destroy(C);
  dsa_group_free(g);
  return 0;
}

// -------------------------------------------------------------- llr2d (2-D)

struct Grid {
  double E1top, E1bot, step1, hw1, c0, c1, c2, step2, hw2;
  int nperp, N1, N2;
  double E1(int i) const { return E1top - i * step1; }
  double E2(int i, int j) const {
    const double e = E1(i);
    return c0 + c1 * e + c2 * e * e + (j - nperp) * step2;
  }
};

Grid parse_grid(char **a) {  // a[0..9] = E1top E1bot step1 hw1 c0 c1 c2 step2 hw2 nperp
  Grid G;
  G.E1top = atof(a[0]);
  G.E1bot = atof(a[1]);
  G.step1 = atof(a[2]);
  G.hw1 = atof(a[3]);
  G.c0 = atof(a[4]);
  G.c1 = atof(a[5]);
  G.c2 = atof(a[6]);
  G.step2 = atof(a[7]);
  G.hw2 = atof(a[8]);
  G.nperp = atoi(a[9]);
  G.N1 = (int)((G.E1top - G.E1bot) / G.step1 + 0.5) + 1;
  G.N2 = 2 * G.nperp + 1;
  return G;
}

dsa_llr_params cell_params(uint32_t D, uint32_t Nt, uint32_t Nx, const Grid &G, int i, int j, double a1, double a2,
                           unsigned K, unsigned NRM, int32_t seed) {
  dsa_llr_params p = base_params(D, Nt, Nx);
  p.axis = 3;
  p.Ewin = G.E1(i);
  p.hw = G.hw1;
  p.Ewin2 = G.E2(i, j);
  p.hw2 = G.hw2;
  p.a0 = a1;
  p.a20 = a2;
  p.K = K;
  p.NRM = NRM;
  p.cap = 4000;  // llr2d.cpp:323
  p.seed = seed;
  return p;
}

// The ridge walks of llr2d's full and seed passes (llr2d.cpp:324-396,
// 452-472).  One walk per (repeat, start): ridge cells in order, each seeded
// from the walk's current configuration (a failed ridge cell leaves the
// configuration where its drive-in stopped, llr2d.cpp:333); the transverse
// cells of a solved ridge cell start from its configuration and tilts, and
// each direction stops at its first cell that cannot be reached.  on_cell
// receives (repeat, start, i, j, result) for every reached cell, in the order
// the reference reaches them.
struct Walks {
  const Grid &G;
  uint32_t D, Nt, Nx;
  double a0;
  int iseed;
  unsigned K, NRM;
  std::function<void(unsigned, int, int, int, const Result &)> on_cell;
  Runner *run;
  // per walk: results of the ridge step's fans, in the reference's order
  struct Fan {
    std::vector<Result> res;
    std::vector<bool> done;
    int pending = 0;
  };
  std::map<std::pair<int, int>, Fan> fans;  // key: (walk id, ridge step)
  std::map<int, int> emitted_upto;          // per walk id: ridge steps emitted
  std::map<int, std::map<int, Result>> ridge_res;
  std::map<int, std::map<int, bool>> ridge_ok;

  int walk_id(unsigned r, int cold) const { return (int)(2 * r + (cold ? 0 : 1)); }
  int ridge_i(int cold, int s) const { return cold ? G.N1 - 1 - s : s; }

  void start(unsigned r, int cold, const std::vector<int32_t> &init) { ridge(r, cold, 0, init); }

  void ridge(unsigned r, int cold, int s, std::vector<int32_t> init) {
    if (s >= G.N1) return;
    const int i = ridge_i(cold, s);
    Task t;
    t.label = std::to_string(r) + (cold ? " cold " : " hot ") + std::to_string(i) + " " + std::to_string(G.nperp);
    t.p = cell_params(D, Nt, Nx, G, i, G.nperp, a0, a0, K, NRM, task_seed(iseed, 2 * r + (cold ? 0 : 1), (uint32_t)i, 0));
    t.p.start = cold ? DSA_START_COLD : DSA_START_HOT;
    t.init = std::move(init);
    t.done = [this, r, cold, s, i](const Result &res) {
      const int w = walk_id(r, cold);
      ridge_res[w][s] = res;
      ridge_ok[w][s] = res.reached;
      Fan &f = fans[{w, s}];
      if (res.reached) {
        // transverse cells: both directions, from the ridge configuration
        f.res.assign(2 * G.nperp, Result());
        f.done.assign(2 * G.nperp, false);
        f.pending = 2 * G.nperp;
        int q = 0;
        for (int dir = -1; dir <= 1; dir += 2)
          for (int tt = 1; tt <= G.nperp; ++tt, ++q) {
            const int j = G.nperp + dir * tt;
            Task c;
            c.label = std::to_string(r) + (cold ? " cold " : " hot ") + std::to_string(i) + " " + std::to_string(j);
            c.p = cell_params(D, Nt, Nx, G, i, j, res.a, res.a2, K, NRM,
                              task_seed(iseed, 2 * r + (cold ? 0 : 1), (uint32_t)i, (uint32_t)(j + 1)));
            c.init = res.links;
            c.done = [this, w, s, q](const Result &cr) {
              Fan &ff = fans[{w, s}];
              ff.res[q] = cr;
              ff.done[q] = true;
              --ff.pending;
              emit(w);
            };
            run->submit(std::move(c));
          }
      }
      emit(w);
      // the walk continues from the ridge configuration (or, after a failed
      // drive-in, from where it stopped)
      ridge(r, cold, s + 1, res.links);
    };
    run->submit(std::move(t));
  }

  // report ridge steps whose fans are complete, in walk order
  void emit(int w) {
    int &s = emitted_upto[w];
    const unsigned r = (unsigned)(w / 2);
    const int cold = (w % 2) == 0;
    while (ridge_ok[w].count(s)) {
      Fan &f = fans[{w, s}];
      if (ridge_ok[w][s] && f.pending) return;
      const int i = ridge_i(cold, s);
      if (ridge_ok[w][s]) {
        on_cell(r, cold, i, G.nperp, ridge_res[w][s]);
        int q = 0;
        for (int dir = -1; dir <= 1; dir += 2) {
          bool stop = false;
          for (int tt = 1; tt <= G.nperp; ++tt, ++q) {
            if (stop) continue;
            if (!f.res[q].reached) {
              stop = true;
              continue;
            }
            on_cell(r, cold, i, G.nperp + dir * tt, f.res[q]);
          }
        }
      }
      ridge_res[w].erase(s);
      fans.erase({w, s});
      ++s;
    }
  }
};

int main_llr2d_full(int argc, char **argv) {
  if (argc < 17) {
    fprintf(stderr,
            "usage: %s group D Nt Nx E1top E1bot step1 hw1 c0 c1 c2 step2 hw2 nperp a0 seed [K] [NRM] [R]\n"
            "  windows BOTH S1 and S2; tiles a band around the curved ridge\n"
            "  E2 = c0 + c1*E1 + c2*E1^2, offset by (j-nperp)*step2 (j=0..2*nperp).\n",
            argv[0]);
    return 1;
  }
  const char *gf = argv[1];
  const uint32_t D = (uint32_t)atoi(argv[2]), Nt = (uint32_t)atoi(argv[3]), Nx = (uint32_t)atoi(argv[4]);
  const Grid G = parse_grid(argv + 5);
  const double a0 = atof(argv[15]);
  const int iseed = atoi(argv[16]);
  const unsigned K = argc > 17 ? (unsigned)atoi(argv[17]) : 40;
  const unsigned NRM = argc > 18 ? (unsigned)atoi(argv[18]) : 70;
  const unsigned R = argc > 19 ? (unsigned)atoi(argv[19]) : 1;
  printf("LLR2D(grp,D,Nt,Nx): %s %d %d %d | E1=[%g,%g] step1=%g hw1=%g | ridge c0=%g c1=%g c2=%g step2=%g hw2=%g "
         "nperp=%d | a0=%g seed=%d K=%u NRM=%u R=%u | %dx%d cells\n",
         gf, (int)D, (int)Nt, (int)Nx, G.E1top, G.E1bot, G.step1, G.hw1, G.c0, G.c1, G.c2, G.step2, G.hw2, G.nperp,
         a0, iseed, K, NRM, R, G.N1, G.N2);
  uint32_t V = Nt;
  for (uint32_t d = 1; d < D; ++d) V *= Nx;
  printf("NPLAQ: %u\nNRECT: %u\n", V * D * (D - 1) / 2, V * D * (D - 1));
  fflush(stdout);
  dsa_group *g = nullptr;
  if (dsa_group_load(gf, &g)) return die(argv[0], "dsa_group_load");
  if (!lattice_setup_2d(argv[0], g, D, Nt, Nx, iseed)) return 2;
  std::vector<std::vector<std::vector<double>>> a1s(G.N1, std::vector<std::vector<double>>(G.N2)), a2s = a1s;
  std::vector<int> creach(R, 0), hreach(R, 0);
  double drift1 = 0, drift2 = 0;
  {
    Runner run(argv[0], g, D, std::max<uint32_t>(2 * R * (1 + 2 * G.nperp), 1));
    if (!run.ok()) return die(argv[0], "dsa_llr_create");
    Walks W{G, D, Nt, Nx, a0, iseed, K, NRM, {}, &run, {}, {}, {}, {}};
    W.on_cell = [&](unsigned r, int cold, int i, int j, const Result &res) {
      a1s[i][j].push_back(res.a);
      a2s[i][j].push_back(res.a2);
      if (j == G.nperp) ++(cold ? creach : hreach)[r];
    };
    for (unsigned r = 0; r < R; ++r) {
      W.start(r, 1, {});
      W.start(r, 0, {});
    }
    if (int rc = run.run()) return rc;
    drift1 = run.drift1();
    drift2 = run.drift2();
  }
  for (unsigned r = 0; r < R; ++r)
    printf("repeat %u: cold ridge %d/%d, hot ridge %d/%d\n", r, creach[r], G.N1, hreach[r], G.N1);
  printf("\n# E1 E2 a1 a2 sd1 sd2 nrep\n");
  int nreached = 0;
  for (int i = 0; i < G.N1; ++i)
    for (int j = 0; j < G.N2; ++j) {
      const std::vector<double> &x1 = a1s[i][j], &x2 = a2s[i][j];
      if (x1.empty()) continue;
      ++nreached;
      const int n = (int)x1.size();
      double m1 = 0, m2 = 0;
      for (int k = 0; k < n; ++k) {
        m1 += x1[k];
        m2 += x2[k];
      }
      m1 /= n;
      m2 /= n;
      double v1 = 0, v2 = 0;
      for (int k = 0; k < n; ++k) {
        v1 += (x1[k] - m1) * (x1[k] - m1);
        v2 += (x2[k] - m2) * (x2[k] - m2);
      }
      const double sd1 = n > 1 ? sqrt(v1 / (n - 1)) : 0, sd2 = n > 1 ? sqrt(v2 / (n - 1)) : 0;
      printf("ANE2: %9.2f %9.2f %10.6f %10.6f %9.5f %9.5f %d\n", G.E1(i), G.E2(i, j), m1, m2, sd1, sd2, n);
    }
  printf("reached %d/%d cells\n", nreached, G.N1 * G.N2);
  printf("max |incr-recompute|: S1 %.3e  S2 %.3e\n", drift1, drift2);
  dsa_group_free(g);
  return 0;
}

int main_llr2d_seed(int argc, char **argv) {
  if (argc < 18) {
    fprintf(stderr, "usage: %s seed group D Nt Nx E1top E1bot step1 hw1 c0 c1 c2 step2 hw2 nperp a0 seed cfgdir\n",
            argv[0]);
    return 1;
  }
  const char *gf = argv[1];
  const uint32_t D = (uint32_t)atoi(argv[2]), Nt = (uint32_t)atoi(argv[3]), Nx = (uint32_t)atoi(argv[4]);
  const Grid G = parse_grid(argv + 5);
  const double a0 = atof(argv[15]);
  const int iseed = atoi(argv[16]);
  const char *cfgdir = argv[17];
  printf("LLR2D(seed): %s %d %d %d | E1=[%g,%g] step1=%g hw1=%g | ridge c0=%g c1=%g c2=%g step2=%g hw2=%g nperp=%d\n",
         gf, (int)D, (int)Nt, (int)Nx, G.E1top, G.E1bot, G.step1, G.hw1, G.c0, G.c1, G.c2, G.step2, G.hw2, G.nperp);
  fflush(stdout);
  dsa_group *g = nullptr;
  if (dsa_group_load(gf, &g)) return die(argv[0], "dsa_group_load");
  if (!lattice_setup_2d(argv[0], g, D, Nt, Nx, iseed)) return 2;
  std::vector<std::vector<bool>> saved(G.N1, std::vector<bool>(G.N2, false));
  int idx = 0, err = 0;
  {
    Runner run(argv[0], g, D, std::max<uint32_t>(2 * (1 + 2 * G.nperp), 1));
    if (!run.ok()) return die(argv[0], "dsa_llr_create");
    // NRM = 0: each cell only drives in (llr2d.cpp:443-447); the cold walk
    // runs before the hot one as in the reference (the first save wins)
    Walks W{G, D, Nt, Nx, a0, iseed, 1, 0, {}, &run, {}, {}, {}, {}};
    std::map<int, std::vector<std::pair<std::pair<int, int>, Result>>> pend;
    W.on_cell = [&](unsigned, int cold, int i, int j, const Result &res) {
      pend[cold ? 0 : 1].push_back({{i, j}, res});
    };
    W.start(0, 1, {});
    W.start(0, 0, {});
    if (int rc = run.run()) return rc;
    for (int walk = 0; walk < 2; ++walk)
      for (auto &e : pend[walk]) {
        const int i = e.first.first, j = e.first.second;
        if (saved[i][j]) continue;
        char path[1024];
        snprintf(path, sizeof path, "%s/cell_%d_%d.cfg", cfgdir, i, j);
        if (dsa_cfg_save(path, e.second.links.data(), e.second.links.size())) {
          err = die(argv[0], "dsa_cfg_save");
          continue;
        }
        saved[i][j] = true;
        printf("CELL: %d %d %d %.2f %.2f %.2f %.2f %s\n", idx++, i, j, G.E1(i), G.E2(i, j), G.hw1, G.hw2, path);
      }
  }
  fprintf(stderr, "SEED: saved %d cells to %s\n", idx, cfgdir);
  dsa_group_free(g);
  return err;
}

int main_llr2d_solve(int argc, char **argv) {
  if (argc < 12) {
    fprintf(stderr, "usage: %s solve group D Nt Nx manifest lo hi a0 K NRM seed\n", argv[0]);
    return 1;
  }
  const char *gf = argv[1];
  const uint32_t D = (uint32_t)atoi(argv[2]), Nt = (uint32_t)atoi(argv[3]), Nx = (uint32_t)atoi(argv[4]);
  const char *manifest = argv[5];
  const int lo = atoi(argv[6]), hi = atoi(argv[7]);
  const double a0 = atof(argv[8]);
  const unsigned K = (unsigned)atoi(argv[9]), NRM = (unsigned)atoi(argv[10]);
  const int iseed = atoi(argv[11]);
  // (the reference prints its lattice set-up before it finds a manifest
  // missing, llr2d.cpp:493-495; here the missing manifest is reported first)
  FILE *mf = fopen(manifest, "r");
  if (!mf) {
    fprintf(stderr, "solve: cannot open manifest %s\n", manifest);
    return 1;
  }
  dsa_group *g = nullptr;
  if (dsa_group_load(gf, &g)) return die(argv[0], "dsa_group_load");
  if (!lattice_setup_2d(argv[0], g, D, Nt, Nx, iseed)) return 2;
  uint64_t V = Nt;
  for (uint32_t d = 1; d < D; ++d) V *= Nx;
  struct Cell {
    double E1, E2, hw1, hw2;
    Result res;
  };
  std::vector<Cell> cells;
  char line[2048];
  std::vector<std::vector<int32_t>> cfgs;
  while (fgets(line, sizeof line, mf)) {
    if (strncmp(line, "CELL:", 5)) continue;
    int idx, i, j;
    double E1, E2, hw1, hw2;
    char path[1024];
    if (sscanf(line, "CELL: %d %d %d %lf %lf %lf %lf %1023s", &idx, &i, &j, &E1, &E2, &hw1, &hw2, path) != 8) continue;
    if (idx < lo || idx >= hi) continue;
    std::vector<int32_t> a(V * D);
    if (dsa_cfg_load(path, a.data(), a.size())) {
      fprintf(stderr, "solve: missing cfg %s\n", path);
      continue;
    }
    cells.push_back(Cell{E1, E2, hw1, hw2, Result()});
    cfgs.push_back(std::move(a));
  }
  fclose(mf);
  double drift1 = 0, drift2 = 0;
  {
    Runner run(argv[0], g, D, std::max<uint32_t>((uint32_t)cells.size(), 1));
    if (!run.ok()) return die(argv[0], "dsa_llr_create");
    for (size_t c = 0; c < cells.size(); ++c) {
      Task t;
      t.label = "cell " + std::to_string(c);
      t.p = base_params(D, Nt, Nx);
      t.p.axis = 3;
      t.p.Ewin = cells[c].E1;
      t.p.hw = cells[c].hw1;
      t.p.Ewin2 = cells[c].E2;
      t.p.hw2 = cells[c].hw2;
      t.p.a0 = a0;
      t.p.a20 = a0;
      t.p.K = K;
      t.p.NRM = NRM;
      t.p.cap = 4000;
      t.p.seed = task_seed(iseed, (uint32_t)c, 7, 0);
      t.init = std::move(cfgs[c]);
      t.done = [&cells, c](const Result &r) { cells[c].res = r; };
      run.submit(std::move(t));
    }
    if (int rc = run.run()) return rc;
    drift1 = run.drift1();
    drift2 = run.drift2();
  }
  for (const Cell &c : cells)
    printf("ANE2: %9.2f %9.2f %10.6f %10.6f %9.5f %9.5f %d\n", c.E1, c.E2, c.res.a, c.res.a2, 0.0, 0.0, 1);
  fprintf(stderr, "SOLVE[%d,%d): %d cells, drift S1 %.3e S2 %.3e\n", lo, hi, (int)cells.size(), drift1, drift2);
  dsa_group_free(g);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const std::string prog = basename(argv[0]);
  if (prog.find("llr2d") != std::string::npos) {
    if (argc > 1 && strcmp(argv[1], "seed") == 0) return main_llr2d_seed(argc - 1, argv + 1);
    if (argc > 1 && strcmp(argv[1], "solve") == 0) return main_llr2d_solve(argc - 1, argv + 1);
    return main_llr2d_full(argc, argv);
  }
  // llr --replay: the reference's own chain, bit for bit
  for (int k = 1; k < argc; ++k)
    if (strcmp(argv[k], "--replay") == 0) {
      std::vector<char *> rest;
      for (int j = 0; j < argc; ++j)
        if (j != k) rest.push_back(argv[j]);
      rest.push_back(nullptr);
      return main_llr_replay(argc - 1, rest.data());
    }
  return main_llr(argc, argv);
}
