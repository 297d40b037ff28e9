"""Replays one repeat of llr.cpp's two ladders (run_pass, llr.cpp:256-274)
on a chain object with the reference's step API (reseed, start, seed_window,
rm_solve, state, a) and compares it, window by window, with a fixture from
the reference's own functions (tests/golden/llrladder_*.npz)."""
import numpy as np


def ladder_args(f):
    g, D, Nt, Nx, axis, bg, Etop, Ebot, step, hw, a0, seed, K, NRM = [str(x) for x in f["args"]]
    return (g, int(D), int(Nt), int(Nx), int(axis), float(bg), float(Etop), float(Ebot), float(step), float(hw),
            float(a0), int(seed), int(K), int(NRM))


def check_ladder(chain, f):
    """Asserts exact equality with the fixture: RECT_MULT, every window's
    drive-in result and energies, a_n, the energies after rm_solve, the
    running maximum drift, and the links after every window."""
    g, D, Nt, Nx, axis, bg, Etop, Ebot, step, hw, a0, seed, K, NRM = ladder_args(f)
    M = int((Etop - Ebot) / step + 0.5) + 1
    assert chain.rect_mult == float(f["rect_mult"])
    win = f["windows"]
    row = 0
    chain.reseed(0)
    for p in range(2):
        dirn = 1 if p == 0 else -1
        chain.start(hot=dirn > 0)
        assert np.array_equal(chain.a, f[f"start_{p}"]), f"pass {p}: start differs"
        a_prev = a0
        for s in range(M):
            n = s if dirn > 0 else M - 1 - s
            Ewin = Etop - n * step
            ok = chain.seed_window(Ewin - hw, Ewin + hw, a_prev)
            st = chain.state()
            w = win[row]
            assert (int(w[0]), int(w[1]), bool(w[2])) == (p, n, ok), (p, n, ok, w)
            assert (st["S1"], st["S2"]) == (w[3], w[4]), (p, n, "drive-in energies", st, w)
            row += 1
            if not ok:
                break
            an = chain.rm_solve(Ewin, hw, a_prev, K, NRM)
            st = chain.state()
            assert an == w[5], (p, n, "a_n", an, w[5])
            assert (st["S1"], st["S2"], st["maxdrift"]) == (w[6], w[7], w[8]), (p, n, "after rm_solve", st, w)
            assert np.array_equal(chain.a, f[f"links_{p}_{s}"]), f"pass {p} window {n}: links differ"
            a_prev = an
    assert row == len(win)
