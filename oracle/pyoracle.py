"""ctypes wrapper of oracle/build/liboracle.so.

TEST INFRASTRUCTURE ONLY: the checker for tests/, the smoke check and the
CPU-baseline leg of bench.py.  The product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")
_lib = None

RNG = {"minstd": 0, "philox": 1}
ORDER = {"reference": 0, "colour": 1}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} missing: build the oracle (make -C oracle)")
        L = C.CDLL(LIB)
        vp, u32, i32, dbl, u64 = C.c_void_p, C.c_uint32, C.c_int32, C.c_double, C.c_uint64
        sigs = {
            "ogp_group_load": (vp, [C.c_char_p]),
            "ogp_group_free": (None, [vp]),
            "ogp_group_P": (u32, [vp]),
            "ogp_group_id": (i32, [vp]),
            "ogp_group_npool": (u32, [vp]),
            "ogp_group_tables": (None, [vp, vp, vp, vp, vp, vp]),
            "ogp_chain_new": (vp, [vp, u32, u32, u32, dbl, dbl, dbl, i32, i32, i32, i32]),
            "ogp_chain_free": (None, [vp]),
            "ogp_chain_sweep": (None, [vp, u32]),
            "ogp_chain_get_links": (None, [vp, vp]),
            "ogp_chain_set_links": (None, [vp, vp]),
            "ogp_chain_counters": (None, [vp, C.POINTER(u64), C.POINTER(u64)]),
            "ogp_chain_sweeps": (u64, [vp]),
            "ogp_chain_set_sweeps": (None, [vp, u64]),
            "ogp_chain_get_mstate": (None, [vp, vp]),
            "ogp_chain_measure": (None, [vp, u32, u32, i32, vp]),
            "ogp_chain_measure_abs": (None, [vp, u32, u32, vp]),
            "ogp_chain_staples": (None, [vp, vp, vp]),
            "ogp_chain_loops": (None, [vp, i32, u32, u32, vp]),
            "ogp_chain_colouring": (C.c_long, [vp, vp, vp]),
            "ogp_chain_readset": (u32, [vp, u32, u32, vp]),
            "ogp_chain_S1": (dbl, [vp]),
            "ogp_chain_sum_m1": (dbl, [vp]),
            "ogp_philox": (None, [vp, vp, vp]),
            "ogp_exp": (dbl, [dbl]),
            "ogp_exp_many": (None, [vp, vp, C.c_size_t]),
            "ogp_minstd_seed": (u64, [C.c_longlong]),
            "ogp_canonical": (dbl, [C.POINTER(u64)]),
            "ogp_randint": (C.c_int, [C.POINTER(u64), C.c_int, C.c_int]),
        }
        for n, (r, a) in sigs.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data)


class OGroup:
    def __init__(self, path: str):
        self.h = lib().ogp_group_load(path.encode())
        if not self.h:
            raise ValueError(f"oracle: cannot load {path}")
        self.P = lib().ogp_group_P(self.h)
        self.id = lib().ogp_group_id(self.h)
        self.npool = lib().ogp_group_npool(self.h)

    def tables(self):
        P = self.P
        mult = np.empty((P, P), np.int32)
        inv = np.empty(P, np.int32)
        retr = np.empty(P, np.float64)
        imtr = np.empty(P, np.float64)
        pool = np.empty(self.npool, np.int32)
        lib().ogp_group_tables(self.h, _p(mult), _p(inv), _p(retr), _p(imtr), _p(pool))
        return mult, inv, retr, imtr, pool

    def __del__(self):
        if getattr(self, "h", None):
            lib().ogp_group_free(self.h)
            self.h = None


class OChain:
    def __init__(self, g: OGroup, D, Nt, Nx, b0, b1, b2, seed, start="hot", rng="philox", order="colour"):
        self.g = g
        self.h = lib().ogp_chain_new(g.h, D, Nt, Nx, b0, b1, b2, seed, 1 if start == "cold" else 0,
                                     RNG[rng], ORDER[order])
        if not self.h:
            raise ValueError("oracle: chain init failed")
        self.D, self.Nt, self.Nx = D, Nt, Nx
        self.V = Nt * Nx ** (D - 1)

    def sweep(self, n=1):
        lib().ogp_chain_sweep(self.h, n)

    @property
    def a(self):
        out = np.empty(self.V * self.D, np.int32)
        lib().ogp_chain_get_links(self.h, _p(out))
        return out

    @a.setter
    def a(self, v):
        arr = np.ascontiguousarray(v, dtype=np.int32).reshape(-1)
        assert arr.size == self.V * self.D
        lib().ogp_chain_set_links(self.h, _p(arr))

    def counters(self):
        h, a = C.c_uint64(), C.c_uint64()
        lib().ogp_chain_counters(self.h, C.byref(h), C.byref(a))
        return h.value, a.value

    @property
    def sweeps_done(self):
        return lib().ogp_chain_sweeps(self.h)

    @sweeps_done.setter
    def sweeps_done(self, v):
        lib().ogp_chain_set_sweeps(self.h, v)

    def mstate(self):
        out = np.empty(self.V, np.uint32)
        lib().ogp_chain_get_mstate(self.h, _p(out))
        return out

    def measure(self, wk=None, wl=None, mode=0):
        wk = self.Nx if wk is None else wk
        wl = self.Nt if wl is None else wl
        out = np.empty(4 + wk * wl, np.float64)
        lib().ogp_chain_measure(self.h, wk, wl, mode, _p(out))
        return out

    def measure_abs(self, wk=None, wl=None):
        wk = self.Nx if wk is None else wk
        wl = self.Nt if wl is None else wl
        out = np.empty(4 + wk * wl, np.float64)
        lib().ogp_chain_measure_abs(self.h, wk, wl, _p(out))
        return out

    def staples(self):
        ns, nr = 2 * (self.D - 1), 6 * (self.D - 1)
        st = np.empty((self.V * self.D, ns), np.int32)
        re = np.empty((self.V * self.D, nr), np.int32)
        lib().ogp_chain_staples(self.h, _p(st), _p(re))
        return st, re

    def loops(self, kind, k=0, l=0):
        D, V = self.D, self.V
        n = {"plaq": V * D * (D - 1) // 2, "rect": V * D * (D - 1), "poly": V // self.Nt,
             "wilson": V * (D - 1)}[kind]
        out = np.empty(n, np.int32)
        lib().ogp_chain_loops(self.h, {"plaq": 0, "rect": 1, "poly": 2, "wilson": 3}[kind], k, l, _p(out))
        return out

    def colouring(self):
        col = np.empty(self.D * self.V, np.uint8)
        ncol = np.zeros(8, np.uint32)
        conflicts = lib().ogp_chain_colouring(self.h, _p(col), _p(ncol))
        return col.reshape(self.D, self.V), ncol[: self.D], conflicts

    def readset(self, i, d):
        out = np.empty(512, np.uint32)
        n = lib().ogp_chain_readset(self.h, i, d, _p(out))
        return out[:n].copy()

    def S1(self):
        return lib().ogp_chain_S1(self.h)

    def sum_m1(self):
        return lib().ogp_chain_sum_m1(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ogp_chain_free(self.h)
            self.h = None


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.empty(4, np.uint32)
    lib().ogp_philox(_p(c), _p(k), _p(o))
    return o


def exp_many(x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    lib().ogp_exp_many(_p(x), _p(y), x.size)
    return y


class OLlr:
    """The device LLR algorithm restated on the CPU (oracle_llr.c, ogl_dev_*)."""

    def __init__(self, g: OGroup, D, Nt, Nx, axis, bg, Ewin, hw, a0, seed, start="hot", K=20, NRM=50, tol=0.0,
                 cap=0, Ewin2=0.0, hw2=0.0, a20=0.0, links=None):
        L = lib()
        if not hasattr(L.ogp_llr_new, "_sig"):
            vp, u32, i32, dbl = C.c_void_p, C.c_uint32, C.c_int32, C.c_double
            L.ogp_llr_new.restype = vp
            L.ogp_llr_new.argtypes = [vp, u32, u32, u32, i32, dbl, dbl, dbl, dbl, i32, i32, u32, u32, dbl, u32, dbl,
                                      dbl, dbl]
            L.ogp_llr_set_links.argtypes = [vp, vp]
            L.ogp_llr_free.argtypes = [vp]
            L.ogp_llr_step.argtypes = [vp, u32]
            L.ogp_llr_status.argtypes = [vp, vp]
            L.ogp_llr_links.argtypes = [vp, vp]
            L.ogp_llr_new._sig = True
        self.g = g
        self.h = L.ogp_llr_new(g.h, D, Nt, Nx, axis, bg, Ewin, hw, a0, seed, 1 if start == "cold" else 0, K, NRM,
                               tol, cap, Ewin2, hw2, a20)
        if not self.h:
            raise ValueError("oracle: llr init failed")
        self.V = Nt * Nx ** (D - 1)
        self.D = D
        if links is not None:
            arr = np.ascontiguousarray(links, dtype=np.int32)
            assert arr.size == self.V * D
            L.ogp_llr_set_links(self.h, _p(arr))

    def step(self, n=1):
        lib().ogp_llr_step(self.h, n)

    def status(self):
        out = np.empty(15, np.float64)
        lib().ogp_llr_status(self.h, _p(out))
        keys = ["a", "S1", "S2", "last_da", "maxdrift", "phase", "iterations", "seed_sweeps", "reached", "sweeps",
                "hit", "acc", "a2", "drift1", "drift2"]
        return dict(zip(keys, out.tolist()))

    @property
    def a_links(self):
        out = np.empty(self.V * self.D, np.int32)
        lib().ogp_llr_links(self.h, _p(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().ogp_llr_free(self.h)
            self.h = None


class OLlrRef:
    """The reference llr's serial chain restated (oracle_llr.c, ogl_ref_*):
    main's set-up with RECT_MULT calibrated (llr.cpp:296-321), then run_pass's
    steps (llr.cpp:223-274) one call at a time."""

    def __init__(self, g: OGroup, D, Nt, Nx, axis, bg, seed):
        L = lib()
        if not hasattr(L.ogp_lref_new, "_sig"):
            vp, u32, i32, dbl = C.c_void_p, C.c_uint32, C.c_int32, C.c_double
            L.ogp_lref_new.restype = vp
            L.ogp_lref_new.argtypes = [vp, u32, u32, u32, i32, dbl, i32]
            L.ogp_lref_free.argtypes = [vp]
            L.ogp_lref_rect_mult.restype = dbl
            L.ogp_lref_rect_mult.argtypes = [vp]
            L.ogp_lref_reseed.argtypes = [vp, u32]
            L.ogp_lref_start.argtypes = [vp, i32]
            L.ogp_lref_seed_window.restype = C.c_int
            L.ogp_lref_seed_window.argtypes = [vp, dbl, dbl, dbl]
            L.ogp_lref_rm_solve.restype = dbl
            L.ogp_lref_rm_solve.argtypes = [vp, dbl, dbl, dbl, u32, u32]
            L.ogp_lref_state.argtypes = [vp, vp]
            L.ogp_lref_links.argtypes = [vp, vp]
            L.ogp_lref_sweep.argtypes = [vp, dbl, dbl, dbl, i32]
            L.ogp_lref_new._sig = True
        self.g = g
        self.h = L.ogp_lref_new(g.h, D, Nt, Nx, axis, bg, seed)
        if not self.h:
            raise ValueError("oracle: llr reference chain init failed")
        self.V, self.D = Nt * Nx ** (D - 1), D

    @property
    def rect_mult(self):
        return lib().ogp_lref_rect_mult(self.h)

    def reseed(self, r):
        lib().ogp_lref_reseed(self.h, r)

    def start(self, hot):
        lib().ogp_lref_start(self.h, 1 if hot else 0)

    def seed_window(self, Elo, Ehi, a_guess):
        return bool(lib().ogp_lref_seed_window(self.h, Elo, Ehi, a_guess))

    def rm_solve(self, Ewin, hw, a0, K, NRM):
        return lib().ogp_lref_rm_solve(self.h, Ewin, hw, a0, K, NRM)

    def sweep(self, a, Elo, Ehi, drive=False):
        """One update_constrained call (llr.cpp:178-219) at tilt a."""
        lib().ogp_lref_sweep(self.h, a, Elo, Ehi, 1 if drive else 0)

    def state(self):
        out = np.empty(5, np.float64)
        lib().ogp_lref_state(self.h, _p(out))
        return {"S1": out[0], "S2": out[1], "maxdrift": out[2], "hit": int(out[3]), "acc": int(out[4])}

    @property
    def a(self):
        out = np.empty(self.V * self.D, np.int32)
        lib().ogp_lref_links(self.h, _p(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().ogp_lref_free(self.h)
            self.h = None
