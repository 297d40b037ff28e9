"""ctypes binding of libdsa.so and the Python-side mirror of the reference API.

The reference drives this path through globals (group.h:8-16, lattice.h:7-16)
and free functions (load_group, update, plaquette, polyakov, getpoly, wilson).
Here the same operations are methods on handles; argument meanings and the
element/enumeration orders are the reference's.  Errors raise DSAError with the
library's message instead of aborting.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libdsa.so")
GROUP_DIR = os.path.join(_HERE, "groups")

DSA_RNG = {"philox": 0, "minstd": 1}
DSA_ORDER = {"colour": 0, "reference": 1}
DSA_START = {"hot": 0, "cold": 1}
DSA_SUM = {"tree": 0, "ordered": 1}
DSA_KERNEL = {"auto": 0, "table": 1, "class": 2, "class-narrow": 3, "matrix": 4, "field": 5}
LOOP_KIND = {"plaq": 0, "rect": 1, "poly": 2, "wilson": 3}


class DSAError(RuntimeError):
    """A libdsa call returned a negative status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"dsa error {code}: {msg}")
        self.code = code


class GroupInfo(C.Structure):
    _fields_ = [
        ("P", C.c_uint32),
        ("id", C.c_int32),
        ("npool", C.c_uint32),
        ("min_retr", C.c_double),
        ("nn_retr", C.c_double),
    ]


class ChainParams(C.Structure):
    _fields_ = [
        ("D", C.c_uint32),
        ("Nt", C.c_uint32),
        ("Nx", C.c_uint32),
        ("beta0", C.c_double),
        ("beta1", C.c_double),
        ("beta2", C.c_double),
        ("seed", C.c_int32),
        ("start", C.c_int32),
        ("rng", C.c_int32),
        ("order", C.c_int32),
    ]


_lib: Optional[C.CDLL] = None

_SIGS = {
    "dsa_group_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "dsa_group_info_get": (C.c_int, [C.c_void_p, C.POINTER(GroupInfo)]),
    "dsa_group_tables": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "dsa_group_free": (None, [C.c_void_p]),
    "dsa_chain_create": (C.c_int, [C.c_void_p, C.POINTER(ChainParams), C.POINTER(C.c_void_p)]),
    "dsa_chain_destroy": (None, [C.c_void_p]),
    "dsa_chain_sweep": (C.c_int, [C.c_void_p, C.c_uint32]),
    "dsa_chain_counters": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "dsa_chain_get_sweep": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "dsa_chain_set_sweep": (C.c_int, [C.c_void_p, C.c_uint64]),
    "dsa_chain_measure": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int32, C.c_void_p]),
    "dsa_chain_measure_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int32, C.c_void_p]),
    "dsa_chain_get_links": (C.c_int, [C.c_void_p, C.c_void_p]),
    "dsa_chain_set_links": (C.c_int, [C.c_void_p, C.c_void_p]),
    "dsa_chain_get_rng_state": (C.c_int, [C.c_void_p, C.c_void_p]),
    "dsa_chain_set_rng_state": (C.c_int, [C.c_void_p, C.c_void_p]),
    "dsa_chain_staples": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "dsa_chain_loops": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32, C.c_void_p]),
    "dsa_chain_schedule_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_void_p]),
    "dsa_chain_last_sweep_ms": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "dsa_schedule_query": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_uint32), C.c_void_p, C.c_void_p, C.c_void_p]),
    "dsa_batch_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "dsa_batch_add": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(ChainParams), C.POINTER(C.c_uint32)]),
    "dsa_batch_finalize": (C.c_int, [C.c_void_p]),
    "dsa_batch_set_kernel": (C.c_int, [C.c_void_p, C.c_int32, C.c_double]),
    "dsa_batch_exact_decisions": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "dsa_batch_set_task_links": (C.c_int, [C.c_void_p, C.c_uint32]),
    "dsa_batch_task_links": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_void_p, C.c_uint32]),
    "dsa_batch_destroy": (None, [C.c_void_p]),
    "dsa_batch_sweep": (C.c_int, [C.c_void_p, C.c_uint32]),
    "dsa_batch_measure": (C.c_int, [C.c_void_p, C.c_void_p]),
    "dsa_batch_run_cycles": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    "dsa_batch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32)]),
    "dsa_batch_get_links": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "dsa_batch_counters": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "dsa_batch_last_ms": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "dsa_chain_set_profiling": (C.c_int, [C.c_void_p, C.c_int32]),
    "dsa_chain_set_kernel": (C.c_int, [C.c_void_p, C.c_int32, C.c_double]),
    "dsa_group_matrices": (C.c_int, [C.c_void_p, C.c_void_p]),
    "dsa_group_field": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "dsa_cfg_save": (C.c_int, [C.c_char_p, C.c_void_p, C.c_uint64]),
    "dsa_cfg_load": (C.c_int, [C.c_char_p, C.c_void_p, C.c_uint64]),
    "dsa_chain_save_cfg": (C.c_int, [C.c_void_p, C.c_char_p]),
    "dsa_chain_load_cfg": (C.c_int, [C.c_void_p, C.c_char_p]),
    "dsa_chain_save_kentucky": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.c_char_p, C.c_uint64]),
    "dsa_chain_exact_decisions": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "dsa_chain_kernel_form": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "dsa_chain_kernel_time": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "dsa_probe_gather": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_double)]),
    "dsa_probe_copy": (C.c_int, [C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]),
    "dsa_last_error": (C.c_char_p, []),
    "dsa_version": (C.c_char_p, []),
}


def lib() -> C.CDLL:
    """Load build/libdsa.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise DSAError(rc, lib().dsa_last_error().decode())


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def group_path(name: str) -> str:
    """Path of a packaged group table (e.g. 'mys1080-v4', 'S216')."""
    return os.path.join(GROUP_DIR, name + ".gz")


class Group:
    """A loaded group table (load_group, group.c:13-81)."""

    def __init__(self, path: str):
        h = C.c_void_p()
        _check(lib().dsa_group_load(path.encode(), C.byref(h)))
        self._h = h
        info = GroupInfo()
        _check(lib().dsa_group_info_get(h, C.byref(info)))
        self.P = info.P
        self.id = info.id
        self.npool = info.npool
        self.min_retr = info.min_retr
        self.nn_retr = info.nn_retr
        self.path = path

    def tables(self):
        """(mult[P,P], inv[P], ReTr[P], ImTr[P], pool[npool]) as numpy arrays."""
        P = self.P
        mult = np.empty((P, P), np.int32)
        inv = np.empty(P, np.int32)
        retr = np.empty(P, np.float64)
        imtr = np.empty(P, np.float64)
        pool = np.empty(self.npool, np.int32)
        _check(lib().dsa_group_tables(self._h, _ptr(mult), _ptr(inv), _ptr(retr), _ptr(imtr), _ptr(pool)))
        return mult, inv, retr, imtr, pool

    def matrices(self) -> np.ndarray:
        """The defining-representation matrices appended to the group file, as
        complex [P, 3, 3] (dym-mod-metro-savecfg.cpp:249-262); raises if the
        file has none."""
        raw = np.empty(self.P * 18, np.float64)
        _check(lib().dsa_group_matrices(self._h, _ptr(raw)))
        m = raw.reshape(self.P, 3, 3, 2)
        return m[..., 0] + 1j * m[..., 1]

    def field(self):
        """The group's GF(4) form: (mat, key), mat[e] the 3x3 matrix of e over
        GF(4) as a uint8 [P, 3, 3] array of entries 2 e1 + e0 (for e1 w + e0),
        key the element of each first-two-rows code (4096 uint16, 0xFFFF where
        none); raises if the group has none."""
        mat = np.empty(self.P, np.uint32)
        key = np.empty(4096, np.uint16)
        _check(lib().dsa_group_field(self._h, _ptr(mat), _ptr(key)))
        ent = (mat[:, None] >> (2 * np.arange(9, dtype=np.uint32))) & 3
        return ent.astype(np.uint8).reshape(self.P, 3, 3), key

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().dsa_group_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_group(path: str) -> Group:
    return Group(path)


def schedule(D: int, Nt: int, Nx: int, order: str = "colour"):
    """[(mu, sites ndarray), ...]: the passes of one sweep, in order (host only)."""
    n, m = C.c_uint32(), C.c_uint32()
    o = DSA_ORDER[order]
    _check(lib().dsa_schedule_query(D, Nt, Nx, o, C.byref(n), C.byref(m), None, None, None))
    mu = np.empty(n.value, np.uint32)
    cnt = np.empty(n.value, np.uint32)
    sites = np.empty(m.value, np.uint32)
    _check(lib().dsa_schedule_query(D, Nt, Nx, o, C.byref(n), C.byref(m), _ptr(mu), _ptr(cnt), _ptr(sites)))
    out, off = [], 0
    for q in range(n.value):
        out.append((int(mu[q]), sites[off: off + cnt[q]]))
        off += int(cnt[q])
    return out


class Chain:
    """One Markov chain on the device (dym-mod-metro.cpp:222-268 set-up)."""

    def __init__(self, group: Group, D: int, Nt: int, Nx: int, beta0: float, beta1: float, beta2: float,
                 seed: int, start: str = "hot", rng: str = "philox", order: str = "colour"):
        p = ChainParams(D, Nt, Nx, beta0, beta1, beta2, seed, DSA_START[start], DSA_RNG[rng], DSA_ORDER[order])
        h = C.c_void_p()
        _check(lib().dsa_chain_create(group._h, C.byref(p), C.byref(h)))
        self._h = h
        self.group = group
        self.D, self.Nt, self.Nx = D, Nt, Nx
        self.V = Nt * Nx ** (D - 1)

    # -- dym-mod-metro.cpp:146-197
    def update(self, n: int = 1) -> None:
        _check(lib().dsa_chain_sweep(self._h, n))

    sweep = update

    def counters(self):
        hit, acc = C.c_uint64(), C.c_uint64()
        _check(lib().dsa_chain_counters(self._h, C.byref(hit), C.byref(acc)))
        return hit.value, acc.value

    @property
    def sweeps_done(self) -> int:
        s = C.c_uint64()
        _check(lib().dsa_chain_get_sweep(self._h, C.byref(s)))
        return s.value

    @sweeps_done.setter
    def sweeps_done(self, v: int) -> None:
        _check(lib().dsa_chain_set_sweep(self._h, v))

    def save_cfg(self, path: str) -> None:
        """Raw int32 configuration file (llr2d.cpp:262-265)."""
        _check(lib().dsa_chain_save_cfg(self._h, os.fsencode(path)))

    def load_cfg(self, path: str) -> None:
        """Loads a raw int32 configuration file (llr2d.cpp:266-271)."""
        _check(lib().dsa_chain_load_cfg(self._h, os.fsencode(path)))

    def save_kentucky(self, prefix: str, num: int) -> str:
        """SU(3) configuration in savekyconfig's layout and byte order
        (dym-mod-metro-savecfg.cpp:55-86); returns the file name."""
        name = C.create_string_buffer(1024)
        _check(lib().dsa_chain_save_kentucky(self._h, os.fsencode(prefix), num, name, 1024))
        return name.value.decode()

    def set_kernel(self, kernel: str = "auto", margin_scale: float = 1.0) -> None:
        """Sweep kernel: "table" (element ReTr via L2 gathers), "class" (class
        values in LDS, exact re-decision near the bound), "class-narrow" (the
        byte class-index form), "matrix" (trial actions as FP32 traces of the group's SU(3)
        matrices in LDS, exact re-decision near the rounding bound; groups
        whose file carries the matrices), "field" (the byte class-index form
        with the staple products as GF(4) matrix products; groups with a
        GF(4) form), or "auto"; all give identical links.  A huge
        margin_scale re-decides every hit exactly."""
        _check(lib().dsa_chain_set_kernel(self._h, DSA_KERNEL[kernel], margin_scale))

    def kernel_form(self) -> str:
        """The kernel the chain's passes run: "table", "class" (the int16-value
        form), "class-narrow", "matrix" or "field"."""
        f = C.c_int32()
        _check(lib().dsa_chain_kernel_form(self._h, C.byref(f)))
        return {v: k for k, v in DSA_KERNEL.items()}[f.value]

    def exact_decisions(self) -> int:
        n = C.c_uint64()
        _check(lib().dsa_chain_exact_decisions(self._h, C.byref(n)))
        return n.value

    def last_sweep_ms(self) -> float:
        ms = C.c_double()
        _check(lib().dsa_chain_last_sweep_ms(self._h, C.byref(ms)))
        return ms.value

    # -- dym-mod-metro.cpp:281-318
    def measure(self, wk: Optional[int] = None, wl: Optional[int] = None, sum: str = "tree") -> np.ndarray:
        wk = self.Nx if wk is None else wk
        wl = self.Nt if wl is None else wl
        out = np.empty(4 + wk * wl, np.float64)
        _check(lib().dsa_chain_measure(self._h, wk, wl, DSA_SUM[sum], _ptr(out)))
        return out

    def measure_noalloc(self, wk: Optional[int] = None, wl: Optional[int] = None, sum: str = "tree") -> np.ndarray:
        """measure() with no device allocation: the work space is the
        library's static per-GPU array (raises DSAError DSA_E_RANGE when the
        lattice and grid need more than it holds)."""
        wk = self.Nx if wk is None else wk
        wl = self.Nt if wl is None else wl
        out = np.empty(4 + wk * wl, np.float64)
        _check(lib().dsa_chain_measure_noalloc(self._h, wk, wl, DSA_SUM[sum], _ptr(out)))
        return out

    # -- the gauge field a[i*D+mu] (lattice.h:9)
    @property
    def a(self) -> np.ndarray:
        out = np.empty(self.V * self.D, np.int32)
        _check(lib().dsa_chain_get_links(self._h, _ptr(out)))
        return out

    @a.setter
    def a(self, value) -> None:
        arr = np.ascontiguousarray(value, dtype=np.int32).reshape(-1)
        if arr.size != self.V * self.D:
            raise ValueError(f"expected {self.V * self.D} links, got {arr.size}")
        _check(lib().dsa_chain_set_links(self._h, _ptr(arr)))

    def rng_state(self) -> np.ndarray:
        out = np.empty(self.V, np.uint32)
        _check(lib().dsa_chain_get_rng_state(self._h, _ptr(out)))
        return out

    def set_rng_state(self, st) -> None:
        arr = np.ascontiguousarray(st, dtype=np.uint32)
        _check(lib().dsa_chain_set_rng_state(self._h, _ptr(arr)))

    # -- get_staples / get_rect_loops (dym-mod-metro.cpp:63-142)
    def staples(self):
        ns, nr = 2 * (self.D - 1), 6 * (self.D - 1)
        st = np.empty((self.V * self.D, ns), np.int32)
        re = np.empty((self.V * self.D, nr), np.int32)
        _check(lib().dsa_chain_staples(self._h, _ptr(st), _ptr(re)))
        return st, re

    # -- lattice.c:76-124 loop elements
    def loops(self, kind: str, k: int = 0, l: int = 0) -> np.ndarray:
        D, V = self.D, self.V
        n = {"plaq": V * D * (D - 1) // 2, "rect": V * D * (D - 1), "poly": V // self.Nt,
             "wilson": V * (D - 1)}[kind]
        out = np.empty(n, np.int32)
        _check(lib().dsa_chain_loops(self._h, LOOP_KIND[kind], k, l, _ptr(out)))
        return out

    def schedule_info(self):
        n = C.c_uint32()
        cols = (C.c_uint32 * 8)()
        _check(lib().dsa_chain_schedule_info(self._h, C.byref(n), C.cast(cols, C.c_void_p)))
        return n.value, list(cols)[: self.D]

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().dsa_chain_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Many independent chains swept together (dsa_batch_*): the device-side
    replacement of the reference's one-process-per-point scans."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().dsa_batch_create(C.byref(h)))
        self._h = h
        self.shapes = []
        self._groups = []
        self.final = False

    def add(self, group: Group, D: int, Nt: int, Nx: int, beta0: float, beta1: float, beta2: float, seed: int,
            start: str = "hot", rng: str = "philox", order: str = "colour") -> int:
        p = ChainParams(D, Nt, Nx, beta0, beta1, beta2, seed, DSA_START[start], DSA_RNG[rng], DSA_ORDER[order])
        idx = C.c_uint32()
        _check(lib().dsa_batch_add(self._h, group._h, C.byref(p), C.byref(idx)))
        self.shapes.append((D, Nt, Nx))
        self._groups.append(group)
        return idx.value

    def set_kernel(self, kernel: str = "auto", margin_scale: float = 1.0) -> None:
        """Before finalize: "table", "class" / "class-narrow", "field" (the
        class update with GF(4) products) or "auto" (the class update where it
        fits, with GF(4) products where the group has the form); identical
        results."""
        _check(lib().dsa_batch_set_kernel(self._h, DSA_KERNEL[kernel], margin_scale))

    def exact_decisions(self) -> int:
        n = C.c_uint64()
        _check(lib().dsa_batch_exact_decisions(self._h, C.byref(n)))
        return n.value

    def set_task_links(self, links: int = 0) -> None:
        """Before finalize: links per sweep block (256, 512, 1024) or 0 to tune."""
        _check(lib().dsa_batch_set_task_links(self._h, links))

    def task_links(self):
        """(links per block in use, ms per sweep of 256/512/1024 when tuned)."""
        n = C.c_uint32()
        ms = np.zeros(3, np.float64)
        _check(lib().dsa_batch_task_links(self._h, C.byref(n), _ptr(ms), 3))
        return n.value, ms

    def finalize(self) -> "Batch":
        _check(lib().dsa_batch_finalize(self._h))
        self.final = True
        return self

    def sweep(self, n: int = 1) -> None:
        _check(lib().dsa_batch_sweep(self._h, n))

    update = sweep

    def measure(self) -> np.ndarray:
        out = np.empty((len(self.shapes), 4), np.float64)
        _check(lib().dsa_batch_measure(self._h, _ptr(out)))
        return out

    def run_cycles(self, ncycles: int, sweeps_per_cycle: int) -> np.ndarray:
        out = np.empty((ncycles, len(self.shapes), 4), np.float64)
        _check(lib().dsa_batch_run_cycles(self._h, ncycles, sweeps_per_cycle, _ptr(out)))
        return out

    def links(self, chain: int) -> np.ndarray:
        D, Nt, Nx = self.shapes[chain]
        out = np.empty(Nt * Nx ** (D - 1) * D, np.int32)
        _check(lib().dsa_batch_get_links(self._h, chain, _ptr(out)))
        return out

    def counters(self):
        n = len(self.shapes)
        hit = np.empty(n, np.uint64)
        acc = np.empty(n, np.uint64)
        _check(lib().dsa_batch_counters(self._h, _ptr(hit), _ptr(acc)))
        return hit, acc

    def info(self):
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(lib().dsa_batch_info(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def last_ms(self) -> float:
        ms = C.c_double()
        _check(lib().dsa_batch_last_ms(self._h, C.byref(ms)))
        return ms.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().dsa_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class LlrParams(C.Structure):
    _fields_ = [
        ("D", C.c_uint32), ("Nt", C.c_uint32), ("Nx", C.c_uint32),
        ("axis", C.c_int32), ("bg", C.c_double), ("Ewin", C.c_double), ("hw", C.c_double), ("a0", C.c_double),
        ("seed", C.c_int32), ("start", C.c_int32), ("K", C.c_uint32), ("NRM", C.c_uint32),
        ("tol", C.c_double), ("cap", C.c_uint32),
        ("Ewin2", C.c_double), ("hw2", C.c_double), ("a20", C.c_double),
    ]


class LlrStatus(C.Structure):
    _fields_ = [
        ("a", C.c_double), ("S1", C.c_double), ("S2", C.c_double), ("last_da", C.c_double),
        ("maxdrift", C.c_double), ("phase", C.c_uint32), ("iterations", C.c_uint32),
        ("seed_sweeps", C.c_uint32), ("reached", C.c_uint32), ("sweeps", C.c_uint64), ("hit", C.c_uint64),
        ("acc", C.c_uint64), ("a2", C.c_double), ("drift1", C.c_double), ("drift2", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


LLR_PHASE = {0: "seed", 1: "rm", 2: "done", 3: "idle", 4: "rejected"}


def llr_footprint(params: "LlrParams") -> int:
    """Device bytes one LLR chain of these parameters holds (what serving
    admission counts against the engine's budget)."""
    L = lib()
    L.dsa_llr_footprint.argtypes = [C.POINTER(LlrParams), C.POINTER(C.c_uint64)]
    n = C.c_uint64()
    _check(L.dsa_llr_footprint(C.byref(params), C.byref(n)))
    return n.value


def llr_params(D, Nt, Nx, axis, bg, Ewin, hw, a0, seed, start="hot", K=20, NRM=50, tol=0.0, cap=0, Ewin2=0.0,
               hw2=0.0, a20=0.0):
    """axis 1/2: one window (llr.cpp); axis 3: the box (Ewin, hw) x (Ewin2, hw2)
    on (S1, S2) with tilts (a0, a20) (llr2d.cpp); bg is unused there."""
    return LlrParams(D, Nt, Nx, axis, bg, Ewin, hw, a0, seed, DSA_START[start], K, NRM, tol, cap, Ewin2, hw2, a20)


class LlrEngine:
    """Device-resident LLR chains (dsa_llr_*): llr.cpp's constrained update,
    drive-in and Robbins-Monro iteration for many windows at once."""

    def __init__(self, D: int = 4, max_chains: int = 64, hbm_budget: int = 0):
        L = lib()
        if not hasattr(L.dsa_llr_create, "_sig"):
            vp = C.c_void_p
            L.dsa_llr_create.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(vp)]
            L.dsa_llr_destroy.argtypes = [vp]
            L.dsa_llr_destroy.restype = None
            L.dsa_llr_add.argtypes = [vp, vp, C.POINTER(LlrParams), vp, C.POINTER(C.c_uint32)]
            L.dsa_llr_remove.argtypes = [vp, C.c_uint32]
            L.dsa_llr_step.argtypes = [vp, C.c_uint32]
            L.dsa_llr_status_get.argtypes = [vp, C.c_uint32, C.POINTER(LlrStatus)]
            L.dsa_llr_get_links.argtypes = [vp, C.c_uint32, vp]
            L.dsa_llr_footprint.argtypes = [C.POINTER(LlrParams), C.POINTER(C.c_uint64)]
            L.dsa_llr_serve.argtypes = [vp, vp, vp, vp, C.c_uint32, C.c_uint32, vp, vp, vp]
            L.dsa_llr_set_kernel.argtypes = [vp, C.c_int32, C.c_double]
            L.dsa_llr_create._sig = True
        h = C.c_void_p()
        _check(L.dsa_llr_create(D, max_chains, hbm_budget, C.byref(h)))
        self._h = h
        self.D = D
        self._shapes = {}
        self._keep = []

    def set_kernel(self, kernel: str = "auto", margin_scale: float = 1.0) -> None:
        """Pass kernel of chains added from now on: "table" (element values),
        "class" / "class-narrow" (byte class tables in LDS, certified),
        "field" (the class path with the staple-times-link products from the
        group's GF(4) form) or "auto" (the class path, with GF(4) products
        where the group has the form); identical results.  A huge
        margin_scale decides every hit from element values."""
        _check(lib().dsa_llr_set_kernel(self._h, DSA_KERNEL[kernel], margin_scale))

    def add(self, group: Group, params: LlrParams, links=None) -> int:
        slot = C.c_uint32()
        ptr = None
        if links is not None:
            arr = np.ascontiguousarray(links, dtype=np.int32)
            ptr = _ptr(arr)
        _check(lib().dsa_llr_add(self._h, group._h, C.byref(params), ptr, C.byref(slot)))
        self._shapes[slot.value] = (params.D, params.Nt, params.Nx)
        self._keep.append(group)
        return slot.value

    def remove(self, slot: int) -> None:
        _check(lib().dsa_llr_remove(self._h, slot))
        self._shapes.pop(slot, None)

    def step(self, n: int = 1) -> None:
        _check(lib().dsa_llr_step(self._h, n))

    def status(self, slot: int) -> dict:
        s = LlrStatus()
        _check(lib().dsa_llr_status_get(self._h, slot, C.byref(s)))
        return s.as_dict()

    def links(self, slot: int) -> np.ndarray:
        D, Nt, Nx = self._shapes[slot]
        out = np.empty(Nt * Nx ** (D - 1) * D, np.int32)
        _check(lib().dsa_llr_get_links(self._h, slot, _ptr(out)))
        return out

    def serve(self, groups, params, arrivals, chunk: int = 5):
        """Continuous batching over a request stream; returns (admit_s, done_s, statuses)."""
        n = len(params)
        garr = (C.c_void_p * n)(*[g._h.value for g in groups])
        parr = (LlrParams * n)(*params)
        arr = np.ascontiguousarray(arrivals, np.float64)
        admit = np.zeros(n, np.float64)
        done = np.zeros(n, np.float64)
        st = (LlrStatus * n)()
        self._keep.extend(groups)
        _check(lib().dsa_llr_serve(self._h, C.cast(garr, C.c_void_p), C.cast(parr, C.c_void_p), _ptr(arr), n, chunk,
                                   _ptr(admit), _ptr(done), C.cast(st, C.c_void_p)))
        return admit, done, [s.as_dict() for s in st]

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().dsa_llr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class LlrRef:
    """The reference llr's serial chain replayed on the device
    (dsa_llr_ref_*): main's set-up (llr.cpp:296-321, RECT_MULT calibrated)
    at creation, then run_pass's steps (llr.cpp:223-274) one call at a time,
    bit for bit as the reference runs them at one thread."""

    def __init__(self, group: Group, D: int, Nt: int, Nx: int, axis: int, bg: float, seed: int):
        L = lib()
        if not hasattr(L.dsa_llr_ref_create, "_sig"):
            vp, u32, i32, dbl = C.c_void_p, C.c_uint32, C.c_int32, C.c_double
            L.dsa_llr_ref_create.argtypes = [vp, u32, u32, u32, i32, dbl, i32, C.POINTER(vp)]
            L.dsa_llr_ref_destroy.argtypes = [vp]
            L.dsa_llr_ref_destroy.restype = None
            L.dsa_llr_ref_rect_mult.argtypes = [vp, C.POINTER(dbl)]
            L.dsa_llr_ref_reseed.argtypes = [vp, u32]
            L.dsa_llr_ref_start.argtypes = [vp, i32]
            L.dsa_llr_ref_seed_window.argtypes = [vp, dbl, dbl, dbl, C.POINTER(i32)]
            L.dsa_llr_ref_rm_solve.argtypes = [vp, dbl, dbl, dbl, u32, u32, C.POINTER(dbl)]
            L.dsa_llr_ref_state.argtypes = [vp, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl), C.POINTER(C.c_uint64),
                                            C.POINTER(C.c_uint64)]
            L.dsa_llr_ref_get_links.argtypes = [vp, vp]
            L.dsa_llr_ref_create._sig = True
        h = C.c_void_p()
        _check(L.dsa_llr_ref_create(group._h, D, Nt, Nx, axis, bg, seed, C.byref(h)))
        self._h = h
        self.group = group
        self.V, self.D = Nt * Nx ** (D - 1), D

    @property
    def rect_mult(self) -> float:
        v = C.c_double()
        _check(lib().dsa_llr_ref_rect_mult(self._h, C.byref(v)))
        return v.value

    def reseed(self, repeat: int) -> None:
        _check(lib().dsa_llr_ref_reseed(self._h, repeat))

    def start(self, hot: bool) -> None:
        _check(lib().dsa_llr_ref_start(self._h, DSA_START["hot" if hot else "cold"]))

    def seed_window(self, Elo: float, Ehi: float, a_guess: float) -> bool:
        r = C.c_int32()
        _check(lib().dsa_llr_ref_seed_window(self._h, Elo, Ehi, a_guess, C.byref(r)))
        return bool(r.value)

    def rm_solve(self, Ewin: float, hw: float, a0: float, K: int, NRM: int) -> float:
        a = C.c_double()
        _check(lib().dsa_llr_ref_rm_solve(self._h, Ewin, hw, a0, K, NRM, C.byref(a)))
        return a.value

    def state(self) -> dict:
        s1, s2, md = C.c_double(), C.c_double(), C.c_double()
        hit, acc = C.c_uint64(), C.c_uint64()
        _check(lib().dsa_llr_ref_state(self._h, C.byref(s1), C.byref(s2), C.byref(md), C.byref(hit), C.byref(acc)))
        return {"S1": s1.value, "S2": s2.value, "maxdrift": md.value, "hit": hit.value, "acc": acc.value}

    @property
    def a(self) -> np.ndarray:
        out = np.empty(self.V * self.D, np.int32)
        _check(lib().dsa_llr_ref_get_links(self._h, _ptr(out)))
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().dsa_llr_ref_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
