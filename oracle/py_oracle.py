"""ctypes binding of the CPU oracle (oracle/sg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
cpu_baseline leg of bench.py, always as the checker or as the timed CPU
baseline, never as a product code path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libsg_oracle.so")

MISS = 0xFFFFFFFF


class Params(C.Structure):
    _fields_ = [
        ("lambda_", C.c_float),
        ("top_divisor", C.c_int),
        ("lambda_leaf", C.c_float),
        ("leaf_min_refs", C.c_uint32),
        ("res_override", C.c_int * 3),
        ("nthreads", C.c_int),
    ]


U32P = C.POINTER(C.c_uint32)


class GridC(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("status", C.c_int),
        ("res", C.c_int * 3),
        ("bmin", C.c_float * 3),
        ("bmax", C.c_float * 3),
        ("cs", C.c_float * 3),
        ("rcp_build", C.c_float * 3),
        ("cell_rcp", C.c_float * 3),
        ("n_tris", C.c_uint32),
        ("C1", C.c_uint32), ("P1", C.c_uint32), ("P1_valid", C.c_uint32),
        ("C2", C.c_uint32), ("P2", C.c_uint32), ("P2_valid", C.c_uint32),
        ("tri_counts", U32P), ("top_keys", U32P), ("top_vals", U32P), ("top_ranges", U32P),
        ("top_cells", U32P), ("leaf_counts", U32P), ("pair_counts", U32P), ("leaf_keys", U32P),
        ("leaf_vals", U32P), ("leaves", U32P), ("indices", U32P),
        ("n_indices", C.c_uint32),
    ]


P9 = C.c_void_p * 9
P7 = C.c_void_p * 7

_lib = None


def build() -> str:
    """Compile the oracle (make in oracle/); returns the library path."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    lib.or_det_pow.restype = C.c_double
    lib.or_det_pow.argtypes = [C.c_double, C.c_double]
    lib.or_cbrtpow.restype = C.c_float
    lib.or_cbrtpow.argtypes = [C.c_float]
    lib.or_sincos_2pi.restype = None
    lib.or_sincos_2pi.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.or_exclusive_scan_u32.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    lib.or_inclusive_scan_u32.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    lib.or_sort_pairs.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.or_ranges.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_uint32]
    lib.or_bounds.argtypes = [P9, C.c_uint32, C.c_float * 3, C.c_float * 3]
    lib.or_resolution.argtypes = [C.c_float * 3, C.c_float * 3, C.c_uint32, C.c_float, C.c_int, C.c_int * 3]
    lib.or_build_uniform.restype = C.POINTER(GridC)
    lib.or_build_uniform.argtypes = [P9, C.c_uint32, C.POINTER(Params)]
    lib.or_build_twolevel.restype = C.POINTER(GridC)
    lib.or_build_twolevel.argtypes = [P9, C.c_uint32, C.POINTER(Params)]
    lib.or_grid_free.argtypes = [C.POINTER(GridC)]
    lib.or_trace.argtypes = [C.POINTER(GridC), P9, P7, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    lib.or_trace_shadow.argtypes = lib.or_trace.argtypes
    lib.or_trace_store.argtypes = [C.POINTER(GridC), P9, P7, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    lib.or_brute_f32.argtypes = [P9, C.c_uint32, P7, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int]
    lib.or_brute_f64.argtypes = [P9, C.c_uint32, P7, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int]
    lib.or_bounce.restype = C.c_uint32
    lib.or_bounce.argtypes = [P9, P7, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64, P7, C.c_void_p]
    lib.or_bounce_ordered.restype = C.c_uint32
    F3 = C.c_float * 3
    lib.or_bounce_ordered.argtypes = [P9, P7, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, F3, F3, F3,
                                      P7, C.c_void_p]
    lib.or_hash3.restype = C.c_uint64
    lib.or_hash3.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    lib.or_u01.restype = C.c_float
    lib.or_u01.argtypes = [C.c_uint64]
    _lib = lib
    return lib


def _p9(soa9: np.ndarray):
    assert soa9.dtype == np.float32 and soa9.flags.c_contiguous and soa9.shape[0] == 9
    return P9(*[soa9[i].ctypes.data for i in range(9)])


def _p7(r: np.ndarray):
    assert r.dtype == np.float32 and r.flags.c_contiguous and r.shape[0] == 7
    return P7(*[r[i].ctypes.data for i in range(7)])


def _u32(p, n) -> np.ndarray:
    if n == 0 or not p:
        return np.zeros(0, np.uint32)
    return np.ctypeslib.as_array(p, shape=(n,)).copy()


class OracleGrid:
    """A grid built by the oracle; arrays are copied into numpy."""

    def __init__(self, ptr, soa9: np.ndarray):
        self._p = ptr
        self._soa9 = soa9        # keep the triangles alive for tracing
        g = ptr.contents
        self.kind = g.kind
        self.status = g.status
        self.res = tuple(g.res)
        self.bmin = np.array(list(g.bmin), np.float32)
        self.bmax = np.array(list(g.bmax), np.float32)
        self.cs = np.array(list(g.cs), np.float32)
        self.rcp_build = np.array(list(g.rcp_build), np.float32)
        self.cell_rcp = np.array(list(g.cell_rcp), np.float32)
        self.n_tris = g.n_tris
        self.C1, self.P1, self.P1_valid = g.C1, g.P1, g.P1_valid
        self.C2, self.P2, self.P2_valid = g.C2, g.P2, g.P2_valid
        self.tri_counts = _u32(g.tri_counts, g.n_tris)
        self.top_keys = _u32(g.top_keys, g.P1)
        self.top_vals = _u32(g.top_vals, g.P1)
        self.top_ranges = _u32(g.top_ranges, 2 * g.C1)
        self.top_cells = _u32(g.top_cells, 2 * g.C1) if g.kind == 1 else None
        self.leaf_counts = _u32(g.leaf_counts, g.C1) if g.kind == 1 else None
        self.pair_counts = _u32(g.pair_counts, g.P1) if g.kind == 1 else None
        self.leaf_keys = _u32(g.leaf_keys, g.P2) if g.kind == 1 else None
        self.leaf_vals = _u32(g.leaf_vals, g.P2) if g.kind == 1 else None
        self.leaves = _u32(g.leaves, 2 * g.C2) if g.kind == 1 else None
        self.indices = _u32(g.indices, g.n_indices)

    def trace(self, rays7: np.ndarray, nthreads: int = 8, with_stats: bool = False, shadow: bool = False):
        """Closest hit (or, with shadow=True, the any-hit shadow traversal)."""
        rays7 = np.ascontiguousarray(rays7, np.float32)
        n = rays7.shape[1]
        prim = np.empty(n, np.uint32)
        t = np.empty(n, np.float32)
        stats = np.empty((n, 4), np.uint32) if with_stats else None
        fn = load().or_trace_shadow if shadow else load().or_trace
        fn(self._p, _p9(self._soa9), _p7(rays7), n, prim.ctypes.data, t.ctypes.data,
           stats.ctypes.data if with_stats else None, nthreads)
        return (prim, t, stats) if with_stats else (prim, t)

    def trace_store(self, rays7: np.ndarray, nthreads: int = 8) -> np.ndarray:
        """SimpleRayBuffer words (8, n) float32 of a closest-hit trace"""
        rays7 = np.ascontiguousarray(rays7, np.float32)
        n = rays7.shape[1]
        prim = np.empty(n, np.uint32)
        t = np.empty(n, np.float32)
        buf = np.empty((8, n), np.float32)
        load().or_trace_store(self._p, _p9(self._soa9), _p7(rays7), n, prim.ctypes.data, t.ctypes.data,
                              buf.ctypes.data, nthreads)
        return buf

    def close(self):
        if self._p:
            load().or_grid_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def params(lam=5.0, top_divisor=6, lambda_leaf=1.2, leaf_min_refs=16, res_override=(0, 0, 0), nthreads=8):
    p = Params()
    p.lambda_ = lam
    p.top_divisor = top_divisor
    p.lambda_leaf = lambda_leaf
    p.leaf_min_refs = leaf_min_refs
    for i in range(3):
        p.res_override[i] = int(res_override[i])
    p.nthreads = nthreads
    return p


def build_uniform(soa9: np.ndarray, **kw) -> OracleGrid:
    soa9 = np.ascontiguousarray(soa9, np.float32)
    p = params(**kw)
    g = load().or_build_uniform(_p9(soa9), soa9.shape[1], C.byref(p))
    return OracleGrid(g, soa9)


def build_twolevel(soa9: np.ndarray, **kw) -> OracleGrid:
    soa9 = np.ascontiguousarray(soa9, np.float32)
    p = params(**kw)
    g = load().or_build_twolevel(_p9(soa9), soa9.shape[1], C.byref(p))
    return OracleGrid(g, soa9)


def brute_f32(soa9, rays7, nthreads=8):
    soa9 = np.ascontiguousarray(soa9, np.float32)
    rays7 = np.ascontiguousarray(rays7, np.float32)
    n = rays7.shape[1]
    prim = np.empty(n, np.uint32)
    t = np.empty(n, np.float32)
    load().or_brute_f32(_p9(soa9), soa9.shape[1], _p7(rays7), n, prim.ctypes.data, t.ctypes.data, nthreads)
    return prim, t


def brute_f64(soa9, rays7, nthreads=8):
    soa9 = np.ascontiguousarray(soa9, np.float32)
    rays7 = np.ascontiguousarray(rays7, np.float32)
    n = rays7.shape[1]
    prim = np.empty(n, np.uint32)
    t = np.empty(n, np.float64)
    load().or_brute_f64(_p9(soa9), soa9.shape[1], _p7(rays7), n, prim.ctypes.data, t.ctypes.data, nthreads)
    return prim, t


def bounce(soa9, rays7, prim, t, seed, order=None, box=None, axis=None):
    """bounce rays of every hit; order None: direction-class rounds, "exit":
    longest exit distance from box = (min, max) first, "axis": smallest
    |d . axis| first"""
    soa9 = np.ascontiguousarray(soa9, np.float32)
    rays7 = np.ascontiguousarray(rays7, np.float32)
    prim = np.ascontiguousarray(prim, np.uint32)
    t = np.ascontiguousarray(t, np.float32)
    n = rays7.shape[1]
    out = np.empty((7, max(n, 1)), np.float32)
    src = np.empty(max(n, 1), np.uint32)
    if order is None:
        m = load().or_bounce(_p9(soa9), _p7(rays7), prim.ctypes.data, t.ctypes.data, n, seed, _p7(out),
                             src.ctypes.data)
    else:
        F3 = C.c_float * 3
        lo, hi = box if box is not None else ((0.0,) * 3, (0.0,) * 3)
        ax = axis if axis is not None else (0.0, 0.0, 0.0)
        m = load().or_bounce_ordered(_p9(soa9), _p7(rays7), prim.ctypes.data, t.ctypes.data, n, seed,
                                     {"exit": 1, "axis": 2}[order], F3(*lo), F3(*hi), F3(*ax), _p7(out),
                                     src.ctypes.data)
    return out[:, :m].copy(), src[:m].copy()


def exclusive_scan(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint32)
    out = np.empty_like(a)
    load().or_exclusive_scan_u32(a.ctypes.data, out.ctypes.data, a.size)
    return out


def inclusive_scan(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint32)
    out = np.empty_like(a)
    load().or_inclusive_scan_u32(a.ctypes.data, out.ctypes.data, a.size)
    return out


def sort_pairs(keys, vals, nthreads=1):
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    load().or_sort_pairs(k.ctypes.data, v.ctypes.data, k.size, nthreads)
    return k, v


def ranges(sorted_keys, ncells):
    k = np.ascontiguousarray(sorted_keys, np.uint32)
    cells = np.empty(2 * ncells, np.uint32)
    load().or_ranges(k.ctypes.data, k.size, cells.ctypes.data, ncells)
    return cells


def cbrtpow(x: float) -> float:
    return float(load().or_cbrtpow(x))


def det_pow(x: float, y: float) -> float:
    return float(load().or_det_pow(x, y))


def sincos_2pi(u: float):
    s, c = C.c_float(), C.c_float()
    load().or_sincos_2pi(u, C.byref(s), C.byref(c))
    return s.value, c.value


def bounds(soa9):
    soa9 = np.ascontiguousarray(soa9, np.float32)
    mn, mx = (C.c_float * 3)(), (C.c_float * 3)()
    load().or_bounds(_p9(soa9), soa9.shape[1], mn, mx)
    return np.array(list(mn), np.float32), np.array(list(mx), np.float32)


def resolution(bmin, bmax, n, lam, divisor):
    mn = (C.c_float * 3)(*[float(x) for x in bmin])
    mx = (C.c_float * 3)(*[float(x) for x in bmax])
    res = (C.c_int * 3)()
    st = load().or_resolution(mn, mx, n, lam, divisor, res)
    return st, tuple(res)
