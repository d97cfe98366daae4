"""shinobi_amd -- MI355X (gfx950) grid-acceleration hot path of Shinobi.

Python mirror of the reference's host-side interface for this path, over the
C ABI of libshinobi_grid.so (include/shinobi_grid.h):

* ``FaceSoup`` -- primitive storage (src/Control/Structure/FaceSoup.h:41-247):
  SoA triangles resident on the device; ``indices`` of a built grid map
  references to triangle ids.
* ``FastSimpleGridBuilder`` / ``FastTwoLevelGridBuilder`` -- the builder
  concept ``init / build / rebuild``
  (src/Control/Algorithm/FastSimpleGridBuilder.h:54-452,
  src/Control/Algorithm/FastTwoLevelGridBuilder.h:56-740), producing a
  ``SimpleGrid`` / ``TwoLevelGrid``.
* ``Grid.trace`` -- the trace kernel contract
  (src/Engine/RayTracingKernels.h:37-136): rays in, (triangle id, t) out.

All arrays exchanged with the host are numpy; device memory is owned by
``DeviceBuffer`` objects.  Everything runs on the GPU through the HIP
library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import ShinobiError, check

MISS = 0xFFFFFFFF
FLT_MAX = np.float32(3.402823466e38)

__all__ = [
    "Context", "DeviceBuffer", "FaceSoup", "FastSimpleGridBuilder", "FastTwoLevelGridBuilder", "Grid",
    "GridInfo", "RayBatch", "ShinobiError", "MISS", "FLT_MAX", "gen_soup", "gen_heightfield",
    "gen_cluster", "gen_camera_rays", "gen_random_rays", "heightfield_height", "build_params",
    "device_count", "lib", "load_obj", "load_camera", "Batch", "Graph", "autotune", "rebuild_many",
]


def lib() -> C.CDLL:
    return L.load()


def device_count() -> int:
    n = C.c_int(0)
    lib().sg_device_count(C.byref(n))
    return n.value


# --------------------------------------------------------------------- memory
class Context:
    """A device context (sg_ctx) plus the stream the host API enqueues on."""

    def __init__(self, device: int = 0, own_stream: bool = True):
        self._l = lib()
        h = C.c_void_p()
        check(self._l.sg_ctx_create(device, C.byref(h)), "sg_ctx_create")
        self.h = h
        self.device = device
        self.stream = C.c_void_p()
        self._owned = weakref.WeakSet()
        if own_stream:
            check(self._l.sg_stream_create(self.h, C.byref(self.stream)), "sg_stream_create")

    def _own(self, obj) -> None:
        self._owned.add(obj)

    # objects that depend on others close first
    _CLOSE_ORDER = ("Server", "Graph", "Batch", "Grid", "FaceSoup", "RayBatch", "DeviceBuffer", "PinnedArray",
                    "Event")

    def close(self) -> None:
        """Close every object still open on this context, then the context."""
        if getattr(self, "h", None):
            rank = {n: i for i, n in enumerate(self._CLOSE_ORDER)}
            for obj in sorted(list(self._owned), key=lambda o: rank.get(type(o).__name__, len(rank))):
                try:
                    obj.close()
                except Exception:
                    pass
            if self.stream:
                self._l.sg_stream_sync(self.h, self.stream)
                self._l.sg_stream_destroy(self.h, self.stream)
                self.stream = C.c_void_p()
            self._l.sg_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self) -> None:
        check(self._l.sg_stream_sync(self.h, self.stream), "sg_stream_sync")

    def alloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def upload(self, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = DeviceBuffer(self, a.nbytes)
        b.upload(a)
        return b

    def event(self) -> "Event":
        return Event(self)


class Event:
    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = C.c_void_p()
        check(ctx._l.sg_event_create(ctx.h, C.byref(self.h)), "sg_event_create")
        ctx._own(self)

    def record(self, stream=None) -> None:
        check(self.ctx._l.sg_event_record(self.ctx.h, self.h, self.ctx.stream if stream is None else stream),
              "sg_event_record")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(self.ctx._l.sg_event_elapsed_ms(self.ctx.h, self.h, end.h, C.byref(ms)), "sg_event_elapsed_ms")
        return float(ms.value)

    def close(self) -> None:
        if self.h and self.ctx.h:
            self.ctx._l.sg_event_destroy(self.ctx.h, self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """A device allocation (sg_malloc); freed when the object is collected."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        check(ctx._l.sg_malloc(ctx.h, max(self.nbytes, 1), C.byref(self.ptr)), "sg_malloc")
        ctx._own(self)

    @property
    def addr(self) -> int:
        return int(self.ptr.value)

    def at(self, byte_offset: int) -> C.c_void_p:
        return C.c_void_p(self.addr + int(byte_offset))

    def upload(self, a: np.ndarray, byte_offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        assert byte_offset + a.nbytes <= self.nbytes
        check(self.ctx._l.sg_memcpy_htod(self.ctx.h, self.at(byte_offset), a.ctypes.data_as(C.c_void_p),
                                         a.nbytes, self.ctx.stream), "sg_memcpy_htod")

    def download(self, dtype, count: int | None = None, byte_offset: int = 0) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - byte_offset) // dtype.itemsize
        out = np.empty(count, dtype)
        assert byte_offset + out.nbytes <= self.nbytes
        check(self.ctx._l.sg_memcpy_dtoh(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.at(byte_offset),
                                         out.nbytes, self.ctx.stream), "sg_memcpy_dtoh")
        return out

    def free(self) -> None:
        if self.ptr and self.ctx.h:
            self.ctx._l.sg_free(self.ctx.h, self.ptr)
        self.ptr = C.c_void_p()

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """A numpy array in pinned host memory (sg_host_alloc); .a is the array."""

    def __init__(self, ctx: Context, shape, dtype=np.float32):
        self.ctx = ctx
        dtype = np.dtype(dtype)
        n = int(np.prod(shape)) * dtype.itemsize
        self.ptr = C.c_void_p()
        check(ctx._l.sg_host_alloc(ctx.h, max(n, 1), C.byref(self.ptr)), "sg_host_alloc")
        ctx._own(self)
        buf = (C.c_char * max(n, 1)).from_address(self.ptr.value)
        # the array keeps this object (and so the pinned memory) alive: a
        # view of .a outliving its PinnedArray must not read freed memory
        buf._owner = self
        self.a = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)

    def close(self) -> None:
        if self.ptr:
            self.a = None
            if self.ctx.h:
                self.ctx._l.sg_host_free(self.ctx.h, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------- rays
class RayBatch:
    """SoA rays (ox,oy,oz,dx,dy,dz,tmax) resident on the device."""

    def __init__(self, ctx: Context, rays7: np.ndarray | None = None, capacity: int | None = None):
        self.ctx = ctx
        if rays7 is not None:
            rays7 = np.ascontiguousarray(rays7, dtype=np.float32)
            assert rays7.shape[0] == 7
            self.n = rays7.shape[1]
        else:
            self.n = int(capacity)
        self.cap = self.n
        self.buf = DeviceBuffer(ctx, 7 * max(self.cap, 1) * 4)
        if rays7 is not None and self.n:
            self.buf.upload(rays7)

    def field(self, k: int) -> C.c_void_p:
        return self.buf.at(4 * k * self.cap)

    def struct(self) -> L.Rays:
        r = L.Rays()
        for i in range(3):
            r.o[i] = self.field(i)
            r.d[i] = self.field(3 + i)
        r.tmax = self.field(6)
        return r

    def download(self, count: int | None = None) -> np.ndarray:
        a = self.buf.download(np.float32, 7 * self.cap).reshape(7, self.cap)
        return a[:, : (self.n if count is None else count)].copy()


# -------------------------------------------------------------------- storage
class FaceSoup:
    """Primitive storage: SoA triangles on the device (sg_scene)."""

    def __init__(self, ctx: Context, soa9: np.ndarray):
        self.ctx = ctx
        soa9 = np.ascontiguousarray(soa9, dtype=np.float32)
        assert soa9.ndim == 2 and soa9.shape[0] == 9
        self.n = int(soa9.shape[1])
        ptrs = L.Float9(*[soa9[i].ctypes.data for i in range(9)])
        h = C.c_void_p()
        check(ctx._l.sg_scene_create(ctx.h, ptrs, self.n, 0, ctx.stream, C.byref(h)), "sg_scene_create")
        self.h = h
        ctx._own(self)

    def upload_vertices(self, soa9: np.ndarray) -> None:
        """FaceSoup::uploadVertices: new positions, same topology."""
        soa9 = np.ascontiguousarray(soa9, dtype=np.float32)
        assert soa9.shape == (9, self.n)
        ptrs = L.Float9(*[soa9[i].ctypes.data for i in range(9)])
        check(self.ctx._l.sg_scene_update(self.h, ptrs, 0, self.ctx.stream), "sg_scene_update")

    def device_soa(self) -> list[int]:
        out = L.Float9()
        check(self.ctx._l.sg_scene_device_soa(self.h, out), "sg_scene_device_soa")
        return [int(x) for x in out]

    def close(self) -> None:
        if getattr(self, "h", None):
            self.ctx._l.sg_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------- grids
def build_params(kind: int, **kw) -> L.BuildParams:
    p = L.BuildParams()
    lib().sg_build_params_default(C.byref(p), kind)
    for k, v in kw.items():
        if k == "lam":
            k = "lambda_"
        if k == "res_override":
            for i in range(3):
                p.res_override[i] = int(v[i])
        else:
            setattr(p, k, v)
    return p


@dataclass
class GridInfo:
    kind: int
    status: int
    res: tuple
    bmin: np.ndarray
    bmax: np.ndarray
    cell_size: np.ndarray
    cell_size_rcp: np.ndarray
    build_rcp: np.ndarray
    n_tris: int
    C1: int
    P1: int
    P1_valid: int
    C2: int
    P2: int
    P2_valid: int
    caps: tuple
    device_bytes: int


ARRAYS = {
    "top_keys": L.SG_ARRAY_TOP_KEYS, "top_vals": L.SG_ARRAY_TOP_VALS, "top_ranges": L.SG_ARRAY_TOP_RANGES,
    "top_cells": L.SG_ARRAY_TOP_CELLS, "leaf_keys": L.SG_ARRAY_LEAF_KEYS, "leaf_vals": L.SG_ARRAY_LEAF_VALS,
    "leaves": L.SG_ARRAY_LEAVES, "indices": L.SG_ARRAY_INDICES, "tri_counts": L.SG_ARRAY_TRI_COUNTS,
    "leaf_counts": L.SG_ARRAY_LEAF_COUNTS, "pair_counts": L.SG_ARRAY_PAIR_COUNTS,
}


class Grid:
    """A built SimpleGrid or TwoLevelGrid (sg_grid) over a FaceSoup."""

    def __init__(self, ctx: Context, scene: FaceSoup, h: C.c_void_p, params: L.BuildParams):
        self.ctx, self.scene, self.h, self.params = ctx, scene, h, params
        ctx._own(self)

    def info(self) -> GridInfo:
        gi = L.GridInfo()
        check(self.ctx._l.sg_grid_get_info(self.h, C.byref(gi)), "sg_grid_get_info")
        f3 = lambda a: np.array(list(a), np.float32)  # noqa: E731
        return GridInfo(gi.kind, gi.status, tuple(gi.res), f3(gi.bmin), f3(gi.bmax), f3(gi.cell_size),
                        f3(gi.cell_size_rcp), f3(gi.build_rcp), gi.n_tris, gi.C1, gi.P1, gi.P1_valid, gi.C2,
                        gi.P2, gi.P2_valid, (gi.cap_C1, gi.cap_P1, gi.cap_C2, gi.cap_P2), gi.device_bytes)

    def download(self, which: str) -> np.ndarray:
        inf = self.info()
        sizes = {"top_keys": inf.P1, "top_vals": inf.P1, "top_ranges": 2 * inf.C1, "top_cells": 2 * inf.C1,
                 "leaf_keys": inf.P2, "leaf_vals": inf.P2, "leaves": 2 * inf.C2,
                 "indices": inf.P2 if inf.kind == L.SG_GRID_TWO_LEVEL else inf.P1,
                 "tri_counts": inf.n_tris, "leaf_counts": inf.C1, "pair_counts": inf.P1}
        out = np.zeros(max(sizes[which], 1), np.uint32)
        check(self.ctx._l.sg_grid_download(self.h, ARRAYS[which], out.ctypes.data_as(C.c_void_p), out.nbytes),
              f"sg_grid_download({which})")
        return out[: sizes[which]]

    def rebuild(self, free_resolution: bool = False, stream=None) -> None:
        f = self.ctx._l.sg_rebuild_free_res if free_resolution else self.ctx._l.sg_rebuild
        check(f(self.ctx.h, self.scene.h, self.h, self.ctx.stream if stream is None else stream), "sg_rebuild")

    def rebuild_phased(self, events: list["Event"]) -> None:
        """rebuild() recording events at the phase boundaries (see
        sg_rebuild_phased in include/shinobi_grid.h)."""
        arr = (C.c_void_p * len(events))(*[e.h.value for e in events])
        check(self.ctx._l.sg_rebuild_phased(self.ctx.h, self.scene.h, self.h, self.ctx.stream, arr, len(events)),
              "sg_rebuild_phased")

    def trace(self, rays: RayBatch, prim: DeviceBuffer | None = None, t: DeviceBuffer | None = None,
              stats: DeviceBuffer | None = None, d_count: C.c_void_p | None = None, n: int | None = None,
              shadow: bool = False):
        """Closest hit (sg_trace_closest) or, with shadow=True, the any-hit
        shadow traversal (sg_trace_shadow: occluded where t < 0.9999)."""
        n = rays.n if n is None else n
        prim = prim or DeviceBuffer(self.ctx, 4 * max(n, 1))
        t = t or DeviceBuffer(self.ctx, 4 * max(n, 1))
        r = rays.struct()
        name = "sg_trace_shadow" if shadow else "sg_trace_closest"
        check(getattr(self.ctx._l, name)(self.ctx.h, self.h, C.byref(r), n, d_count, prim.ptr, t.ptr,
                                         stats.ptr if stats is not None else None, self.ctx.stream), name)
        return prim, t

    def device_view(self) -> L.GridDeviceView:
        """device pointers of the grid and its scene (sg_grid_get_device_view)"""
        v = L.GridDeviceView()
        check(self.ctx._l.sg_grid_get_device_view(self.h, C.byref(v)), "sg_grid_get_device_view")
        return v

    def trace_store(self, rays: RayBatch, raybuf: DeviceBuffer | None = None, prim: DeviceBuffer | None = None,
                    t: DeviceBuffer | None = None, d_count: C.c_void_p | None = None) -> DeviceBuffer:
        """Closest hit with SimpleRayBuffer output (sg_trace_closest_store):
        8 x n floats -- hit point xyz, dir xyz, t, bestHit bits (the reference
        index, 0xFFFFFFFF on a miss); prim / t optional."""
        n = rays.n
        raybuf = raybuf or DeviceBuffer(self.ctx, 8 * 4 * max(n, 1))
        r = rays.struct()
        check(self.ctx._l.sg_trace_closest_store(self.ctx.h, self.h, C.byref(r), n, d_count, raybuf.ptr,
                                                 prim.ptr if prim is not None else None,
                                                 t.ptr if t is not None else None, self.ctx.stream),
              "sg_trace_closest_store")
        return raybuf

    def trace_camera(self, cam: L.Pinhole, tile=(1, 1), prim: DeviceBuffer | None = None,
                     t: DeviceBuffer | None = None, raybuf: DeviceBuffer | None = None,
                     rays_out: RayBatch | None = None, stream=None) -> None:
        """Primary rays generated on the device and traced tile by tile
        (sg_trace_camera); every output is indexed by pixel id."""
        tw, th = (tile, tile) if np.isscalar(tile) else tile
        ro = rays_out.struct() if rays_out is not None else None
        check(self.ctx._l.sg_trace_camera(self.ctx.h, self.h, C.byref(cam), tw, th,
                                          prim.ptr if prim is not None else None,
                                          t.ptr if t is not None else None,
                                          raybuf.ptr if raybuf is not None else None,
                                          C.byref(ro) if ro is not None else None,
                                          self.ctx.stream if stream is None else stream), "sg_trace_camera")

    def trace_host(self, rays7: np.ndarray, with_stats: bool = False, shadow: bool = False):
        """Convenience: upload rays, trace, download (prim, t[, stats])."""
        rb = RayBatch(self.ctx, rays7)
        stats = DeviceBuffer(self.ctx, 16 * max(rb.n, 1)) if with_stats else None
        prim, t = self.trace(rb, stats=stats, shadow=shadow)
        out = (prim.download(np.uint32, rb.n), t.download(np.float32, rb.n))
        if with_stats:
            out = out + (stats.download(np.uint32, 4 * rb.n).reshape(rb.n, 4),)
        return out

    def close(self) -> None:
        if getattr(self, "h", None):
            self.ctx._l.sg_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Builder:
    KIND = L.SG_GRID_UNIFORM

    def __init__(self, ctx: Context, **params):
        self.ctx = ctx
        self.params = build_params(self.KIND, **params)
        self.scene: FaceSoup | None = None
        self.grid: Grid | None = None

    def init(self, scene: FaceSoup) -> None:
        """Builder::init -- bind the storage (bounds and resolution are
        computed on the device by build)."""
        self.scene = scene

    def build(self) -> Grid:
        assert self.scene is not None, "init() first"
        h = C.c_void_p()
        check(self.ctx._l.sg_build(self.ctx.h, self.scene.h, C.byref(self.params), self.ctx.stream, C.byref(h)),
              "sg_build")
        self.grid = Grid(self.ctx, self.scene, h, self.params)
        return self.grid

    def rebuild(self) -> Grid:
        assert self.grid is not None, "build() first"
        self.grid.rebuild()
        return self.grid


class FastSimpleGridBuilder(_Builder):
    """Uniform grid builder (src/Control/Algorithm/FastSimpleGridBuilder.h:43-486)."""
    KIND = L.SG_GRID_UNIFORM


class FastTwoLevelGridBuilder(_Builder):
    """Two-level grid builder (src/Control/Algorithm/FastTwoLevelGridBuilder.h:45-783)."""
    KIND = L.SG_GRID_TWO_LEVEL


def bounce_order(order: str = "class", grid: "Grid | None" = None, box=None, axis=None) -> L.BounceOrder:
    """sg_bounce_order: "class" (sg_bounce's direction-class rounds), "exit"
    (longest exit distance from the grid's bounds, or `box` = (min, max),
    first) or "axis" (smallest |d . axis| first)."""
    q = L.BounceOrder()
    q.order = {"class": L.SG_BOUNCE_CLASS_ROUNDS, "exit": L.SG_BOUNCE_EXIT_FIRST, "axis": L.SG_BOUNCE_AXIS_FIRST}[order]
    if order == "exit":
        if box is None:
            i = grid.info()
            box = (tuple(i.bmin), tuple(i.bmax))
        for k in range(3):
            q.box_min[k], q.box_max[k] = box[0][k], box[1][k]
    if order == "axis":
        for k in range(3):
            q.axis[k] = axis[k]
    return q


def bounce(ctx: Context, scene: FaceSoup, rays: RayBatch, prim: DeviceBuffer, t: DeviceBuffer, seed: int,
           out: RayBatch, src: DeviceBuffer | None = None, count: DeviceBuffer | None = None,
           d_n: C.c_void_p | None = None, stream=None, order: L.BounceOrder | None = None) -> DeviceBuffer:
    """One cosine-weighted diffuse bounce per hit, compacted; returns the
    device counter holding the number of bounce rays.  d_n: optional device
    count of the input rays (<= rays.n); stream: defaults to the context's;
    order: a bounce_order() (default: sg_bounce's direction-class rounds)."""
    count = count or DeviceBuffer(ctx, 4)
    r = rays.struct()
    o = out.struct()
    st = ctx.stream if stream is None else stream
    sp = src.ptr if src is not None else None
    if order is None:
        check(ctx._l.sg_bounce(ctx.h, scene.h, C.byref(r), prim.ptr, t.ptr, rays.n, d_n, seed, C.byref(o), sp,
                               count.ptr, st), "sg_bounce")
    else:
        check(ctx._l.sg_bounce_ordered(ctx.h, scene.h, C.byref(r), prim.ptr, t.ptr, rays.n, d_n, seed,
                                       C.byref(order), C.byref(o), sp, count.ptr, st), "sg_bounce_ordered")
    return count


def trace_paths(ctx: Context, scene: FaceSoup, grid: "Grid", rays: RayBatch, depth: int, seed: int,
                keep: bool = False, order: L.BounceOrder | None = None):
    """Diffuse paths: trace the rays, bounce every hit, trace the bounce rays,
    ... `depth` traced segments in all, every count kept on the device
    (sg_bounce with a device input count, in `order` when given).  Segment
    d > 0 uses bounce seed seed + d - 1.  Returns the device counters of
    segments 1..depth-1 (and, with keep, each segment's (rays, prim, t))."""
    n = rays.n
    bufs = [RayBatch(ctx, capacity=n), RayBatch(ctx, capacity=n)]
    prims = [DeviceBuffer(ctx, 4 * n) for _ in range(2)]
    ts = [DeviceBuffer(ctx, 4 * n) for _ in range(2)]
    counts = [DeviceBuffer(ctx, 4) for _ in range(max(depth - 1, 0))]
    segs = []
    cur, cur_n, k = rays, None, 0
    for d in range(depth):
        p, t = (prims[k], ts[k]) if not keep else (DeviceBuffer(ctx, 4 * n), DeviceBuffer(ctx, 4 * n))
        grid.trace(cur, prim=p, t=t, d_count=cur_n, n=n)
        if keep:
            segs.append((cur, p, t, cur_n))
        if d + 1 < depth:
            nxt = bufs[k] if not keep else RayBatch(ctx, capacity=n)
            bounce(ctx, scene, cur, p, t, seed + d, nxt, count=counts[d], d_n=cur_n, order=order)
            cur, cur_n, k = nxt, counts[d].ptr, k ^ 1
    return (counts, segs) if keep else counts


def scan_u32(ctx: Context, a: np.ndarray, inclusive: bool = False) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint32)
    din = ctx.upload(a)
    dout = DeviceBuffer(ctx, a.nbytes)
    check(ctx._l.sg_scan_u32(ctx.h, din.ptr, dout.ptr, a.size, int(inclusive), ctx.stream), "sg_scan_u32")
    return dout.download(np.uint32, a.size)


def sort_pairs(ctx: Context, keys: np.ndarray, vals: np.ndarray, key_bits: int = 32, radix_bits: int = 8):
    keys = np.ascontiguousarray(keys, np.uint32)
    vals = np.ascontiguousarray(vals, np.uint32)
    dk, dv = ctx.upload(keys), ctx.upload(vals)
    check(ctx._l.sg_sort_pairs(ctx.h, dk.ptr, dv.ptr, keys.size, key_bits, radix_bits, ctx.stream),
          "sg_sort_pairs")
    return dk.download(np.uint32, keys.size), dv.download(np.uint32, keys.size)


# ------------------------------------------------------------------ batching
def rebuild_many(ctx: Context, grids: list, branches: int = 0, stream=None) -> None:
    """Rebuild each grid from its scene (Grid.rebuild for all of them),
    spread over up to `branches` concurrent streams of the context (0: the
    maximum, SG_REBUILD_BRANCHES_MAX) and joined back into `stream`
    (sg_rebuild_many); captures into a Graph as parallel branches."""
    n = len(grids)
    scenes = (C.c_void_p * max(n, 1))(*[g.scene.h.value for g in grids])
    hs = (C.c_void_p * max(n, 1))(*[g.h.value for g in grids])
    check(ctx._l.sg_rebuild_many(ctx.h, n, scenes, hs, int(branches), ctx.stream if stream is None else stream),
          "sg_rebuild_many")


class Batch:
    """Many (grid, ray batch) items traced by one launch per grid kind
    (sg_batch_*)."""

    def __init__(self, ctx: Context, max_items: int):
        self.ctx = ctx
        self.h = C.c_void_p()
        check(ctx._l.sg_batch_create(ctx.h, max_items, C.byref(self.h)), "sg_batch_create")
        ctx._own(self)
        self._keep: list = []

    def clear(self) -> None:
        check(self.ctx._l.sg_batch_clear(self.h), "sg_batch_clear")
        self._keep = []

    def add(self, grid: "Grid", rays: RayBatch, prim: DeviceBuffer, t: DeviceBuffer, n: int | None = None) -> None:
        r = rays.struct()
        check(self.ctx._l.sg_batch_add(self.h, grid.h, C.byref(r), rays.n if n is None else n, prim.ptr, t.ptr),
              "sg_batch_add")
        self._keep += [grid, rays, prim, t]

    def rebuild(self, branches: int = 0, stream=None) -> None:
        """Rebuild every grid of the committed items from its scene's
        current vertices (sg_batch_rebuild): the uniform grids with one
        launch per build phase for all of them, the two-level grids on
        concurrent branches."""
        check(self.ctx._l.sg_batch_rebuild(self.h, int(branches), self.ctx.stream if stream is None else stream),
              "sg_batch_rebuild")

    def commit(self, stream=None) -> None:
        check(self.ctx._l.sg_batch_commit(self.h, self.ctx.stream if stream is None else stream), "sg_batch_commit")

    def launch(self, shadow: bool = False, stream=None) -> None:
        check(self.ctx._l.sg_batch_launch(self.h, int(shadow), self.ctx.stream if stream is None else stream),
              "sg_batch_launch")

    def status(self, n: int, stream=None) -> list[int]:
        """per item (sg_batch_add order), SG_OK or the status of a grid the
        last launch traversed as all misses (sg_batch_get_status)"""
        out = (C.c_int * max(n, 1))()
        check(self.ctx._l.sg_batch_get_status(self.h, self.ctx.stream if stream is None else stream, out, n),
              "sg_batch_get_status")
        return [int(out[i]) for i in range(n)]

    def close(self) -> None:
        if getattr(self, "h", None):
            self.ctx._l.sg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A captured, replayable sequence of library calls on the context's
    stream (sg_graph_*): `with Graph.capture(ctx) as g: ...; g.launch()`."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx._own(self)

    @classmethod
    def capture(cls, ctx: Context) -> "Graph":
        return cls(ctx)

    def __enter__(self) -> "Graph":
        check(self.ctx._l.sg_graph_begin(self.ctx.h, self.ctx.stream), "sg_graph_begin")
        return self

    def __exit__(self, et, ev, tb):
        st = self.ctx._l.sg_graph_end(self.ctx.h, self.ctx.stream, C.byref(self.h))
        if et is None:
            check(st, "sg_graph_end")

    def launch(self) -> None:
        check(self.ctx._l.sg_graph_launch(self.h, self.ctx.stream), "sg_graph_launch")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.ctx._l.sg_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(ctx: Context, scene: FaceSoup, sample: RayBatch, n_rays: int, kinds: int = 3) -> L.TuneResult:
    """Measure candidate (density, radix bits, block size) settings for one
    scene and return the cheapest (sg_autotune); kinds bit 0 = uniform,
    bit 1 = two-level."""
    out = L.TuneResult()
    r = sample.struct()
    check(ctx._l.sg_autotune(ctx.h, scene.h, C.byref(r), sample.n, int(n_rays), kinds, ctx.stream, C.byref(out)),
          "sg_autotune")
    return out


# ------------------------------------------------------------------- serving
@dataclass
class ServerStats:
    requests: int
    rays: int
    waves: int
    p50_us: float
    p99_us: float
    mean_us: float
    max_us: float
    wall_s: float
    mrays_per_s: float
    regrown: int


class Server:
    """Request loop for ray queries against resident grids (sg_server_*).
    mode: "batched" (continuous batching, multi-batch launches) or "plain"
    (one request per launch, synchronous)."""

    def __init__(self, ctx: Context, mode: str = "batched", hbm_budget: int = 1 << 30, max_requests: int = 256,
                 max_waves: int = 2, max_wave_rays: int = 0):
        self.ctx = ctx
        cfg = L.ServerConfig()
        cfg.mode = L.SG_SERVE_PLAIN if mode == "plain" else L.SG_SERVE_BATCHED
        cfg.hbm_budget = int(hbm_budget)
        cfg.max_requests = int(max_requests)
        cfg.max_waves = int(max_waves)
        cfg.max_wave_rays = int(max_wave_rays)
        self.h = C.c_void_p()
        check(ctx._l.sg_server_create(ctx.h, C.byref(cfg), C.byref(self.h)), "sg_server_create")
        ctx._own(self)
        self._keep: dict[int, tuple] = {}
        self._grids: list = []

    def add_grid(self, grid: "Grid") -> int:
        sid = C.c_uint32()
        check(self.ctx._l.sg_server_add_grid(self.h, grid.h, C.byref(sid)), "sg_server_add_grid")
        self._grids.append(grid)
        return int(sid.value)

    def add_dynamic(self, soa9: np.ndarray, **params) -> int:
        soa9 = np.ascontiguousarray(soa9, np.float32)
        p = build_params(params.pop("kind", L.SG_GRID_TWO_LEVEL), **params)
        sid = C.c_uint32()
        check(self.ctx._l.sg_server_add_dynamic(self.h, L.Float9(*[soa9[i].ctypes.data for i in range(9)]),
                                                soa9.shape[1], C.byref(p), C.byref(sid)), "sg_server_add_dynamic")
        return int(sid.value)

    def update_dynamic(self, scene_id: int, soa9: np.ndarray) -> None:
        soa9 = np.ascontiguousarray(soa9, np.float32)
        check(self.ctx._l.sg_server_update_dynamic(self.h, scene_id,
                                                   L.Float9(*[soa9[i].ctypes.data for i in range(9)])),
              "sg_server_update_dynamic")

    def submit(self, scene_id: int, rays7: np.ndarray, prim: np.ndarray | None = None,
               t: np.ndarray | None = None) -> int:
        rays7 = np.ascontiguousarray(rays7, np.float32)
        n = rays7.shape[1]
        prim = np.empty(n, np.uint32) if prim is None else prim
        t = np.empty(n, np.float32) if t is None else t
        rid = C.c_uint64()
        check(self.ctx._l.sg_server_submit(self.h, scene_id, L.Float7(*[rays7[i].ctypes.data for i in range(7)]), n,
                                           prim.ctypes.data, t.ctypes.data, C.byref(rid)), "sg_server_submit")
        self._keep[int(rid.value)] = (rays7, prim, t)
        return int(rid.value)

    def wait(self, rid: int, with_version: bool = False):
        """-> (prim, t, latency_us), or with_version (prim, t, latency_us,
        mesh version of the request's scene it ran on)"""
        lat = C.c_double()
        ver = C.c_uint64()
        st = self.ctx._l.sg_server_wait_version(self.h, rid, C.byref(lat), C.byref(ver))
        _, prim, t = self._keep.pop(rid)
        check(st, "sg_server_wait")
        return (prim, t, float(lat.value), int(ver.value)) if with_version else (prim, t, float(lat.value))

    def drain(self) -> None:
        check(self.ctx._l.sg_server_drain(self.h), "sg_server_drain")
        self._keep.clear()

    def stats(self) -> ServerStats:
        st = L.ServerStats()
        check(self.ctx._l.sg_server_get_stats(self.h, C.byref(st)), "sg_server_get_stats")
        return ServerStats(*[getattr(st, f) for f, _ in L.ServerStats._fields_])

    def reset_stats(self) -> None:
        check(self.ctx._l.sg_server_reset_stats(self.h), "sg_server_reset_stats")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.ctx._l.sg_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- generators
def _soa(n: int, rows: int):
    a = np.empty((rows, n), np.float32)
    ptr_t = L.Float9 if rows == 9 else L.Float7
    return a, ptr_t(*[a[i].ctypes.data for i in range(rows)])


def gen_soup(seed: int, n: int, edge: float = 0.01) -> np.ndarray:
    a, p = _soa(n, 9)
    check(lib().sg_gen_soup(seed, n, edge, p), "sg_gen_soup")
    return a


def gen_heightfield(seed: int, grid: int = 1024, amplitude: float = 0.2) -> np.ndarray:
    n = 2 * (grid - 1) ** 2
    a, p = _soa(n, 9)
    check(lib().sg_gen_heightfield(seed, grid, amplitude, p), "sg_gen_heightfield")
    return a


def heightfield_height(seed: int, x: float, y: float, amplitude: float = 0.2) -> float:
    return float(lib().sg_heightfield_height(seed, x, y, amplitude))


def gen_cluster(seed: int, n_small: int = 500_000, n_large: int = 2_000) -> np.ndarray:
    a, p = _soa(n_small + n_large, 9)
    check(lib().sg_gen_cluster(seed, n_small, n_large, p), "sg_gen_cluster")
    return a


def pinhole(eye, fwd, up, fov_deg: float, res_x: int, res_y: int) -> L.Pinhole:
    """Camera::init (src/Primitive/Camera.h:41-60) for the device generator"""
    c = L.Pinhole()
    f3 = C.c_float * 3
    check(lib().sg_pinhole_init(f3(*eye), f3(*fwd), f3(*up), fov_deg, res_x, res_y, C.byref(c)), "sg_pinhole_init")
    return c


def gen_camera_rays(eye, fwd, up, fov_deg: float, res_x: int, res_y: int) -> np.ndarray:
    a, p = _soa(res_x * res_y, 7)
    f3 = C.c_float * 3
    check(lib().sg_gen_camera_rays(f3(*eye), f3(*fwd), f3(*up), fov_deg, res_x, res_y, p), "sg_gen_camera_rays")
    return a


def load_obj(path: str):
    """Triangles of a Wavefront OBJ file as a (9, T) float32 SoA array (the
    reference's FWObject reader, src/Loaders/FWObjectReader.cpp:181-502) and
    the reader's counts {n_vertices, n_faces, n_triangles, n_bad_lines}."""
    h = C.c_void_p()
    info = L.ObjInfo()
    check(lib().sg_obj_load(os.fsencode(path), C.byref(h), C.byref(info)), "sg_obj_load")
    try:
        a, p = _soa(info.n_triangles, 9)
        check(lib().sg_obj_triangles(h, p), "sg_obj_triangles")
    finally:
        lib().sg_obj_destroy(h)
    return a, {k: getattr(info, k) for k, _ in L.ObjInfo._fields_}


def load_camera(path: str) -> dict:
    """camera.cfg (src/Loaders/CameraManager.cpp:93-165): position,
    orientation, up, rotation, res_x, res_y, fov_deg"""
    cam = L.Camera()
    check(lib().sg_camera_load(os.fsencode(path), C.byref(cam)), "sg_camera_load")
    return {"position": tuple(cam.position), "orientation": tuple(cam.orientation), "up": tuple(cam.up),
            "rotation": tuple(cam.rotation), "res_x": cam.res_x, "res_y": cam.res_y, "fov_deg": cam.fov_deg}


def gen_random_rays(seed: int, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    a, p = _soa(n, 7)
    check(lib().sg_gen_random_rays(seed, n, lo, hi, p), "sg_gen_random_rays")
    return a
