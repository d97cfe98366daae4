"""ctypes declarations of the C ABI in include/shinobi_grid.h.

The library is loaded from shinobi_amd/lib/libshinobi_grid.so (built in-tree
by shinobi_amd/build.py).  There is no fallback: if the library is missing
the import fails with an error naming the build command.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libshinobi_grid.so")

SG_OK = 0
STATUS_NAMES = {
    0: "SG_OK", 1: "SG_ERR_INVALID_ARG", 2: "SG_ERR_INVALID_SCENE", 3: "SG_ERR_OUT_OF_MEMORY",
    4: "SG_ERR_HIP", 5: "SG_ERR_CAPACITY", 6: "SG_ERR_LEAF_RES", 7: "SG_ERR_TOO_MANY_CELLS",
    8: "SG_ERR_NEG_EXTENT", 9: "SG_ERR_TIMEOUT", 10: "SG_ERR_BUDGET", 11: "SG_ERR_NO_DEVICE",
    12: "SG_ERR_IO", 13: "SG_ERR_SCRATCH_LIMIT",
}
SG_ERR_SCRATCH_LIMIT = 13
SG_REBUILD_BRANCHES_MAX = 16
SG_BOUNCE_NOALLOC_MAX_RAYS = 16777216

SG_GRID_UNIFORM = 0
SG_GRID_TWO_LEVEL = 1

(SG_ARRAY_TOP_KEYS, SG_ARRAY_TOP_VALS, SG_ARRAY_TOP_RANGES, SG_ARRAY_TOP_CELLS, SG_ARRAY_LEAF_KEYS,
 SG_ARRAY_LEAF_VALS, SG_ARRAY_LEAVES, SG_ARRAY_INDICES, SG_ARRAY_TRI_COUNTS, SG_ARRAY_LEAF_COUNTS,
 SG_ARRAY_PAIR_COUNTS) = range(11)

c_float_p = C.POINTER(C.c_float)
c_u32_p = C.POINTER(C.c_uint32)
Float9 = C.c_void_p * 9
Float7 = C.c_void_p * 7
Float3 = C.c_float * 3


class BuildParams(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("lambda_", C.c_float),
        ("top_divisor", C.c_int),
        ("lambda_leaf", C.c_float),
        ("leaf_min_refs", C.c_uint32),
        ("res_override", C.c_int * 3),
        ("radix_bits", C.c_int),
        ("capacity_slack", C.c_float),
        ("keep_debug", C.c_int),
        ("trace_block", C.c_int),
    ]


class GridInfo(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("status", C.c_int),
        ("res", C.c_int * 3),
        ("bmin", C.c_float * 3),
        ("bmax", C.c_float * 3),
        ("cell_size", C.c_float * 3),
        ("cell_size_rcp", C.c_float * 3),
        ("build_rcp", C.c_float * 3),
        ("n_tris", C.c_uint32),
        ("C1", C.c_uint32), ("P1", C.c_uint32), ("P1_valid", C.c_uint32),
        ("C2", C.c_uint32), ("P2", C.c_uint32), ("P2_valid", C.c_uint32),
        ("cap_C1", C.c_uint32), ("cap_P1", C.c_uint32), ("cap_C2", C.c_uint32), ("cap_P2", C.c_uint32),
        ("device_bytes", C.c_uint64),
    ]


class Rays(C.Structure):
    _fields_ = [("o", C.c_void_p * 3), ("d", C.c_void_p * 3), ("tmax", C.c_void_p)]


class TuneResult(C.Structure):
    _fields_ = [("best", BuildParams), ("build_ms", C.c_float), ("trace_ms", C.c_float), ("cost_ms", C.c_float),
                ("candidates", C.c_uint32)]


SG_SERVE_BATCHED = 0
SG_SERVE_PLAIN = 1


class ServerConfig(C.Structure):
    _fields_ = [("mode", C.c_int), ("hbm_budget", C.c_size_t), ("max_requests", C.c_uint32),
                ("max_waves", C.c_uint32), ("max_wave_rays", C.c_uint32)]


class ObjInfo(C.Structure):
    _fields_ = [("n_vertices", C.c_uint32), ("n_faces", C.c_uint32), ("n_triangles", C.c_uint32),
                ("n_bad_lines", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("orientation", C.c_float * 3), ("up", C.c_float * 3),
                ("rotation", C.c_float * 3), ("res_x", C.c_uint32), ("res_y", C.c_uint32), ("fov_deg", C.c_float)]


class Pinhole(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("top_left", C.c_float * 3), ("step_x", C.c_float * 3),
                ("step_y", C.c_float * 3), ("res_x", C.c_uint32), ("res_y", C.c_uint32)]


class GridDeviceView(C.Structure):
    _fields_ = [("kind", C.c_int), ("status", C.c_int), ("res", C.c_int * 3), ("bmin", C.c_float * 3),
                ("bmax", C.c_float * 3), ("cell_size", C.c_float * 3), ("cell_size_rcp", C.c_float * 3),
                ("cells", C.c_void_p), ("leaves", C.c_void_p), ("n_cells", C.c_uint32), ("n_leaves", C.c_uint32),
                ("indices", C.c_void_p), ("n_refs", C.c_uint32), ("tri", C.c_void_p * 9), ("n_tris", C.c_uint32)]


SG_BOUNCE_CLASS_ROUNDS, SG_BOUNCE_EXIT_FIRST, SG_BOUNCE_AXIS_FIRST = 0, 1, 2


class BounceOrder(C.Structure):
    _fields_ = [("order", C.c_int), ("box_min", C.c_float * 3), ("box_max", C.c_float * 3),
                ("axis", C.c_float * 3)]


class ServerStats(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("rays", C.c_uint64), ("waves", C.c_uint64),
                ("p50_us", C.c_double), ("p99_us", C.c_double), ("mean_us", C.c_double), ("max_us", C.c_double),
                ("wall_s", C.c_double), ("mrays_per_s", C.c_double), ("regrown", C.c_uint64)]


# name -> (restype, argtypes); the authoritative list of exported entry points
SIGNATURES = {
    "sg_status_str": (C.c_char_p, [C.c_int]),
    "sg_build_params_default": (None, [C.POINTER(BuildParams), C.c_int]),
    "sg_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "sg_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "sg_ctx_destroy": (C.c_int, [C.c_void_p]),
    "sg_scene_create": (C.c_int, [C.c_void_p, Float9, C.c_uint32, C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    "sg_scene_update": (C.c_int, [C.c_void_p, Float9, C.c_int, C.c_void_p]),
    "sg_scene_destroy": (C.c_int, [C.c_void_p]),
    "sg_scene_size": (C.c_uint32, [C.c_void_p]),
    "sg_scene_device_soa": (C.c_int, [C.c_void_p, Float9]),
    "sg_build": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BuildParams), C.c_void_p, C.POINTER(C.c_void_p)]),
    "sg_rebuild": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sg_rebuild_free_res": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sg_rebuild_phased": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p),
                                    C.c_int]),
    "sg_grid_get_info": (C.c_int, [C.c_void_p, C.POINTER(GridInfo)]),
    "sg_grid_download": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]),
    "sg_grid_get_device_view": (C.c_int, [C.c_void_p, C.POINTER(GridDeviceView)]),
    "sg_grid_destroy": (C.c_int, [C.c_void_p]),
    "sg_trace_closest": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Rays), C.c_uint32, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_void_p]),
    "sg_trace_shadow": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Rays), C.c_uint32, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p]),
    "sg_bounce": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Rays), C.c_void_p, C.c_void_p, C.c_uint32,
                            C.c_void_p, C.c_uint64, C.POINTER(Rays), C.c_void_p, C.c_void_p, C.c_void_p]),
    "sg_trace_closest_store": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Rays), C.c_uint32, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p]),
    "sg_pinhole_init": (C.c_int, [Float3, Float3, Float3, C.c_float, C.c_uint32, C.c_uint32, C.POINTER(Pinhole)]),
    "sg_trace_camera": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Pinhole), C.c_uint32, C.c_uint32, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.POINTER(Rays), C.c_void_p]),
    "sg_bounce_ordered": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Rays), C.c_void_p, C.c_void_p, C.c_uint32,
                                    C.c_void_p, C.c_uint64, C.POINTER(BounceOrder), C.POINTER(Rays), C.c_void_p,
                                    C.c_void_p, C.c_void_p]),
    "sg_bounce_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Rays), C.c_void_p, C.c_void_p, C.c_uint32,
                                    C.c_void_p, C.c_uint64, C.POINTER(Rays), C.c_void_p, C.c_void_p, C.c_void_p]),
    "sg_scan_u32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]),
    "sg_sort_pairs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.c_void_p]),
    "sg_malloc": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "sg_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sg_host_alloc": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "sg_host_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sg_memcpy_htod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "sg_memcpy_dtoh": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "sg_stream_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "sg_stream_destroy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sg_stream_sync": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sg_event_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "sg_event_destroy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sg_event_record": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "sg_rebuild_many": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_uint32,
                                  C.c_void_p]),
    "sg_batch_rebuild": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "sg_stream_wait_event": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "sg_event_elapsed_ms": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "sg_batch_create": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "sg_batch_destroy": (C.c_int, [C.c_void_p]),
    "sg_batch_clear": (C.c_int, [C.c_void_p]),
    "sg_batch_add": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Rays), C.c_uint32, C.c_void_p, C.c_void_p]),
    "sg_batch_commit": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sg_batch_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "sg_batch_get_status": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.c_uint32]),
    "sg_graph_begin": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sg_graph_end": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "sg_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sg_graph_destroy": (C.c_int, [C.c_void_p]),
    "sg_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Rays), C.c_uint32, C.c_uint64, C.c_int,
                              C.c_void_p, C.POINTER(TuneResult)]),
    "sg_server_create": (C.c_int, [C.c_void_p, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "sg_server_destroy": (C.c_int, [C.c_void_p]),
    "sg_server_add_grid": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)]),
    "sg_server_add_dynamic": (C.c_int, [C.c_void_p, Float9, C.c_uint32, C.POINTER(BuildParams),
                                        C.POINTER(C.c_uint32)]),
    "sg_server_update_dynamic": (C.c_int, [C.c_void_p, C.c_uint32, Float9]),
    "sg_server_submit": (C.c_int, [C.c_void_p, C.c_uint32, Float7, C.c_uint32, C.c_void_p, C.c_void_p,
                                   C.POINTER(C.c_uint64)]),
    "sg_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_double)]),
    "sg_server_wait_version": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "sg_server_drain": (C.c_int, [C.c_void_p]),
    "sg_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "sg_server_reset_stats": (C.c_int, [C.c_void_p]),
    "sg_gen_soup": (C.c_int, [C.c_uint64, C.c_uint32, C.c_float, Float9]),
    "sg_gen_heightfield": (C.c_int, [C.c_uint64, C.c_uint32, C.c_float, Float9]),
    "sg_heightfield_height": (C.c_float, [C.c_uint64, C.c_float, C.c_float, C.c_float]),
    "sg_gen_cluster": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, Float9]),
    "sg_gen_camera_rays": (C.c_int, [C.c_float * 3, C.c_float * 3, C.c_float * 3, C.c_float, C.c_uint32,
                                     C.c_uint32, Float7]),
    "sg_gen_random_rays": (C.c_int, [C.c_uint64, C.c_uint32, C.c_float, C.c_float, Float7]),
    "sg_obj_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(ObjInfo)]),
    "sg_obj_triangles": (C.c_int, [C.c_void_p, Float9]),
    "sg_obj_destroy": (None, [C.c_void_p]),
    "sg_camera_load": (C.c_int, [C.c_char_p, C.POINTER(Camera)]),
}

_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python shinobi_amd/build.py`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class ShinobiError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        name = STATUS_NAMES.get(status, str(status))
        msg = load().sg_status_str(status).decode()
        super().__init__(f"{what}: {name} ({msg})")


def check(status: int, what: str) -> None:
    if status != SG_OK:
        raise ShinobiError(status, what)
