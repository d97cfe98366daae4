"""Synthetic workloads of the grid-acceleration path (scene, rays, build
parameters), shared by the tests, the smoke check and bench.py.

* ``soup10k``   -- 10,000 triangles, centroid U[0,1]^3, vertex offsets
  U[-0.01,0.01]^3 (seed 1); 256x256 pinhole rays from (0.5,0.5,-1) toward +z,
  60 degree vertical FOV; uniform grid, lambda 5.
* ``heightfield2m`` -- 1024x1024-vertex height field over [0,1]^2 =
  2,093,058 triangles, 8-octave value noise of amplitude 0.2 (seed 2);
  1920x1080 rays from (0.5,-0.3,0.6) toward the field centre (row-major, or
  tile by tile) plus one cosine-weighted bounce per hit (seed 102); two-level
  grid, lambda 5 / 6 top, lambda 2 leaf, 16-reference leaf threshold.
* ``cluster`` -- 500,000 tiny triangles in a sphere of radius 0.02 at
  (0.5,0.5,0.5) plus 2,000 large triangles in [0,10]^3 (seed 3); two-level.
* ``scene_file`` -- an OBJ scene and its camera.cfg (the reference's ruins
  scene: 38,665 triangles, 1024x1024 camera); two-level grid with the
  reference builder's defaults (lambda 5 / 6 top, lambda 1.2 leaf).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

import gzip
import os
import shutil
import tempfile

from . import gen_camera_rays, gen_cluster, gen_heightfield, gen_soup, heightfield_height, load_camera, load_obj
from ._lib import SG_GRID_TWO_LEVEL, SG_GRID_UNIFORM


@dataclass
class Workload:
    name: str
    kind: int
    scene: np.ndarray                 # (9, N) float32
    rays: np.ndarray                  # (7, R) float32 primary rays
    build: dict = field(default_factory=dict)
    bounce_seed: int | None = None
    # oracle keyword arguments for the same build
    oracle: dict = field(default_factory=dict)
    # the pinhole camera of the primary rays: (eye, forward, up, fov_deg,
    # res_x, res_y), for the device generator (shinobi_amd.pinhole)
    camera: tuple | None = None


def soup10k(n: int = 10_000, res: int = 256, seed: int = 1) -> Workload:
    scene = gen_soup(seed, n, 0.01)
    cam = ((0.5, 0.5, -1.0), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0), 60.0, res, res)
    rays = gen_camera_rays(*cam)
    return Workload("soup10k", SG_GRID_UNIFORM, scene, rays, dict(lam=5.0),
                    oracle=dict(lam=5.0, top_divisor=1), camera=cam)


def heightfield_pinhole(seed: int, res_x: int, res_y: int, amplitude: float = 0.2) -> tuple:
    """camera of the height field: from (0.5,-0.3,0.6) toward the field centre"""
    hc = heightfield_height(seed, 0.5, 0.5, amplitude)
    eye = (0.5, -0.3, 0.6)
    fwd = (0.5 - eye[0], 0.5 - eye[1], hc - eye[2])
    return (eye, fwd, (0.0, 0.0, 1.0), 60.0, res_x, res_y)


def heightfield_camera(seed: int, res_x: int, res_y: int, amplitude: float = 0.2) -> np.ndarray:
    return gen_camera_rays(*heightfield_pinhole(seed, res_x, res_y, amplitude))


def tile_order(res_x: int, res_y: int, tile) -> np.ndarray:
    """Row-major pixel ids visited tile by tile: tiles of tile = w (square)
    or (w, h) pixels in row-major order, row-major inside a tile (partial
    tiles at the right and bottom edges keep their pixels).  tile = 1 is the
    identity."""
    tw, th = (tile, tile) if np.isscalar(tile) else tile
    ids = np.arange(res_x * res_y, dtype=np.int64).reshape(res_y, res_x)
    if tw <= 1 and th <= 1:
        return ids.reshape(-1)
    out = [ids[y0:y0 + th, x0:x0 + tw].reshape(-1)
           for y0 in range(0, res_y, th) for x0 in range(0, res_x, tw)]
    return np.concatenate(out)


def heightfield2m(grid: int = 1024, res_x: int = 1920, res_y: int = 1080, seed: int = 2,
                  tile=1) -> Workload:
    """tile (w or (w, h)) other than 1: the same primary rays ordered tile by
    tile (tile_order), a renderer's choice of which pixels share a wave"""
    scene = gen_heightfield(seed, grid, 0.2)
    rays = heightfield_camera(seed, res_x, res_y)
    if tile != 1:
        rays = np.ascontiguousarray(rays[:, tile_order(res_x, res_y, tile)])
    return Workload("heightfield2m", SG_GRID_TWO_LEVEL, scene, rays,
                    dict(lam=5.0, top_divisor=6, lambda_leaf=2.0, leaf_min_refs=16), bounce_seed=seed + 100,
                    oracle=dict(lam=5.0, top_divisor=6, lambda_leaf=2.0, leaf_min_refs=16),
                    camera=heightfield_pinhole(seed, res_x, res_y))


def cluster(n_small: int = 500_000, n_large: int = 2_000, res_x: int = 1280, res_y: int = 720,
            seed: int = 3) -> Workload:
    scene = gen_cluster(seed, n_small, n_large)
    cam = ((0.5, 0.5, -0.6), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0), 60.0, res_x, res_y)
    rays = gen_camera_rays(*cam)
    return Workload("cluster", SG_GRID_TWO_LEVEL, scene, rays,
                    dict(lam=5.0, top_divisor=6, lambda_leaf=2.0, leaf_min_refs=16), bounce_seed=seed + 100,
                    oracle=dict(lam=5.0, top_divisor=6, lambda_leaf=2.0, leaf_min_refs=16), camera=cam)


def sphere_mesh(n_tris: int = 200_000, frame: int = 0, seed: int = 5) -> np.ndarray:
    """Latitude/longitude tessellation of a sphere (centre (0.5,0.5,0.5),
    radius 0.3) with about n_tris triangles, its radius displaced by a seeded
    sine wave that moves with `frame` -- the dynamic mesh of the serving
    workload.  Returns (9, N) float32 triangle SoA."""
    rng = np.random.default_rng(seed)
    k = rng.uniform(4.0, 9.0, 3)
    ph = rng.uniform(0.0, 2.0 * np.pi, 3)
    n_lat = max(4, int(np.sqrt(n_tris / 4.0)))
    n_lon = max(8, n_tris // (2 * n_lat))
    th = np.linspace(0.0, np.pi, n_lat + 1)
    phi = np.linspace(0.0, 2.0 * np.pi, n_lon + 1)
    T, P = np.meshgrid(th, phi, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)])
    w = 0.25 * frame
    r = 0.3 + 0.02 * (np.sin(k[0] * d[0] + ph[0] + w) * np.sin(k[1] * d[1] + ph[1] + 0.7 * w)
                      + 0.5 * np.sin(k[2] * d[2] + ph[2] - w))
    v = (0.5 + r * d).astype(np.float32)            # (3, n_lat+1, n_lon+1)
    a, b = v[:, :-1, :-1], v[:, 1:, :-1]
    c, e = v[:, 1:, 1:], v[:, :-1, 1:]
    t1 = np.concatenate([a, b, c]).reshape(9, -1)
    t2 = np.concatenate([a, c, e]).reshape(9, -1)
    return np.ascontiguousarray(np.concatenate([t1, t2], axis=1), np.float32)


def scene_file(obj_path: str, camera_path: str, name: str = "ruins", tile=1) -> Workload:
    """An OBJ scene (plain or .gz) with the primary rays of its camera.cfg
    (row-major, or tile by tile as tile_order for tile other than 1)."""
    if obj_path.endswith(".gz"):
        with tempfile.TemporaryDirectory() as d:
            plain = os.path.join(d, "scene.obj")
            with gzip.open(obj_path, "rb") as f, open(plain, "wb") as g:
                shutil.copyfileobj(f, g)
            scene, _ = load_obj(plain)
    else:
        scene, _ = load_obj(obj_path)
    cam = load_camera(camera_path)
    pin = (tuple(cam["position"]), tuple(cam["orientation"]), tuple(cam["up"]), cam["fov_deg"], cam["res_x"],
           cam["res_y"])
    rays = gen_camera_rays(*pin)
    if tile != 1:
        rays = np.ascontiguousarray(rays[:, tile_order(cam["res_x"], cam["res_y"], tile)])
    return Workload(name, SG_GRID_TWO_LEVEL, scene, rays,
                    dict(lam=5.0, top_divisor=6, lambda_leaf=1.2, leaf_min_refs=16), bounce_seed=107,
                    oracle=dict(lam=5.0, top_divisor=6, lambda_leaf=1.2, leaf_min_refs=16), camera=pin)
