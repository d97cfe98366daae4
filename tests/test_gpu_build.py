"""GPU parity of the grid builds against the oracle: every array bit-exact.

Inputs are the seeded synthetic workloads; the oracle builds the same scene
with the same parameters on the CPU.
"""
import numpy as np
import pytest

import shinobi_amd as sa
from shinobi_amd import workloads as W

pytestmark = pytest.mark.gpu


def _gpu_build(ctx, soa9, kind, **kw):
    scene = sa.FaceSoup(ctx, soa9)
    B = sa.FastTwoLevelGridBuilder if kind == 1 else sa.FastSimpleGridBuilder
    b = B(ctx, keep_debug=1, **kw)
    b.init(scene)
    g = b.build()
    return scene, g


def _assert_same_params(info, og):
    assert info.status == 0
    assert tuple(info.res) == tuple(og.res)
    np.testing.assert_array_equal(info.bmin, og.bmin)
    np.testing.assert_array_equal(info.bmax, og.bmax)
    np.testing.assert_array_equal(info.cell_size, og.cs)
    np.testing.assert_array_equal(info.build_rcp, og.rcp_build)
    np.testing.assert_array_equal(info.cell_size_rcp, og.cell_rcp)
    assert (info.C1, info.P1, info.P1_valid) == (og.C1, og.P1, og.P1_valid)


def _assert_top_level(g, og):
    np.testing.assert_array_equal(g.download("tri_counts"), og.tri_counts)
    np.testing.assert_array_equal(g.download("top_keys"), og.top_keys)
    np.testing.assert_array_equal(g.download("top_vals"), og.top_vals)
    np.testing.assert_array_equal(g.download("top_ranges"), og.top_ranges)


def _assert_leaf_level(g, og):
    info = g.info()
    assert (info.C2, info.P2, info.P2_valid) == (og.C2, og.P2, og.P2_valid)
    np.testing.assert_array_equal(g.download("leaf_counts"), og.leaf_counts)
    np.testing.assert_array_equal(g.download("top_cells"), og.top_cells)
    np.testing.assert_array_equal(g.download("pair_counts"), og.pair_counts)
    np.testing.assert_array_equal(g.download("leaf_keys"), og.leaf_keys)
    np.testing.assert_array_equal(g.download("leaf_vals"), og.leaf_vals)
    np.testing.assert_array_equal(g.download("leaves"), og.leaves)
    np.testing.assert_array_equal(g.download("indices"), og.indices)


def test_uniform_soup10k_bit_exact(ctx, oracle):
    w = W.soup10k()
    og = oracle.build_uniform(w.scene, **w.oracle)
    assert og.status == 0
    _, g = _gpu_build(ctx, w.scene, 0, **w.build)
    _assert_same_params(g.info(), og)
    _assert_top_level(g, og)
    np.testing.assert_array_equal(g.download("indices"), og.indices)


@pytest.mark.parametrize("grid", [64, 256])
def test_twolevel_heightfield_small_bit_exact(ctx, oracle, grid):
    w = W.heightfield2m(grid=grid, res_x=64, res_y=32)
    og = oracle.build_twolevel(w.scene, **w.oracle)
    assert og.status == 0
    _, g = _gpu_build(ctx, w.scene, 1, **w.build)
    _assert_same_params(g.info(), og)
    _assert_top_level(g, og)
    _assert_leaf_level(g, og)


def test_twolevel_heightfield2m_bit_exact(ctx, oracle):
    w = W.heightfield2m(res_x=64, res_y=32)
    assert w.scene.shape[1] == 2_093_058
    og = oracle.build_twolevel(w.scene, **w.oracle)
    assert og.status == 0
    _, g = _gpu_build(ctx, w.scene, 1, **w.build)
    _assert_same_params(g.info(), og)
    _assert_top_level(g, og)
    _assert_leaf_level(g, og)


def test_twolevel_cluster_bit_exact(ctx, oracle):
    """the cluster's dense top cell is a leaf tile too large for LDS: its
    pairs take the parallel emission and the global leaf sort, in the build
    and in every rebuild"""
    w = W.cluster(n_small=60_000, n_large=400, res_x=32, res_y=32)
    og = oracle.build_twolevel(w.scene, **w.oracle)
    _, g = _gpu_build(ctx, w.scene, 1, **w.build)
    info = g.info()
    assert info.status == og.status
    _assert_top_level(g, og)
    _assert_leaf_level(g, og)
    g.rebuild()
    _assert_top_level(g, og)
    _assert_leaf_level(g, og)


def test_twolevel_cluster_full_size_bit_exact(ctx, oracle):
    """the dense-cluster configuration at full size (500,000 tiny triangles
    in a sphere of radius 0.02 plus 2,000 large ones in [0, 10]^3): every
    top- and leaf-level array equals the oracle's, after the build and after
    a rebuild"""
    w = W.cluster(res_x=16, res_y=16)
    assert w.scene.shape[1] == 502_000
    og = oracle.build_twolevel(w.scene, **w.oracle)
    assert og.status == 0
    _, g = _gpu_build(ctx, w.scene, 1, **w.build)
    _assert_same_params(g.info(), og)
    _assert_top_level(g, og)
    _assert_leaf_level(g, og)
    g.rebuild()
    _assert_same_params(g.info(), og)
    _assert_top_level(g, og)
    _assert_leaf_level(g, og)


def test_rebuild_into_an_oversize_tile(ctx, oracle):
    """a grid built without oversize leaf tiles and rebuilt after 5,000 of its
    triangles moved into one top cell: the one-block emission of such tiles
    (used until a build of the grid has met one) gives the oracle's grid"""
    a = sa.gen_soup(71, 20_000, 0.01)
    b = a.copy()
    k = 5_000
    c = np.float32(0.301)
    for j in range(3):   # shrink triangles 0..k-1 around (0.301, 0.301, 0.301)
        b[3 * j:3 * j + 3, :k] = c + (a[3 * j:3 * j + 3, :k] - np.float32(0.5)) * np.float32(0.004)
    scene, g = _gpu_build(ctx, a, 1, lam=5.0, lambda_leaf=2.0)
    res = tuple(g.info().res)
    scene.upload_vertices(b)
    g.rebuild()
    og = oracle.build_twolevel(b, lam=5.0, top_divisor=6, lambda_leaf=2.0, res_override=res)
    info = g.info()
    assert info.status == og.status == 0
    assert tuple(info.res) == tuple(og.res)
    _assert_top_level(g, og)
    _assert_leaf_level(g, og)
    # the traversal reads the reference records the block wrote for the
    # oversize tile: rays through the dense cell hit what the oracle hits
    rays = sa.gen_random_rays(72, 30_000, 0.28, 0.32)
    pg, tg = g.trace_host(rays)
    po, to = og.trace(rays)
    assert (po != sa.MISS).sum() > 1000
    np.testing.assert_array_equal(pg, po)
    np.testing.assert_array_equal(tg.view(np.uint32), to.view(np.uint32))


def test_twolevel_reference_leaf_lambda(ctx, oracle):
    """lambda_leaf = 1.2 (TwoLevelBuildKernels.h:49) on the 10k soup"""
    w = W.soup10k()
    og = oracle.build_twolevel(w.scene, lam=5.0, top_divisor=6, lambda_leaf=1.2, leaf_min_refs=16)
    _, g = _gpu_build(ctx, w.scene, 1, lam=5.0, top_divisor=6, lambda_leaf=1.2, leaf_min_refs=16)
    _assert_same_params(g.info(), og)
    _assert_top_level(g, og)
    _assert_leaf_level(g, og)


@pytest.mark.parametrize("radix_bits", [4, 6, 8])
def test_radix_width_does_not_change_the_grid(ctx, oracle, radix_bits):
    w = W.heightfield2m(grid=200, res_x=8, res_y=8)
    og = oracle.build_twolevel(w.scene, **w.oracle)
    _, g = _gpu_build(ctx, w.scene, 1, radix_bits=radix_bits, **w.build)
    _assert_top_level(g, og)
    _assert_leaf_level(g, og)


def test_build_is_deterministic_and_rebuild_matches(ctx, oracle):
    w = W.heightfield2m(grid=300, res_x=8, res_y=8)
    scene, g = _gpu_build(ctx, w.scene, 1, **w.build)
    first = {k: g.download(k) for k in ("top_keys", "top_vals", "leaf_keys", "leaf_vals", "leaves", "top_cells")}
    g.rebuild()
    for k, v in first.items():
        np.testing.assert_array_equal(g.download(k), v)
    # move the vertices (same topology); rebuild keeps the resolution
    res = g.info().res
    moved = w.scene.copy()
    moved[2] *= np.float32(1.1)
    moved[5] *= np.float32(1.1)
    moved[8] *= np.float32(1.1)
    scene.upload_vertices(moved)
    g.rebuild()
    og = oracle.build_twolevel(moved, res_override=res, **w.oracle)
    info = g.info()
    assert info.status == 0 and tuple(info.res) == tuple(og.res)
    _assert_top_level(g, og)
    _assert_leaf_level(g, og)


def test_edge_scenes(ctx, oracle):
    # one triangle
    one = np.array([[0.], [0.], [0.], [1.], [0.], [0.2], [0.], [1.], [0.5]], np.float32)
    for kind in (0, 1):
        og = (oracle.build_twolevel if kind else oracle.build_uniform)(one, lam=5.0, top_divisor=6 if kind else 1)
        _, g = _gpu_build(ctx, one, kind, lam=5.0)
        _assert_top_level(g, og)
        if kind:
            _assert_leaf_level(g, og)
    # degenerate (zero-area) triangles mixed in: plane test rejects them when they span cells
    w = sa.gen_soup(7, 3000, 0.05)
    w[3:6, ::7] = w[0:3, ::7]
    w[6:9, ::7] = w[0:3, ::7] + np.float32(0.2)
    og = oracle.build_twolevel(w, lam=5.0, top_divisor=6, lambda_leaf=2.0)
    _, g = _gpu_build(ctx, w, 1, lam=5.0, lambda_leaf=2.0)
    _assert_top_level(g, og)
    _assert_leaf_level(g, og)


def test_invalid_scenes_are_reported(ctx):
    flat = sa.gen_soup(3, 100, 0.01)
    flat[[2, 5, 8]] = 0.25      # zero extent in z
    scene = sa.FaceSoup(ctx, flat)
    b = sa.FastSimpleGridBuilder(ctx)
    b.init(scene)
    with pytest.raises(sa.ShinobiError) as e:
        b.build()
    assert e.value.status == 2


def test_rebuild_many_graph_matches_oracle(ctx, oracle):
    """sg_rebuild_many over six grids of both kinds, captured into a graph
    (concurrent branches) and replayed: after every replay each grid equals
    the oracle's build of its moved vertices at the grid's resolution, and
    the traversal after the join reads the rebuilt grids"""
    specs = [(0, 3_000, 81), (1, 40_000, 82), (0, 12_000, 83), (1, 2_000, 84), (1, 25_000, 85), (0, 80_000, 86)]
    items = []
    for kind, n, seed in specs:
        a = sa.gen_soup(seed, n, 0.01)
        kw = dict(lam=5.0, lambda_leaf=2.0) if kind else dict(lam=3.0)
        scene, g = _gpu_build(ctx, a, kind, **kw)
        moved = a.copy()
        for j in (2, 5, 8):
            moved[j] *= np.float32(1.1)
        items.append((kind, kw, scene, g, a, moved, tuple(g.info().res)))
    for kind, kw, scene, g, a, moved, res in items:
        scene.upload_vertices(moved)
    rays = sa.gen_random_rays(87, 20_000, 0.0, 1.0)
    rb = sa.RayBatch(ctx, rays)
    outs = [(sa.DeviceBuffer(ctx, 4 * rb.n), sa.DeviceBuffer(ctx, 4 * rb.n)) for _ in items]
    batch = sa.Batch(ctx, len(items))
    for it, (p, t) in zip(items, outs):
        batch.add(it[3], rb, p, t)
    batch.commit()
    ctx.sync()
    with sa.Graph.capture(ctx) as graph:
        batch.rebuild(branches=4)
        batch.launch()
    expect = []
    for kind, kw, scene, g, a, moved, res in items:
        if kind:
            og = oracle.build_twolevel(moved, top_divisor=6, res_override=res, **kw)
        else:
            og = oracle.build_uniform(moved, top_divisor=1, res_override=res, **kw)
        assert og.status == 0
        expect.append(og)
    for rep in range(2):
        # eager work between replays: grid 1 rebuilt from the original
        # vertices, the outputs poisoned
        items[1][2].upload_vertices(items[1][4])
        items[1][3].rebuild()
        items[1][2].upload_vertices(items[1][5])
        for p, _ in outs:
            p.upload(np.full(rb.n, 7, np.uint32))
        graph.launch()
        ctx.sync()
        for (kind, kw, scene, g, a, moved, res), og, (p, t) in zip(items, expect, outs):
            assert g.info().status == 0 and tuple(g.info().res) == res
            _assert_top_level(g, og)
            if kind:
                _assert_leaf_level(g, og)
            else:
                np.testing.assert_array_equal(g.download("indices"), og.indices)
            po, to = og.trace(rays)
            np.testing.assert_array_equal(p.download(np.uint32, rb.n), po, err_msg=f"replay {rep}")
            np.testing.assert_array_equal(t.download(np.float32, rb.n).view(np.uint32), to.view(np.uint32))
    graph.close()
    # sg_rebuild_many itself, eager, on 1 and on 4 branches (the original
    # vertices back in between): every grid again equals the oracle's
    for branches in (1, 4):
        for it in items:
            it[2].upload_vertices(it[4])
        sa.rebuild_many(ctx, [it[3] for it in items], branches=2)
        for it in items:
            it[2].upload_vertices(it[5])
        sa.rebuild_many(ctx, [it[3] for it in items], branches=branches)
        ctx.sync()
        for (kind, kw, scene, g, a, moved, res), og in zip(items, expect):
            _assert_top_level(g, og)
            if kind:
                _assert_leaf_level(g, og)
            else:
                np.testing.assert_array_equal(g.download("indices"), og.indices)
    # the argument checks
    sa.rebuild_many(ctx, [], branches=4)
    with pytest.raises(sa.ShinobiError):
        sa.rebuild_many(ctx, [items[0][3], items[0][3]])
    with pytest.raises(sa.ShinobiError):
        sa.rebuild_many(ctx, [items[0][3]], branches=17)
    ctx.sync()
    # a grid rebuilt from another scene (same triangle count) after the
    # commit: the batch refuses to rebuild until it is committed again, a
    # cleared batch rebuilds nothing, and after the new commit the grid is
    # rebuilt from the new scene
    kind, kw, scene, g, a, moved, res = items[0]
    assert kind == 0
    other = sa.gen_soup(91, a.shape[1], 0.01)
    oscene = sa.FaceSoup(ctx, other)
    sa.check(ctx._l.sg_rebuild(ctx.h, oscene.h, g.h, ctx.stream), "sg_rebuild")
    with pytest.raises(sa.ShinobiError):
        batch.rebuild(branches=4)
    batch.clear()
    batch.rebuild(branches=4)
    # a cleared batch launches nothing: the outputs keep what they hold
    outs[0][0].upload(np.full(rb.n, 5, np.uint32))
    batch.launch()
    ctx.sync()
    assert (outs[0][0].download(np.uint32, rb.n) == 5).all()
    for it, (p, t) in zip(items, outs):
        batch.add(it[3], rb, p, t)
    batch.commit()
    oscene.upload_vertices(other * np.float32(0.9))
    batch.rebuild(branches=4)
    ctx.sync()
    og = oracle.build_uniform(other * np.float32(0.9), top_divisor=1, res_override=res, **kw)
    assert og.status == 0
    _assert_top_level(g, og)
    np.testing.assert_array_equal(g.download("indices"), og.indices)
    batch.clear()
