"""The quadtree built on the GPU (bfhipQuadtreeInitWithLeafSize, GPU).

The reference builds its quadtree by recursive in-place sifts
(quadtree_node.c:123-306) with a leaf threshold of 1 (quadtree_node.c:17).
The device builder computes each point's quadrant path as a Morton key,
radix-sorts the keys and emits the nodes level by level.  Checked here:
at leaf size 1 the whole tree -- every node record byte for byte (boxes,
splits, offsets, links, depths), the level table and the permutation --
equals the host builder's, and the permutation equals the oracle's
restatement of the sift; with larger leaves the node records are still
byte-equal and every leaf holds the same points (in key order); a
factorization on device-built trees applies to the accuracy of one on
host-built trees.  Inputs include points lying exactly on split lines
(ties go low) and a 2^16-point curve, where the host builder stitches
subtrees built on threads.
"""
import ctypes as C
import time

import numpy as np
import pytest

import oracle as orc
from conftest import box_problem, curve_problem

pytestmark = pytest.mark.gpu


def node_bytes(bf, t):
    q = bf.QuadtreeStruct.from_address(t.ptr)
    return C.string_at(q.nodes, q.numNodes * C.sizeof(bf.QuadtreeNodeStruct))


def level_table(bf, t):
    q = bf.QuadtreeStruct.from_address(t.ptr)
    return [q.levelOffset[d] for d in range(q.numLevels + 1)]


def point_sets():
    rng = np.random.default_rng(7)
    g = np.arange(9) / 8.0
    grid = np.stack(np.meshgrid(g, g), -1).reshape(-1, 2)  # every split line of [0,1]^2 carries points
    return {
        "box": box_problem(5000, seed=3)[0],
        "curve": curve_problem(4096)[0],
        "grid_ties": grid,
        "uniform": rng.uniform(-2, 3, (3001, 2)),
        "curve_2^16": curve_problem(1 << 16)[0],
    }


@pytest.mark.parametrize("name", ["box", "curve", "grid_ties", "uniform", "curve_2^16"])
def test_device_tree_equals_host_tree_at_leaf_one(gpu_lib, name):
    bf = gpu_lib
    pts = point_sets()[name]
    th, td = bf.Quadtree(pts, 1), bf.Quadtree(pts, 1, device=True)
    assert td.num_nodes == th.num_nodes
    assert level_table(bf, td) == level_table(bf, th)
    assert node_bytes(bf, td) == node_bytes(bf, th)
    assert np.array_equal(td.perm, th.perm)
    if len(pts) <= 5000:
        perm, nodes, _ = orc.tree(pts, 1)
        assert np.array_equal(td.perm, perm)
        assert td.num_nodes == len(nodes)


@pytest.mark.parametrize("leaf", [8, 64])
def test_device_tree_has_the_host_nodes_and_leaf_sets(gpu_lib, leaf):
    bf = gpu_lib
    for name, pts in point_sets().items():
        th, td = bf.Quadtree(pts, leaf), bf.Quadtree(pts, leaf, device=True)
        assert node_bytes(bf, td) == node_bytes(bf, th), name
        assert level_table(bf, td) == level_table(bf, th), name
        ph, pd = th.perm, td.perm
        assert np.array_equal(np.sort(pd), np.arange(len(pts)))
        leaves = 0
        for i in range(th.num_nodes):
            nd = th.node(i)
            if all(c < 0 for c in nd.child):
                a, b = nd.offset[0], nd.offset[4]
                assert np.array_equal(np.sort(pd[a:b]), np.sort(ph[a:b])), (name, i)
                leaves += 1
        assert leaves > 0


def test_device_trees_factor_like_host_trees(gpu_lib):
    bf = gpu_lib
    n = 8192
    src, tgt, k = curve_problem(n)
    opts = bf.options(tol=1e-10, rank_const=1.25, pass_through=True, pinv_rcond=1e-11)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    rows = np.arange(0, n, 41)
    yd = orc.dense_rows(k, tgt[rows], src, x)
    ys = {}
    for dev in (False, True):
        ts, tt = bf.Quadtree(src, 64, device=dev), bf.Quadtree(tgt, 64, device=dev)
        F = bf.FacHelm2.make_multilevel(bf.Helm2(k), ts, tt, opts)
        y = np.empty(n, complex)
        y[tt.perm] = F @ x[ts.perm]
        ys[dev] = y
        err = np.linalg.norm(y[rows] - yd) / np.linalg.norm(yd)
        print("device trees" if dev else "host trees", "vs dense rows:", err)
        assert err < 1e-8
    assert np.linalg.norm(ys[True] - ys[False]) / np.linalg.norm(ys[False]) < 1e-9


def test_device_tree_at_2_20_points(gpu_lib):
    bf = gpu_lib
    pts = curve_problem(1 << 20)[0]
    bf.Quadtree(pts[:4096], 64, device=True)  # first use: module load
    t0 = time.perf_counter()
    td = bf.Quadtree(pts, 64, device=True)
    t1 = time.perf_counter()
    th = bf.Quadtree(pts, 64)
    t2 = time.perf_counter()
    print(f"2^20 points, leaf 64: device {1e3 * (t1 - t0):.1f} ms, host {1e3 * (t2 - t1):.1f} ms, "
          f"{td.num_nodes} nodes")
    assert node_bytes(bf, td) == node_bytes(bf, th)
    t3 = time.perf_counter()
    t1d = bf.Quadtree(pts, 1, device=True)
    t4 = time.perf_counter()
    t1h = bf.Quadtree(pts, 1)
    t5 = time.perf_counter()
    print(f"2^20 points, leaf 1: device {1e3 * (t4 - t3):.1f} ms, host {1e3 * (t5 - t4):.1f} ms, "
          f"{t1d.num_nodes} nodes")
    assert np.array_equal(t1d.perm, t1h.perm)
    assert node_bytes(bf, t1d) == node_bytes(bf, t1h)


def test_device_tree_rejects_bad_input(gpu_lib):
    bf = gpu_lib
    with pytest.raises(bf.BfError):
        bf.Quadtree(np.zeros((1, 2)), 1, device=True)
    with pytest.raises(bf.BfError):
        bf.Quadtree(np.random.default_rng(0).uniform(size=(10, 2)), 0, device=True)
