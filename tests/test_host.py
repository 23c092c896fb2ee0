"""Host-side checks of the product library that need no GPU: it loads and
exports what the header declares, its RNG, quadtree and planner agree with
the oracle integer-for-integer, and the device H0 source compiled for the
host matches mpmath."""
import ctypes as C
import json
import os
import re
import subprocess

import numpy as np
import pytest

import butterfly_amd as bf
import oracle as orc
from conftest import ROOT, box_problem, curve_problem

HEADER = os.path.join(ROOT, "include", "bf_amd", "bf.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(bf[A-Za-z0-9_]+)\s*\(", text))
    return sorted(n for n in names if not n.startswith("bfk_"))


def test_library_exports_every_declared_symbol():
    L = bf.lib()
    names = declared_functions()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_error_state_is_read_and_cleared():
    L = bf.lib()
    L.bfSetError(3)
    assert L.bfGetError() == 3
    assert L.bfGetError() == 0


def test_rng_matches_oracle_bit_for_bit():
    for s in (0, 1, 12345):
        bf.seed(s)
        orc.seed(s)
        a = bf.real_uniform(1000)
        b = np.zeros(1000)
        orc.lib().orc_uniform_fill(1000, orc._dp(b))
        assert np.array_equal(a, b)
        bf.seed(s)
        orc.seed(s)
        assert np.array_equal(bf.complex_randn(777), orc.complex_randn(777))


@pytest.mark.parametrize("leaf", [1, 7, 64])
def test_quadtree_perm_matches_oracle(leaf):
    src, tgt, _ = box_problem(5000, seed=3)
    for pts in (src, tgt, curve_problem(4096)[0]):
        t = bf.Quadtree(pts, leaf)
        perm, nodes, _ = orc.tree(pts, leaf)
        assert np.array_equal(t.perm, perm)
        assert t.num_nodes == len(nodes)


@pytest.mark.parametrize("leaf", [16, 64])
def test_quadtree_threaded_build_matches_oracle(leaf):
    """From 2^16 points on, the subtrees below the second level are built on
    threads and stitched: same permutation and node count as the serial
    oracle."""
    rng = np.random.default_rng(leaf)
    for pts in (curve_problem(1 << 17)[0], rng.uniform(-1, 1, ((1 << 16) + 3, 2))):
        t = bf.Quadtree(pts, leaf)
        perm, nodes, _ = orc.tree(pts, leaf)
        assert np.array_equal(t.perm, perm)
        assert t.num_nodes == len(nodes)


def test_quadtree_rejects_bad_input():
    with pytest.raises(bf.BfError):
        bf.Quadtree(np.zeros((1, 2)), 1)


def _plan_pair(src, tgt, k, leaf, tol, pt):
    ts, tt = bf.Quadtree(src, leaf), bf.Quadtree(tgt, leaf)
    F = bf.FacHelm2.plan_only(bf.Helm2(k), ts, tt, bf.options(tol=tol, pass_through=pt))
    O = orc.Fac(k, src, tgt, leaf, tol, 128 * 128, pt, plan_only=True)
    return F, O


@pytest.mark.parametrize("case", ["box_ref", "box_pass", "curve_pass", "curve_ref", "curve_pass_c125", "self_pass_c125"])
def test_plan_matches_oracle(case):
    if case.startswith("box"):
        src, tgt, k = box_problem(4096)
    elif case.startswith("self"):
        # the curve against itself: near pairs across narrow gaps
        src, _, k = curve_problem(4096)
        tgt = src
    else:
        src, tgt, k = curve_problem(4096)
    leaf, tol, pt = (1, 1e-15, False) if case.endswith("ref") else (64, 1e-8, True)
    C = 1.25 if case.endswith("c125") else 1.0
    ts, tt = bf.Quadtree(src, leaf), bf.Quadtree(tgt, leaf)
    F = bf.FacHelm2.plan_only(bf.Helm2(k), ts, tt, bf.options(tol=tol, pass_through=pt, rank_const=C))
    O = orc.Fac(k, src, tgt, leaf, tol, 128 * 128, pt, plan_only=True, rank_const=C)
    for a, b in zip(F.plan(), O.plan()):
        assert a.shape == b.shape and np.array_equal(a, b)


def _stored_per_leaf(lt, ft, bt):
    """Stored entries of every top-level leaf (identity blocks store none)."""
    out, f, b = [], 0, 0
    for row in lt:
        if not row[0]:
            out.append(row[3] * row[4])
            continue
        st = 0
        for _ in range(row[5]):
            nb = ft[f, 3]
            blk = bt[b:b + nb]
            st += int((blk[:, 2] * blk[:, 3])[blk[:, 4] != 2].sum())
            b += nb
            f += 1
        out.append(st)
    return np.array(out)


def test_pass_through_keeps_only_compressing_butterflies():
    """The pass-through planner replaces a butterfly that would store more
    than its dense block (on the curve against itself a few such pairs
    occur at N = 16384) by the dense block; no leaf stores more than m n."""
    src, _, k = curve_problem(16384)
    t = bf.Quadtree(src, 64)
    F = bf.FacHelm2.plan_only(bf.Helm2(k), t, t, bf.options(tol=1e-8, pass_through=True, rank_const=1.25))
    lt, ft, bt = F.plan()
    st = _stored_per_leaf(lt, ft, bt)
    assert (st <= lt[:, 3] * lt[:, 4]).all()
    lsq = bt[bt[:, 4] == 0]
    assert lsq[:, 2].max() < 128
    O = orc.Fac(k, src, src, 64, 1e-8, 128 * 128, True, plan_only=True, rank_const=1.25)
    for a, b in zip((lt, ft, bt), O.plan()):
        assert a.shape == b.shape and np.array_equal(a, b)


def test_plan_config0_census():
    # config 0 under the reference rules: 256 butterflies of 4 factors,
    # 4096 + 32768 + 4096 blocks, ranks 16 everywhere
    src, tgt, k = box_problem(16384)
    ts, tt = bf.Quadtree(src, 1), bf.Quadtree(tgt, 1)
    F = bf.FacHelm2.plan_only(bf.Helm2(k), ts, tt)
    lt, ft, bt = F.plan()
    assert len(lt) == 256 and lt[:, 0].all() and (lt[:, 5] == 4).all()
    assert len(bt) == 4096 + 32768 + 4096
    lsq = bt[bt[:, 4] == 0]
    assert (lsq[:, 2] == 16).all()


def _census(lt, ft, bt):
    """(first, middle, last) block counts, stored entries and the ranks of
    the re-expansion blocks of a plan export."""
    f = b = 0
    first = mid = last = stored = 0
    ps = []
    for leaf in lt:
        if leaf[0] == 0:
            stored += int(leaf[3] * leaf[4])
            continue
        nf = int(leaf[5])
        for i in range(nf):
            nb = int(ft[f, 3])
            blk = bt[b:b + nb]
            if i == 0:
                first += nb
            elif i == nf - 1:
                last += nb
            else:
                mid += nb
            stored += int((blk[:, 2] * blk[:, 3])[blk[:, 4] != 2].sum())
            ps += list(blk[blk[:, 4] == 0][:, 2])
            b += nb
            f += 1
    return first, mid, last, stored, np.array(ps)


def test_plan_curve_census_reference_rules():
    """Curve -> circle at N = 2^16, k for 10 points per wavelength, under the
    reference rules (leaf 1, rank estimates at 1e-15, dense below 128^2,
    bfFacHelm2Prepare's level choice, fac_helm2.c:551-651): 188 butterflies
    and 8 dense blocks; 1170 first, 7816 middle and 1374 last blocks;
    1.777e9 stored entries; 8986 re-expansion blocks with p from 31 (median
    100) to 6252.  These counts come from an independent restatement of the
    planning rules; a plan one level off changes every one of them.  Both
    the oracle and the product's planner must give them."""
    src, tgt, k = curve_problem(1 << 16)
    O = orc.Fac(k, src, tgt, plan_only=True)
    ts, tt = bf.Quadtree(src, 1), bf.Quadtree(tgt, 1)
    F = bf.FacHelm2.plan_only(bf.Helm2(k), ts, tt)
    for lt, ft, bt in (O.plan(), F.plan()):
        assert len(lt) == 196 and int(lt[:, 0].sum()) == 188
        first, mid, last, stored, ps = _census(lt, ft, bt)
        assert (first, mid, last) == (1170, 7816, 1374)
        assert round(stored / 1e6) == 1777
        assert len(ps) == 8986 and ps.min() == 31 and int(np.median(ps)) == 100 and ps.max() == 6252
    for a, b in zip(O.plan(), F.plan()):
        assert np.array_equal(a, b)


def test_plan_covers_every_entry_once():
    # the top-level leaves tile the N_tgt x N_src index square exactly
    src, tgt, k = curve_problem(8192)
    ts, tt = bf.Quadtree(src, 64), bf.Quadtree(tgt, 64)
    F = bf.FacHelm2.plan_only(bf.Helm2(k), ts, tt, bf.options(tol=1e-8, pass_through=True))
    lt, _, _ = F.plan()
    assert (lt[:, 3] * lt[:, 4]).sum() == len(src) * len(tgt)
    # no two leaves overlap: sort by (row0, col0) and check each row band
    # of columns is a disjoint cover
    for r0 in np.unique(lt[:, 1]):
        sel = lt[lt[:, 1] == r0]
        assert len(np.unique(sel[:, 2])) == len(sel)


def test_device_h0_source_on_host(tmp_path):
    # compile the device H0 for the host and compare with the mpmath table
    shim = tmp_path / "h0shim.cpp"
    shim.write_text('#include "bf_h0.h"\nextern "C" void h0v(int n, const double *x, double *o)'
                    '{for(int i=0;i<n;i++) bf_h0(x[i], &o[2*i], &o[2*i+1]);}\n')
    so = tmp_path / "h0shim.so"
    subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-I",
                    os.path.join(ROOT, "butterfly_amd", "csrc"), str(shim), "-o", str(so)], check=True)
    L = C.CDLL(str(so))
    tab = json.load(open(os.path.join(ROOT, "tests", "golden", "h0_table.json")))["rows"]
    x = np.array([float(r[0]) for r in tab])
    want = np.array([complex(float(r[1]), float(r[2])) for r in tab])
    out = np.zeros(2 * len(x))
    L.h0v(len(x), x.ctypes.data_as(C.POINTER(C.c_double)), out.ctypes.data_as(C.POINTER(C.c_double)))
    got = out.view(np.complex128)
    err = np.abs(got - want) / np.abs(want)
    assert err.max() < 2e-15, (x[np.argmax(err)], err.max())


def test_device_h1_source_on_host(tmp_path):
    """The device H1 (double layer) compiled for the host against the mpmath
    table, and the double-layer kernel value against its closed form."""
    shim = tmp_path / "h1shim.cpp"
    shim.write_text('#include "bf_h0.h"\nextern "C" void h1v(int n, const double *x, double *o)'
                    '{for(int i=0;i<n;i++) bf_h1(x[i], &o[2*i], &o[2*i+1]);}\n'
                    'extern "C" void dlv(double k, double r, double dot, double *o)'
                    '{bf_helm2_double(k, r, dot, &o[0], &o[1]);}\n')
    so = tmp_path / "h1shim.so"
    subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-I",
                    os.path.join(ROOT, "butterfly_amd", "csrc"), str(shim), "-o", str(so)], check=True)
    L = C.CDLL(str(so))
    tab = json.load(open(os.path.join(ROOT, "tests", "golden", "h1_table.json")))["rows"]
    x = np.array([float(r[0]) for r in tab])
    want = np.array([complex(float(r[1]), float(r[2])) for r in tab])
    out = np.zeros(2 * len(x))
    L.h1v(len(x), x.ctypes.data_as(C.POINTER(C.c_double)), out.ctypes.data_as(C.POINTER(C.c_double)))
    got = out.view(np.complex128)
    err = np.abs(got - want) / np.abs(want)
    assert err.max() < 3e-15, (x[np.argmax(err)], err.max())
    # (i k / 4) H1(k r) dot / r, and 0 below r = 1e-15
    from scipy.special import hankel1
    o = np.zeros(2)
    L.dlv.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]
    for k, r, dot in [(30.0, 0.2, 0.13), (7.0, 2.5, -1.7), (1000.0, 0.01, 0.004)]:
        L.dlv(k, r, dot, o.ctypes.data_as(C.POINTER(C.c_double)))
        ref = 0.25j * k * hankel1(1, k * r) * dot / r
        assert abs(complex(o[0], o[1]) - ref) < 1e-13 * abs(ref)
    L.dlv(30.0, 1e-16, 1.0, o.ctypes.data_as(C.POINTER(C.c_double)))
    assert o[0] == 0 and o[1] == 0


def test_rank_features_closed_form():
    """bfhipRankFeatures (host code): log diameters, k times the larger
    diameter, centre distance over it, depth, log10 tol."""
    import butterfly_amd as bf
    f = bf.rank_features([0, 0, 1, 1], [2, 0, 4, 2], 10.0, 3, 1e-8)
    dc, dr = np.sqrt(2), np.sqrt(8)
    want = [np.log(dc), np.log(dr), 10 * dr, np.hypot(2.5, 0.5) / dr, 3, -8]
    assert np.allclose(f, want, rtol=1e-6)


def _iter(L, tree, node, traversal):
    it = bf.TreeLevelIterStruct()
    L.bfTreeLevelIterInit(C.byref(it), traversal, tree.node_ptr(node))
    return it


def _level_nodes(it, tree):
    lv = it.levelNodes.contents
    base = bf.QuadtreeStruct.from_address(tree.ptr).nodes
    size = C.sizeof(bf.QuadtreeNodeStruct)
    return [(lv.data[i] - base) // size for i in range(lv.num_elts)]


def test_tree_level_iterator_walks_levels_both_ways():
    """bfTreeLevelIter (tree_level_iter.c:10-330) over a subtree: level order
    from the node down (IsDone after the deepest level, whose view stays),
    reverse level order from the deepest level up (the view stays on the
    node's level); the levels are the subtree's nodes by depth, children in
    quadrant order; the point counts and the internal test per level."""
    L = bf.lib()
    src, _, _ = box_problem(3000)
    t = bf.Quadtree(src, 1)
    root = 5  # a depth-2 node
    nodes, depth = [root], {root: 0}
    for v in nodes:
        for c in t.node(v).child:
            if c >= 0:
                nodes.append(c)
                depth[c] = depth[v] + 1
    D = max(depth.values())
    levels = [[v for v in nodes if depth[v] == d] for d in range(D + 1)]
    it = _iter(L, t, root, 1)
    for d in range(D + 1):
        assert not L.bfTreeLevelIterIsDone(C.byref(it))
        assert L.bfTreeLevelIterCurrentDepth(C.byref(it)) == t.node(root).depth + d
        assert _level_nodes(it, t) == levels[d]
        npts = sum(t.node(v).offset[4] - t.node(v).offset[0] for v in levels[d])
        assert L.bfTreeLevelIterGetNumPoints(C.byref(it)) == npts
        internal = all(any(c >= 0 for c in t.node(v).child) for v in levels[d])
        assert L.bfTreeLevelIterCurrentLevelIsInternal(C.byref(it)) == internal
        L.bfTreeLevelIterNext(C.byref(it))
    assert L.bfTreeLevelIterIsDone(C.byref(it)) and _level_nodes(it, t) == levels[D]
    L.bfTreeLevelIterDeinit(C.byref(it))
    it = _iter(L, t, root, 2)
    for d in range(D, -1, -1):
        assert _level_nodes(it, t) == levels[d]
        L.bfTreeLevelIterNext(C.byref(it))
    assert _level_nodes(it, t) == levels[0]
    L.bfTreeLevelIterDeinit(C.byref(it))
    assert L.bfQuadtreeNodeGetQuadtreeConst(t.node_ptr(root)) == t.ptr


def test_prepare_gives_the_planners_factor_counts():
    """bfFacHelm2Prepare (fac_helm2.c:551-651) on every separated pair of the
    curve -> circle plan at N = 2^14 under the reference rules returns the
    factor count of the planner's leaf for that pair (the planner is pinned
    by the census above), and leaves the source iterator on the level the
    first factor uses."""
    L = bf.lib()
    src, tgt, k = curve_problem(1 << 14)
    ts, tt = bf.Quadtree(src, 1), bf.Quadtree(tgt, 1)
    helm = bf.Helm2(k)
    lt, _, _ = bf.FacHelm2.plan_only(helm, ts, tt).plan()

    def index(tree):
        out = {}
        for i in range(tree.num_nodes):
            nd = tree.node(i)
            out.setdefault((nd.offset[0], nd.offset[4] - nd.offset[0], nd.depth), i)
        return out

    si, ti = index(ts), index(tt)
    checked = 0
    for leaf in lt[lt[:, 0] == 1]:
        _, row0, col0, m, n, nf = (int(v) for v in leaf)
        for d in range(2, 30):
            if (col0, n, d) in si and (row0, m, d) in ti:
                s, t = si[(col0, n, d)], ti[(row0, m, d)]
                break
        a, b = bf.TreeLevelIterStruct(), bf.TreeLevelIterStruct()
        got = L.bfFacHelm2Prepare(C.byref(helm.s), ts.node_ptr(s), tt.node_ptr(t), C.byref(a), C.byref(b))
        assert got == nf, (s, t, got, nf)
        assert L.bfTreeLevelIterCurrentDepth(C.byref(a)) == ts.node(s).depth + nf - 2
        assert L.bfTreeLevelIterCurrentDepth(C.byref(b)) == tt.node(t).depth
        L.bfTreeLevelIterDeinit(C.byref(a))
        L.bfTreeLevelIterDeinit(C.byref(b))
        checked += 1
    assert checked == int(lt[:, 0].sum()) and checked > 50


def test_c_caller_compiles_against_the_header(tmp_path):
    """A C11 program written against include/bf_amd/bf.h alone (the reference
    drivers' calls: quadtree, multilevel and single-level factorization,
    bfMatMul, the containers' struct fields) compiles warning-free and
    links against the library; the GPU test runs it."""
    exe = tmp_path / "bf_caller"
    r = subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "c_caller", "bf_caller.c"), "-o", str(exe),
                        "-L", os.path.join(ROOT, "butterfly_amd"), "-lbutterfly_amd",
                        "-Wl,-rpath," + os.path.join(ROOT, "butterfly_amd"), "-lm"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
