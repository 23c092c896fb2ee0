"""The reference's block matrix algebra at the boundary: BfMatProduct,
BfMatBlockDiag, BfMatBlockCoo, BfMatBlockDense (mat_product.c,
mat_block_*.c).

CPU: construction, shapes, offsets and ownership through the C API (no
device work).  GPU: bfMatMul on containers of dense leaves (packed into a
device product and applied by the stage kernels) against numpy's dense
products of the same matrices, a product mixing a device butterfly with a
dense factor (applied factor by factor), and bfhipMatPack.  FP64
throughout: the device sums each row in a different order than numpy, so
results agree to a few ulps of the row norms (1e-13 relative).
"""
import ctypes as C

import numpy as np
import pytest

import butterfly_amd as bf

vp, sz = C.c_void_p, C.c_size_t


class MatS(C.Structure):
    _fields_ = [("vtbl", vp), ("props", C.c_int), ("numRows", sz), ("numCols", sz)]


class MatBlockS(C.Structure):
    _fields_ = [("super", MatS), ("vtbl", vp), ("block", C.POINTER(vp)), ("rowOffset", C.POINTER(sz)),
                ("colOffset", C.POINTER(sz))]


class MatBlockCooS(C.Structure):
    _fields_ = [("super", MatBlockS), ("numBlocks", sz), ("rowInd", C.POINTER(sz)), ("colInd", C.POINTER(sz)),
                ("numBlockRows", sz), ("numBlockCols", sz), ("ownsBlock", C.POINTER(C.c_bool))]


class PtrArrayS(C.Structure):
    _fields_ = [("data", C.POINTER(vp)), ("capacity", sz), ("num_elts", sz), ("isView", C.c_bool)]


def L():
    lib = bf.lib()
    for name, res, args in [
        ("bfMatDenseComplexNewViewFromPtr", vp, [sz, sz, vp]),
        ("bfMatBlockDiagNew", vp, []), ("bfMatBlockDiagInit", None, [vp, sz, sz]),
        ("bfMatBlockDiagSetBlock", None, [vp, sz, vp]), ("bfMatBlockDiagGetBlock", vp, [vp, sz]),
        ("bfMatBlockDiagNumBlocks", sz, [vp]), ("bfMatBlockDiagNewFromBlocks", vp, [vp, C.c_int]),
        ("bfMatBlockDenseNew", vp, []), ("bfMatBlockDenseInit", None, [vp, sz, sz]),
        ("bfMatBlockDenseSetBlock", None, [vp, sz, sz, vp]),
        ("bfMatBlockCooNew", vp, []), ("bfMatBlockCooInit", None, [vp, sz, sz, sz]),
        ("bfMatProductNew", vp, []), ("bfMatProductInit", None, [vp]),
        ("bfMatProductPostMultiply", None, [vp, vp]), ("bfMatProductNumFactors", sz, [vp]),
        ("bfMatProductPopLastFactor", vp, [vp]), ("bfMatProductGetFactor", vp, [vp, sz]),
        ("bfInitPtrArray", None, [vp, sz]), ("bfPtrArrayAppend", None, [vp, vp]),
        ("bfPtrArrayDeinit", None, [vp]), ("bfhipMatPack", vp, [vp]),
    ]:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


class Dense:
    """A BfMatDenseComplex view over a numpy array that stays alive."""

    def __init__(self, A):
        self.A = np.ascontiguousarray(A, np.complex128)
        self.ptr = L().bfMatDenseComplexNewViewFromPtr(self.A.shape[0], self.A.shape[1], self.A.ctypes.data)


def shape(p):
    return L().bfMatGetNumRows(p), L().bfMatGetNumCols(p)


def rand(rng, m, n):
    return rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))


def block_diag(blocks):
    m = sum(b.shape[0] for b in blocks)
    n = sum(b.shape[1] for b in blocks)
    D = np.zeros((m, n), complex)
    i = j = 0
    for b in blocks:
        D[i:i + b.shape[0], j:j + b.shape[1]] = b
        i += b.shape[0]
        j += b.shape[1]
    return D


def make_diag(rng, sizes, keep):
    lib = L()
    d = lib.bfMatBlockDiagNew()
    lib.bfMatBlockDiagInit(d, len(sizes), len(sizes))
    mats = []
    for i, (m, n) in enumerate(sizes):
        A = rand(rng, m, n)
        dn = Dense(A)
        keep.append(dn)
        lib.bfMatBlockDiagSetBlock(d, i, dn.ptr)  # the container owns the view struct
        mats.append(A)
    return d, block_diag(mats)


def make_coo(rng, rowsz, colsz, where, keep):
    """a block-COO matrix filled through its struct fields, as the
    reference's drivers do"""
    lib = L()
    o = lib.bfMatBlockCooNew()
    lib.bfMatBlockCooInit(o, len(rowsz), len(colsz), len(where))
    s = MatBlockCooS.from_address(o)
    ro, co = np.r_[0, np.cumsum(rowsz)], np.r_[0, np.cumsum(colsz)]
    for i, v in enumerate(ro):
        s.super.rowOffset[i] = int(v)
    for i, v in enumerate(co):
        s.super.colOffset[i] = int(v)
    D = np.zeros((ro[-1], co[-1]), complex)
    for b, (r, c) in enumerate(where):
        A = rand(rng, rowsz[r], colsz[c])
        dn = Dense(A)
        keep.append(dn)
        s.rowInd[b], s.colInd[b] = r, c
        s.super.block[b] = dn.ptr
        D[ro[r]:ro[r + 1], co[c]:co[c + 1]] += A
    return o, D


def test_block_containers_shapes_and_ownership():
    """Construction only (no device work)."""
    lib = L()
    rng = np.random.default_rng(1)
    keep = []
    d, D = make_diag(rng, [(5, 7), (4, 4), (6, 3)], keep)
    assert shape(d) == D.shape and lib.bfMatGetType(d) == 17 and lib.bfMatBlockDiagNumBlocks(d) == 3
    assert lib.bfMatBlockDiagGetBlock(d, 1) == keep[1].ptr
    lib.bfMatBlockDiagGetBlock(d, 3)
    assert lib.bfGetError() == 1  # BF_ERROR_INVALID_ARGUMENTS
    o, Do = make_coo(rng, [3, 4, 2], [2, 5], [(0, 1), (2, 0), (1, 1)], keep)
    assert shape(o) == Do.shape and lib.bfMatGetType(o) == 15
    p = lib.bfMatProductNew()
    lib.bfMatProductInit(p)
    lib.bfMatProductPostMultiply(p, d)          # 15 x 14
    e = lib.bfMatBlockDenseNew()
    lib.bfMatBlockDenseInit(e, 2, 1)
    A0, A1 = rand(rng, 6, 9), rand(rng, 8, 9)
    for i, A in enumerate((A0, A1)):
        dn = Dense(A)
        keep.append(dn)
        lib.bfMatBlockDenseSetBlock(e, i, 0, dn.ptr)
    assert shape(e) == (14, 9)
    lib.bfMatProductPostMultiply(p, e)
    assert shape(p) == (15, 9) and lib.bfMatProductNumFactors(p) == 2 and lib.bfMatGetType(p) == 10
    bad = Dense(rand(rng, 3, 3))
    lib.bfMatProductPostMultiply(p, bad.ptr)     # 9 != 3 rows
    assert lib.bfGetError() == 8                 # BF_ERROR_INCOMPATIBLE_SHAPES
    last = lib.bfMatProductPopLastFactor(p)
    assert last == e and shape(p) == (15, 14)
    lib.bfMatProductPostMultiply(p, e)
    # a view container over caller-owned blocks leaves them alone when deleted
    arr = PtrArrayS()
    lib.bfInitPtrArray(C.byref(arr), 2)
    views = [Dense(rand(rng, 2, 3)), Dense(rand(rng, 4, 1))]
    for v in views:
        lib.bfPtrArrayAppend(C.byref(arr), v.ptr)
    dv = lib.bfMatBlockDiagNewFromBlocks(C.byref(arr), 0)  # BF_POLICY_VIEW
    assert shape(dv) == (6, 4)
    lib.bfMatDelete(C.byref(C.c_void_p(dv)))
    assert lib.bfMatGetNumRows(views[0].ptr) == 2    # still alive
    for v in views:
        lib.bfMatDelete(C.byref(C.c_void_p(v.ptr)))
    lib.bfPtrArrayDeinit(C.byref(arr))
    lib.bfMatDelete(C.byref(C.c_void_p(p)))          # deletes d, e and their blocks
    lib.bfMatDelete(C.byref(C.c_void_p(o)))
    lib.bfMatDelete(C.byref(C.c_void_p(bad.ptr)))
    assert lib.bfGetError() == 0


def mul(p, X):
    lib = L()
    X = np.ascontiguousarray(X, np.complex128)
    xm = Dense(X)
    y = lib.bfMatMul(p, xm.ptr)
    assert y, lib.bfGetError()
    out = bf.Mat(y).to_numpy()
    lib.bfMatDelete(C.byref(C.c_void_p(xm.ptr)))
    return out


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.gpu
def test_block_containers_apply_on_the_device(gpu_lib):
    lib = L()
    rng = np.random.default_rng(2)
    keep = []
    d, D = make_diag(rng, [(5, 7), (4, 4), (6, 3)], keep)        # 15 x 14
    X = rand(rng, 14, 3)
    assert rel(mul(d, X), D @ X) < 1e-13
    o, Do = make_coo(rng, [3, 4, 2], [2, 5, 7], [(0, 1), (2, 0), (1, 1), (1, 2), (0, 2)], keep)  # 9 x 14
    assert rel(mul(o, X), Do @ X) < 1e-13
    e = lib.bfMatBlockDenseNew()
    lib.bfMatBlockDenseInit(e, 2, 2)
    De = np.zeros((14, 9), complex)
    for i, (r0, r1) in enumerate([(0, 6), (6, 14)]):
        for j, (c0, c1) in enumerate([(0, 5), (5, 9)]):
            A = rand(rng, r1 - r0, c1 - c0)
            dn = Dense(A)
            keep.append(dn)
            lib.bfMatBlockDenseSetBlock(e, i, j, dn.ptr)
            De[r0:r1, c0:c1] = A
    X9 = rand(rng, 9, 2)
    assert rel(mul(e, X9), De @ X9) < 1e-13
    # the product diag * dense-blocks (15 x 9), then times coo^T-shaped data
    p = lib.bfMatProductNew()
    lib.bfMatProductInit(p)
    lib.bfMatProductPostMultiply(p, d)
    lib.bfMatProductPostMultiply(p, e)
    assert rel(mul(p, X9), D @ (De @ X9)) < 1e-13
    P = lib.bfhipMatPack(p)
    assert P and lib.bfMatGetType(P) == 45  # BF_TYPE_MAT_HIP_PRODUCT, after the reference's BF_TYPE_NONE = 43
    for _ in range(2):
        assert rel(mul(P, X9), D @ (De @ X9)) < 1e-13
    lib.bfMatDelete(C.byref(C.c_void_p(P)))
    lib.bfMatDelete(C.byref(C.c_void_p(p)))
    lib.bfMatDelete(C.byref(C.c_void_p(o)))


@pytest.mark.gpu
def test_product_with_a_device_factorization(gpu_lib):
    """[dense W] * [butterfly]: applied factor by factor (the butterfly
    by its own apply) against the dense product."""
    import oracle as orc
    from conftest import box_problem
    lib = L()
    rng = np.random.default_rng(3)
    src, tgt, k = box_problem(512)
    ts, tt = bf.Quadtree(src, 32), bf.Quadtree(tgt, 32)
    F = bf.FacHelm2.make_multilevel(bf.Helm2(k), ts, tt, bf.options(tol=1e-12, pass_through=True))
    K = orc.kernel_matrix(k, tgt[tt.perm], src[ts.perm])
    W = rand(rng, 7, 512)
    wm = Dense(W)
    p = lib.bfMatProductNew()
    lib.bfMatProductInit(p)
    lib.bfMatProductPostMultiply(p, wm.ptr)
    lib.bfMatProductPostMultiply(p, F.ptr)
    X = rand(rng, 512, 2)
    assert rel(mul(p, X), W @ (K @ X)) < 1e-9
    # the product owns its factors; detach them before it goes
    lib.bfMatProductPopLastFactor(p)
    lib.bfMatProductPopLastFactor(p)
    lib.bfMatDelete(C.byref(C.c_void_p(p)))
    lib.bfMatDelete(C.byref(C.c_void_p(wm.ptr)))


class VecComplexS(C.Structure):
    _fields_ = [("vtbl", vp), ("props", C.c_int), ("size", sz), ("stride", sz), ("data", C.POINTER(C.c_double))]


def vec_lib():
    lib = L()
    for name, res, args in [("bfVecComplexNew", vp, []), ("bfVecComplexInitView", None, [vp, sz, sz, vp]),
                            ("bfVecComplexInit", None, [vp, sz]), ("bfVecDelete", None, [C.POINTER(vp)]),
                            ("bfMatMulVec", vp, [vp, vp])]:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def test_complex_vectors_views_and_ownership():
    lib = vec_lib()
    data = np.arange(12, dtype=np.complex128)
    v = lib.bfVecComplexNew()
    lib.bfVecComplexInitView(v, 4, 3, data.ctypes.data)  # every third entry
    s = VecComplexS.from_address(v)
    assert s.size == 4 and s.stride == 3
    lib.bfVecDelete(C.byref(C.c_void_p(v)))               # a view leaves the data alone
    assert data[9] == 9
    w = lib.bfVecComplexNew()
    lib.bfVecComplexInit(w, 5)
    assert VecComplexS.from_address(w).size == 5
    lib.bfVecDelete(C.byref(C.c_void_p(w)))
    assert lib.bfGetError() == 0


@pytest.mark.gpu
def test_mat_mul_vec(gpu_lib):
    lib = vec_lib()
    rng = np.random.default_rng(4)
    keep = []
    d, D = make_diag(rng, [(3, 5), (6, 2)], keep)
    x = np.zeros(14, complex)
    x[::2] = rand(rng, 7, 1)[:, 0]                          # strided view of 7 entries
    v = lib.bfVecComplexNew()
    lib.bfVecComplexInitView(v, 7, 2, x.ctypes.data)
    y = lib.bfMatMulVec(d, v)
    assert y, lib.bfGetError()
    s = VecComplexS.from_address(y)
    got = np.ctypeslib.as_array(s.data, shape=(2 * s.size,)).view(np.complex128).copy()
    assert s.size == 9 and rel(got, D @ x[::2]) < 1e-13
    lib.bfVecDelete(C.byref(C.c_void_p(y)))
    lib.bfVecDelete(C.byref(C.c_void_p(v)))
    lib.bfMatDelete(C.byref(C.c_void_p(d)))
