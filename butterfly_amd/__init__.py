"""Python host mirror of the MI355X 2-D Helmholtz butterfly library.

Loads the C-ABI library ``libbutterfly_amd.so`` (built from ``csrc/``) and
mirrors the reference's Python wrapper for this path
(wrappers/python/butterfly.pyx): ``Helm2`` (butterfly.pyx:273-328),
``Quadtree``, ``FacHelm2.make_multilevel`` (butterfly.pyx:211-220) and
``Mat.__matmul__`` (butterfly.pyx:390-396).  The factors live in GPU
memory; this module only moves arrays across the C ABI.  There is no CPU
fallback: without the built library or a GPU the calls raise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbutterfly_amd.so")

BF_LAYER_POTENTIAL_SINGLE = 1
BF_LAYER_POTENTIAL_PV_DOUBLE = 2
BF_LAYER_POTENTIAL_PV_NORMAL_DERIV_SINGLE = 3
BF_LAYER_POTENTIAL_COMBINED_FIELD = 5
PRECISION = {"fp64": 0, "fp32": 1, "fp32x2": 2}

ERROR_NAMES = {
    0: "NONE", 1: "INVALID_ARGUMENTS", 2: "RUNTIME_ERROR", 3: "NOT_IMPLEMENTED",
    4: "MEMORY_ERROR", 5: "OUT_OF_RANGE", 6: "FILE_ERROR", 7: "TYPE_ERROR",
    8: "INCOMPATIBLE_SHAPES",
}


class BfError(RuntimeError):
    """Raised when a library call records an error code (error.c:7-24)."""


class Helm2Struct(C.Structure):
    _fields_ = [("k", C.c_double), ("layerPot", C.c_int),
                ("alpha", C.c_double * 2), ("beta", C.c_double * 2)]


class Points2Struct(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_size_t),
                ("capacity", C.c_size_t), ("isView", C.c_bool)]


class Vectors2Struct(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_size_t),
                ("capacity", C.c_size_t), ("isView", C.c_bool)]


class PermStruct(C.Structure):
    _fields_ = [("index", C.POINTER(C.c_size_t)), ("size", C.c_size_t), ("isView", C.c_bool)]


class TreeStruct(C.Structure):
    _fields_ = [("vtable", C.c_void_p), ("root", C.c_void_p), ("perm", C.POINTER(PermStruct))]


class QuadtreeNodeStruct(C.Structure):
    _fields_ = [("bboxMin", C.c_double * 2), ("bboxMax", C.c_double * 2), ("split", C.c_double * 2),
                ("offset", C.c_size_t * 5), ("child", C.c_int32 * 4), ("parent", C.c_int32),
                ("depth", C.c_int32), ("maxDepthBelow", C.c_int32), ("index", C.c_int32)]


class QuadtreeStruct(C.Structure):
    _fields_ = [("super", TreeStruct), ("points", C.c_void_p), ("unitNormals", C.c_void_p),
                ("nodes", C.c_void_p), ("numNodes", C.c_size_t), ("leafSize", C.c_size_t),
                ("levelOffset", C.POINTER(C.c_size_t)), ("numLevels", C.c_size_t)]


class FacOptionsStruct(C.Structure):
    _fields_ = [("tol", C.c_double), ("rankConst", C.c_double),
                ("maxDenseSize", C.c_size_t), ("passThrough", C.c_bool), ("pinvRcond", C.c_double)]


class PtrArrayStruct(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_void_p)), ("capacity", C.c_size_t), ("num_elts", C.c_size_t),
                ("isView", C.c_bool)]


class TreeLevelIterStruct(C.Structure):
    """tree_level_iter.h:7-12"""
    _fields_ = [("traversal", C.c_int), ("nodes", PtrArrayStruct), ("levelNodes", C.POINTER(PtrArrayStruct)),
                ("aux", C.c_void_p)]


class MatStruct(C.Structure):
    """mat.h:187-196"""
    _fields_ = [("vtbl", C.c_void_p), ("props", C.c_int), ("numRows", C.c_size_t), ("numCols", C.c_size_t)]


class MatBlockStruct(C.Structure):
    """mat_block.h:39-66"""
    _fields_ = [("super", MatStruct), ("vtbl", C.c_void_p), ("block", C.POINTER(C.c_void_p)),
                ("rowOffset", C.POINTER(C.c_size_t)), ("colOffset", C.POINTER(C.c_size_t))]


class MatBlockCooStruct(C.Structure):
    """mat_block_coo.h:34-46"""
    _fields_ = [("super", MatBlockStruct), ("numBlocks", C.c_size_t), ("rowInd", C.POINTER(C.c_size_t)),
                ("colInd", C.POINTER(C.c_size_t)), ("numBlockRows", C.c_size_t), ("numBlockCols", C.c_size_t),
                ("ownsBlock", C.c_void_p)]


class MatBlockDiagStruct(C.Structure):
    """mat_block_diag.h:33-37"""
    _fields_ = [("super", MatBlockStruct), ("numBlockRows", C.c_size_t), ("numBlockCols", C.c_size_t),
                ("ownsBlock", C.c_void_p)]


class MatDenseComplexStruct(C.Structure):
    _fields_ = [("vtbl", C.c_void_p), ("props", C.c_int), ("numRows", C.c_size_t),
                ("numCols", C.c_size_t), ("rowStride", C.c_size_t), ("colStride", C.c_size_t),
                ("data", C.c_void_p), ("pyArray", C.c_void_p)]


_lib = None


def lib():
    """Load the C-ABI library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BfError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C butterfly_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp, sz, dp, i64p = C.c_void_p, C.c_size_t, C.POINTER(C.c_double), C.POINTER(C.c_int64)
    fp = C.POINTER(C.c_float)
    sig = {
        "bfGetError": (C.c_int, []),
        "bfSetError": (None, [C.c_int]),
        "bfSeed": (None, [sz]),
        "bfRealUniform1": (C.c_double, []),
        "bfRealUniform": (None, [sz, dp]),
        "bfRealRandn": (None, [sz, dp]),
        "bfComplexRandn": (None, [sz, dp]),
        "bfPoints2NewViewFromPtr": (vp, [dp, sz]),
        "bfPoints2DeinitAndDealloc": (None, [C.POINTER(vp)]),
        "bfQuadtreeNew": (vp, []),
        "bfQuadtreeInitWithLeafSize": (None, [vp, vp, vp, sz]),
        "bfhipQuadtreeInitWithLeafSize": (None, [vp, vp, vp, sz]),
        "bfhipSvdTallMaxN": (C.c_int, []),
        "bfQuadtreeDeinitAndDealloc": (None, [C.POINTER(vp)]),
        "bfHelm2RankEstForTwoCircles": (sz, [vp, vp, vp, C.c_double, C.c_double]),
        "bfHelm2GetKernelMatrix": (vp, [vp, vp, vp, vp, vp]),
        "bfFacHelm2OptionsInitDefault": (None, [vp]),
        "bfFacHelm2MakeMultilevel": (vp, [vp, vp, vp]),
        "bfFacHelm2MakeMultilevelWithOptions": (vp, [vp, vp, vp, vp]),
        "bfhipFacHelm2Plan": (vp, [vp, vp, vp, vp]),
        "bfQuadtreeNodeGetQuadtreeConst": (vp, [vp]),
        "bfTreeLevelIterInit": (None, [vp, C.c_int, vp]),
        "bfTreeLevelIterDeinit": (None, [vp]),
        "bfTreeLevelIterCurrentDepth": (sz, [vp]),
        "bfTreeLevelIterIsDone": (C.c_bool, [vp]),
        "bfTreeLevelIterNext": (None, [vp]),
        "bfTreeLevelIterGetNumPoints": (sz, [vp]),
        "bfTreeLevelIterCurrentLevelIsInternal": (C.c_bool, [vp]),
        "bfFacHelm2Prepare": (sz, [vp, vp, vp, vp, vp]),
        "bfFacHelm2Make": (vp, [vp, vp, vp, vp, vp, sz]),
        "bfFacHelm2MakeSingleLevel": (vp, [vp, vp, vp]),
        "bfFacGetMatProduct": (vp, [vp, C.c_int]),
        "bfMatProductNumFactors": (sz, [vp]),
        "bfMatProductGetFactor": (vp, [vp, sz]),
        "bfhipFacHelm2MakeBatched": (vp, [sz, vp, C.POINTER(vp), C.POINTER(vp), vp]),
        "bfMatGetType": (C.c_int, [vp]),
        "bfMatGetNumRows": (sz, [vp]),
        "bfMatGetNumCols": (sz, [vp]),
        "bfMatNumBytes": (sz, [vp]),
        "bfMatDelete": (None, [C.POINTER(vp)]),
        "bfMatMul": (vp, [vp, vp]),
        "bfMatDenseComplexNew": (vp, []),
        "bfMatDenseComplexInitViewFromPtr": (None, [vp, sz, sz, vp]),
        "bfhipFacLeafNumFactors": (sz, [vp, sz]),
        "bfhipMalloc": (vp, [sz]),
        "bfhipFree": (None, [vp]),
        "bfhipMemcpyHtoD": (C.c_int, [vp, vp, sz]),
        "bfhipMemcpyDtoH": (C.c_int, [vp, vp, sz]),
        "bfhipSynchronize": (C.c_int, []),
        "bfhipDeviceCount": (C.c_int, []),
        "bfhipDeviceCopy": (C.c_int, [vp, vp, sz]),
        "bfhipMatMulDevice": (C.c_int, [vp, vp, vp, sz, C.c_int]),
        "bfhipMatMulDeviceNoAlloc": (C.c_int, [vp, vp, vp, sz, C.c_int]),
        "bfhipFacPrepareMixed": (C.c_int, [vp, C.c_int]),
        "bfhipFacReleaseFp64": (C.c_int, [vp]),
        "bfhipFacOrthogonalize": (C.c_int, [vp]),
        "bfhipFacOrthogonalizeSeconds": (C.c_double, [vp]),
        "bfhipFacGetStats": (None, [vp, dp]),
        "bfhipGetKernelTimes": (None, [dp]),
        "bfhipGetKernelTimeSums": (None, [dp, dp]),
        "bfhipResetKernelTimes": (None, []),
        "bfhipSetApplyConfig": (C.c_int, [C.c_int, C.c_int, C.c_int]),
        "bfhipSetApplyCompensation": (C.c_int, [C.c_int]),
        "bfhipSetApplyGraphs": (C.c_int, [C.c_int]),
        "bfhipSetLstsqWaves": (None, [C.c_int64, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "bfhipFacTuneApply": (C.c_int, [vp, sz, C.c_int, C.c_char_p]),
        "bfhipFacGetApplyConfig": (C.c_int, [vp, sz, C.c_int, C.POINTER(C.c_int32)]),
        "bfhipFacPlanSizes": (None, [vp, C.POINTER(sz), C.POINTER(sz), C.POINTER(sz)]),
        "bfhipFacPlanExport": (None, [vp, i64p, i64p, i64p]),
        "bfhipFacBlockData": (C.c_int, [vp, sz, dp]),
        "bfhipFacDenseData": (C.c_int, [vp, sz, dp]),
        "bfhipFacBlockGeometry": (C.c_int, [vp, sz, dp]),
        "bfhipHankel0": (C.c_int, [sz, dp, dp]),
        "bfhipLstSqBatched": (C.c_int, [sz, sz, sz, dp, dp, dp]),
        "bfMatSave": (None, [vp, C.c_char_p]),
        "bfhipFacLoadFile": (vp, [C.c_char_p]),
        "bfhipSvdBatched": (C.c_int, [sz, C.POINTER(sz), C.POINTER(sz), dp, C.c_double, C.POINTER(sz), dp, dp]),
        "bfFacStreamerNew": (vp, []),
        "bfFacStreamerInit": (None, [vp, vp]),
        "bfFacStreamerDelete": (None, [C.POINTER(vp)]),
        "bfFacStreamerGetNumRows": (sz, [vp]),
        "bfFacStreamerFeed": (C.c_bool, [vp, vp]),
        "bfFacStreamerIsDone": (C.c_bool, [vp]),
        "bfFacStreamerGetFac": (vp, [vp]),
        "bfFacStreamerGetFacSpan": (vp, [vp]),
        "bfFacStreamerGetCurrentColumnNode": (vp, [vp]),
        "bfFacStreamerGetRowTreeReversePerm": (vp, [vp]),
        "bfFacGetMat": (vp, [vp, C.c_int]),
        "bfFacGetNumBytes": (sz, [vp]),
        "bfFacSpanGetMat": (vp, [vp, C.c_int]),
        "bfFacSpanDelete": (None, [C.POINTER(vp)]),
        "bfFacSpanGetNumFacs": (sz, [vp]),
        "bfFacGetNumW": (sz, [vp]),
        "bfFacGetRowNodes": (sz, [vp, C.POINTER(C.c_int32), sz]),
        "bfFacGetFactorShape": (C.c_int, [vp, sz, C.POINTER(sz), C.POINTER(sz), C.POINTER(sz)]),
        "bfFacGetFactorDense": (C.c_int, [vp, sz, dp]),
        "bfhipHelm2BieNew": (vp, [vp, vp, vp, dp, vp, sz, C.c_double]),
        "bfhipBieMulDevice": (C.c_int, [vp, vp, vp, sz, C.c_int]),
        "bfSolveGMRES": (vp, [vp, vp, vp, C.c_double, sz, C.POINTER(sz), vp]),
        "bfhipSolveGmresMixed": (C.c_int, [vp, dp, dp, C.c_double, sz, sz, C.POINTER(sz), C.POINTER(sz), dp]),
        "bfhipSvdBatchedSketch": (C.c_int, [sz, C.POINTER(sz), C.POINTER(sz), dp, C.c_double, C.POINTER(sz),
                                            C.c_uint64, C.POINTER(sz), dp, dp, C.POINTER(C.c_int)]),
        "bfhipRankFeatures": (None, [dp, dp, C.c_double, sz, C.c_double, fp]),
        "bfhipRankModelNew": (vp, [C.c_uint64]),
        "bfhipRankModelDelete": (None, [C.POINTER(vp)]),
        "bfhipRankModelTrain": (C.c_int, [vp, sz, fp, fp, sz, sz, C.c_float, fp]),
        "bfhipRankModelPredict": (C.c_int, [vp, sz, fp, fp]),
        "bfhipRankModelGradient": (C.c_int, [vp, sz, fp, fp, fp, fp]),
        "bfhipRankModelNumParams": (sz, []),
        "bfhipRankModelGetParams": (C.c_int, [vp, fp, fp]),
        "bfhipRankModelSetParams": (C.c_int, [vp, fp, fp]),
        "bfhipFacStreamerSetRankModel": (C.c_int, [vp, vp, C.c_double, sz]),
        "bfhipFacStreamerGetRankSamples": (sz, [vp, fp, fp, sz]),
        "bfhipFacStreamerGetSketchStats": (None, [vp, C.POINTER(sz), C.POINTER(sz), C.POINTER(sz)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def check():
    """Raise BfError if the library recorded an error (and clear it)."""
    e = lib().bfGetError()
    if e:
        raise BfError(ERROR_NAMES.get(e, str(e)))


def seed(s):
    lib().bfSeed(s)


def complex_randn(n):
    """bfComplexRandn: n complex standard normals from the seeded stream."""
    out = np.empty(2 * n)
    lib().bfComplexRandn(n, _dp(out))
    return out.view(np.complex128)


def real_uniform(n):
    out = np.empty(n)
    lib().bfRealUniform(n, _dp(out))
    return out


def sample_box(n, x0, x1, y0, y1):
    """n points uniform in [x0,x1] x [y0,y1], x then y per point (points.c:25-28)."""
    u = real_uniform(2 * n).reshape(n, 2)
    return np.stack([(x1 - x0) * u[:, 0] + x0, (y1 - y0) * u[:, 1] + y0], 1)


class Helm2:
    """Kernel parameters (helm2.h:10-15): the single layer, the double layer
    (BF_LAYER_POTENTIAL_PV_DOUBLE), the combined field alpha S + beta D
    (BF_LAYER_POTENTIAL_COMBINED_FIELD) or the normal derivative of the
    single layer S' (BF_LAYER_POTENTIAL_PV_NORMAL_DERIV_SINGLE, target
    normals)."""

    def __init__(self, k, layer_pot=None, alpha=1.0, beta=1.0):
        if layer_pot is None:
            layer_pot = BF_LAYER_POTENTIAL_SINGLE
        a, b = complex(alpha), complex(beta)
        self.s = Helm2Struct(float(k), layer_pot, (a.real, a.imag), (b.real, b.imag))

    @property
    def k(self):
        return self.s.k

    def get_kernel_matrix(self, src, tgt, src_normals=None, tgt_normals=None):
        """bfHelm2GetKernelMatrix: dense (tgt x src) kernel, sampled on the GPU
        (source unit normals for the double layer and the combined field,
        target unit normals for S')."""
        src = np.ascontiguousarray(src, dtype=np.float64)
        tgt = np.ascontiguousarray(tgt, dtype=np.float64)
        L = lib()
        ps = L.bfPoints2NewViewFromPtr(_dp(src), len(src))
        pt = L.bfPoints2NewViewFromPtr(_dp(tgt), len(tgt))
        nv = tv = None
        if src_normals is not None:
            nrm = np.ascontiguousarray(src_normals, dtype=np.float64)
            nv = Vectors2Struct(nrm.ctypes.data, len(nrm), len(nrm), True)
        if tgt_normals is not None:
            tnrm = np.ascontiguousarray(tgt_normals, dtype=np.float64)
            tv = Vectors2Struct(tnrm.ctypes.data, len(tnrm), len(tnrm), True)
        m = L.bfHelm2GetKernelMatrix(C.byref(self.s), ps, pt, C.byref(nv) if nv is not None else None,
                                     C.byref(tv) if tv is not None else None)
        L.bfPoints2DeinitAndDealloc(C.byref(C.c_void_p(ps)))
        L.bfPoints2DeinitAndDealloc(C.byref(C.c_void_p(pt)))
        check()
        return Mat(m).to_numpy()


class Quadtree:
    """bfQuadtreeInit over an (N, 2) array; keeps the array alive."""

    def __init__(self, points, leaf_size=1, normals=None, device=False):
        """`normals`: unit normals of the points (original order), needed by
        the double layer and the combined field on a source tree and by S' on
        a target tree.  `device`: build the tree on the GPU
        (bfhipQuadtreeInitWithLeafSize)."""
        self.points = np.ascontiguousarray(points, dtype=np.float64)
        L = lib()
        self._pts = L.bfPoints2NewViewFromPtr(_dp(self.points), len(self.points))
        self._nrm = None
        if normals is not None:
            self.normals = np.ascontiguousarray(normals, dtype=np.float64)
            if self.normals.shape != self.points.shape:
                raise ValueError("one unit normal per point")
            self._nrm = Vectors2Struct(self.normals.ctypes.data, len(self.normals), len(self.normals), True)
        self.ptr = L.bfQuadtreeNew()
        init = L.bfhipQuadtreeInitWithLeafSize if device else L.bfQuadtreeInitWithLeafSize
        init(self.ptr, self._pts, C.byref(self._nrm) if self._nrm is not None else None, leaf_size)
        check()

    @property
    def perm(self):
        """Tree permutation: tree position i holds original point perm[i]."""
        q = QuadtreeStruct.from_address(self.ptr)
        p = q.super.perm.contents
        return np.ctypeslib.as_array(p.index, shape=(p.size,)).astype(np.int64)

    @property
    def num_nodes(self):
        return QuadtreeStruct.from_address(self.ptr).numNodes

    def node_ptr(self, i):
        """Address of node i (breadth-first order) for the node-level API."""
        q = QuadtreeStruct.from_address(self.ptr)
        if not 0 <= i < q.numNodes:
            raise IndexError(i)
        return q.nodes + i * C.sizeof(QuadtreeNodeStruct)

    def node(self, i):
        return QuadtreeNodeStruct.from_address(self.node_ptr(i))

    def level(self, d):
        """Indices of the nodes of depth d."""
        q = QuadtreeStruct.from_address(self.ptr)
        return list(range(q.levelOffset[d], q.levelOffset[d + 1]))

    def __del__(self):
        try:
            L = lib()
            L.bfQuadtreeDeinitAndDealloc(C.byref(C.c_void_p(self.ptr)))
            L.bfPoints2DeinitAndDealloc(C.byref(C.c_void_p(self._pts)))
        except Exception:
            pass


def options(tol=1e-15, rank_const=1.0, max_dense=128 * 128, pass_through=False, pinv_rcond=0.0):
    """BfFacHelm2Options; the defaults are the reference's fixed rules."""
    o = FacOptionsStruct()
    lib().bfFacHelm2OptionsInitDefault(C.byref(o))
    o.tol, o.rankConst, o.maxDenseSize, o.passThrough = tol, rank_const, max_dense, pass_through
    o.pinvRcond = pinv_rcond
    return o


def dense_numpy(ptr):
    """A copy of the BfMatDenseComplex at `ptr` (not taking ownership)."""
    d = MatDenseComplexStruct.from_address(ptr)
    n = d.numRows * d.numCols
    if n == 0:
        return np.zeros((d.numRows, d.numCols), np.complex128)
    span = (d.numRows - 1) * d.rowStride + (d.numCols - 1) * d.colStride + 1
    a = np.ctypeslib.as_array(C.cast(d.data, C.POINTER(C.c_double)), shape=(2 * span,)).copy().view(np.complex128)
    return np.lib.stride_tricks.as_strided(a, (d.numRows, d.numCols), (16 * d.rowStride, 16 * d.colStride)).copy()


class Mat:
    """A BfMat handle owned by Python (freed with bfMatDelete)."""

    def __init__(self, ptr, keep=None):
        if not ptr:
            check()
            raise BfError("NULL matrix")
        self.ptr = ptr
        self._keep = keep

    @property
    def shape(self):
        L = lib()
        return (L.bfMatGetNumRows(self.ptr), L.bfMatGetNumCols(self.ptr))

    def to_numpy(self):
        return dense_numpy(self.ptr)

    def __matmul__(self, x):
        """bfMatMul(self, X): X (N_src x nrhs) in source-tree order."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        squeeze = x.ndim == 1
        if squeeze:
            x = x[:, None]
        L = lib()
        xm = L.bfMatDenseComplexNew()
        L.bfMatDenseComplexInitViewFromPtr(xm, x.shape[0], x.shape[1], x.ctypes.data)
        y = L.bfMatMul(self.ptr, xm)
        L.bfMatDelete(C.byref(C.c_void_p(xm)))
        check()
        out = Mat(y).to_numpy()
        return out[:, 0] if squeeze else out

    def __del__(self):
        try:
            if self.ptr:
                lib().bfMatDelete(C.byref(C.c_void_p(self.ptr)))
                self.ptr = None
        except Exception:
            pass


class Factorization(Mat):
    """The device factorization returned by FacHelm2.make_multilevel."""

    def stats(self):
        out = np.zeros(16)
        lib().bfhipFacGetStats(self.ptr, _dp(out))
        keys = ["factor_s", "plan_s", "kernels_s", "entries", "lstsq_blocks", "kernel_blocks",
                "identity_blocks", "dense_leaves", "butterfly_leaves", "stages", "h0_evals",
                "apply_factor_bytes", "apply_vector_bytes", "sample_ms", "compress_ms", "apply_plan_s"]
        return dict(zip(keys, out.tolist()))

    def plan(self):
        """(leafTab [L,6], facTab [F,4], blkTab [B,7]) in the oracle's layout."""
        L = lib()
        a, b, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
        L.bfhipFacPlanSizes(self.ptr, C.byref(a), C.byref(b), C.byref(c))
        lt = np.zeros((a.value, 6), np.int64)
        ft = np.zeros((b.value, 4), np.int64)
        bt = np.zeros((c.value, 7), np.int64)
        p = C.POINTER(C.c_int64)
        L.bfhipFacPlanExport(self.ptr, lt.ctypes.data_as(p), ft.ctypes.data_as(p), bt.ctypes.data_as(p))
        return lt, ft, bt

    def block(self, idx, m, n):
        out = np.zeros(2 * m * n)
        if lib().bfhipFacBlockData(self.ptr, idx, _dp(out)):
            check()
            raise BfError("block copy failed")
        return out.view(np.complex128).reshape(m, n)

    def block_geometry(self, idx):
        out = np.zeros(12)
        lib().bfhipFacBlockGeometry(self.ptr, idx, _dp(out))
        return out

    def dense_leaf(self, idx, m, n):
        out = np.zeros(2 * m * n)
        if lib().bfhipFacDenseData(self.ptr, idx, _dp(out)):
            check()
            raise BfError("dense copy failed")
        return out.view(np.complex128).reshape(m, n)

    def prepare_mixed(self, precision="fp32", release_fp64=False):
        if lib().bfhipFacPrepareMixed(self.ptr, PRECISION[precision]):
            check()
            raise BfError("prepare_mixed failed")
        if release_fp64 and lib().bfhipFacReleaseFp64(self.ptr):
            check()
            raise BfError("release_fp64 failed")

    def save(self, path):
        """bfMatSave: the factorization in the bfMatDump framing."""
        lib().bfMatSave(self.ptr, str(path).encode())
        check()

    @staticmethod
    def load(path):
        """bfhipFacLoadFile: a saved factorization back in HBM."""
        p = lib().bfhipFacLoadFile(str(path).encode())
        if not p:
            check()
            raise BfError("load failed")
        return Factorization(p)

    def orthogonalize(self):
        """bfhipFacOrthogonalize: rotate every intermediate basis so the map
        to the output is orthonormal (same operator); returns the seconds."""
        if lib().bfhipFacOrthogonalize(self.ptr):
            check()
            raise BfError("orthogonalize failed")
        return lib().bfhipFacOrthogonalizeSeconds(self.ptr)

    def tune_apply(self, nrhs, precision="fp64", cache=None):
        """bfhipFacTuneApply: time the launch shapes of every stage for nrhs
        right-hand sides and keep the fastest (cached per stage bucket in
        the text file `cache`, if given)."""
        rc = lib().bfhipFacTuneApply(self.ptr, nrhs, PRECISION[precision],
                                     None if cache is None else str(cache).encode())
        if rc:
            check()
            raise BfError(f"tuning failed ({rc})")

    def apply_config(self, nrhs, precision="fp64"):
        """The launch shape used by each stage (-1 = the global settings)."""
        n = lib().bfhipFacGetApplyConfig(self.ptr, nrhs, PRECISION[precision], None)
        out = (C.c_int32 * max(n, 1))()
        lib().bfhipFacGetApplyConfig(self.ptr, nrhs, PRECISION[precision], out)
        return list(out[:n])


def set_apply_graphs(on):
    """bfhipSetApplyGraphs: replay the apply's stage chain as a hipGraph."""
    lib().bfhipSetApplyGraphs(1 if on else 0)


def set_lstsq_waves(two_above=-1, four_above=-1):
    """bfhipSetLstsqWaves: re-expansion blocks whose LDS class exceeds
    two_above (four_above) bytes are solved by 2 (4) waves; -1 keeps a
    threshold.  Returns the previous (two_above, four_above)."""
    a, b = C.c_int64(), C.c_int64()
    lib().bfhipSetLstsqWaves(int(two_above), int(four_above), C.byref(a), C.byref(b))
    return a.value, b.value


class FacHelm2:
    @staticmethod
    def make_multilevel(helm, src_tree, tgt_tree, opts=None):
        """bfFacHelm2MakeMultilevel (fac_helm2.c:943-1002); opts selects the
        config-driven planner (bfFacHelm2MakeMultilevelWithOptions)."""
        L = lib()
        if opts is None:
            p = L.bfFacHelm2MakeMultilevel(C.byref(helm.s), src_tree.ptr, tgt_tree.ptr)
        else:
            p = L.bfFacHelm2MakeMultilevelWithOptions(C.byref(helm.s), src_tree.ptr, tgt_tree.ptr, C.byref(opts))
        check()
        return Factorization(p, keep=(src_tree, tgt_tree))

    @staticmethod
    def make_batched(helms, src_trees, tgt_trees, opts_list):
        """bfhipFacHelm2MakeBatched: independent problems as one
        block-diagonal operator (inputs and outputs stacked in order)."""
        n = len(helms)
        if not (len(src_trees) == len(tgt_trees) == len(opts_list) == n):
            raise ValueError("one helm, source tree, target tree and options per problem")
        H = (Helm2Struct * n)(*[h.s for h in helms])
        S = (C.c_void_p * n)(*[t.ptr for t in src_trees])
        T = (C.c_void_p * n)(*[t.ptr for t in tgt_trees])
        O = (FacOptionsStruct * n)(*opts_list)
        p = lib().bfhipFacHelm2MakeBatched(n, C.cast(H, C.c_void_p), S, T, C.cast(O, C.c_void_p))
        check()
        if not p:
            raise BfError("batched factorization failed")
        return Factorization(p, keep=(tuple(src_trees), tuple(tgt_trees)))

    @staticmethod
    def make_single_level(helm, src_tree, src_node, tgt_tree, tgt_node):
        """bfFacHelm2MakeSingleLevel (fac_helm2.c:706-729) for node indices
        src_node / tgt_node of the two trees: a host BfMatProduct of block
        factors (the reference's containers) from the source node's points
        to the target node's points, both in tree order."""
        p = lib().bfFacHelm2MakeSingleLevel(C.byref(helm.s), src_tree.node_ptr(src_node),
                                            tgt_tree.node_ptr(tgt_node))
        check()
        return Mat(p, keep=(src_tree, tgt_tree))

    @staticmethod
    def plan_only(helm, src_tree, tgt_tree, opts=None):
        """Host planner only (bfhipFacHelm2Plan): no GPU needed."""
        L = lib()
        if opts is None:
            opts = options()
        p = L.bfhipFacHelm2Plan(C.byref(helm.s), src_tree.ptr, tgt_tree.ptr, C.byref(opts))
        check()
        return Factorization(p, keep=(src_tree, tgt_tree))


def hankel0(x):
    """Device H0^(1) at the given arguments: returns complex J0 + i Y0."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(2 * len(x))
    if lib().bfhipHankel0(len(x), _dp(x), _dp(out)):
        check()
        raise BfError("hankel0 failed")
    return out.view(np.complex128)


def lstsq_batched(A, B):
    """Device pinv(A_b) B_b with the reference's truncation, A: (b,p,p), B: (b,p,n)."""
    A = np.ascontiguousarray(A, dtype=np.complex128)
    B = np.ascontiguousarray(B, dtype=np.complex128)
    b, p, _ = A.shape
    n = B.shape[2]
    X = np.zeros((b, p, n), np.complex128)
    rc = lib().bfhipLstSqBatched(b, p, n, _dp(A.view(np.float64)), _dp(B.view(np.float64)),
                                 _dp(X.view(np.float64)))
    if rc:
        check()
        raise BfError(f"lstsq_batched failed ({rc})")
    return X


def svd_batched(mats, tol):
    """bfhipSvdBatched: truncated SVDs A = U W (W = S V^H) on the device;
    returns [(U, W, k)]."""
    ms = [np.asarray(a, np.complex128) for a in mats]
    m = (C.c_size_t * len(ms))(*[a.shape[0] for a in ms])
    n = (C.c_size_t * len(ms))(*[a.shape[1] for a in ms])
    A = np.concatenate([a.ravel(order="F") for a in ms]) if ms else np.zeros(1, complex)
    full = [min(a.shape) for a in ms]
    U = np.zeros(max(1, sum(a.shape[0] * f for a, f in zip(ms, full))), np.complex128)
    W = np.zeros(max(1, sum(f * a.shape[1] for a, f in zip(ms, full))), np.complex128)
    k = (C.c_size_t * len(ms))()
    if lib().bfhipSvdBatched(len(ms), m, n, _dp(A.view(np.float64)), tol, k, _dp(U.view(np.float64)),
                             _dp(W.view(np.float64))):
        check()
        raise BfError("svd_batched failed")
    out, uo, wo = [], 0, 0
    for a, f, kk in zip(ms, full, k):
        out.append((U[uo:uo + a.shape[0] * kk].reshape(kk, a.shape[0]).T,
                    W[wo:wo + kk * a.shape[1]].reshape(a.shape[1], kk).T, kk))
        uo += a.shape[0] * f
        wo += f * a.shape[1]
    return out


def svd_batched_sketch(mats, tol, widths, seed=1):
    """bfhipSvdBatchedSketch: randomized truncated SVDs of sketch widths
    `widths` (full SVD where a sketch is rejected or not narrower);
    returns [(U, W, k, fell_back)]."""
    ms = [np.asarray(a, np.complex128) for a in mats]
    nb = len(ms)
    m = (C.c_size_t * nb)(*[a.shape[0] for a in ms])
    n = (C.c_size_t * nb)(*[a.shape[1] for a in ms])
    w = (C.c_size_t * nb)(*[int(x) for x in widths])
    A = np.concatenate([a.ravel(order="F") for a in ms]) if ms else np.zeros(1, complex)
    full = [min(a.shape) for a in ms]
    U = np.zeros(max(1, sum(a.shape[0] * f for a, f in zip(ms, full))), np.complex128)
    W = np.zeros(max(1, sum(f * a.shape[1] for a, f in zip(ms, full))), np.complex128)
    k = (C.c_size_t * nb)()
    fell = (C.c_int * nb)()
    if lib().bfhipSvdBatchedSketch(nb, m, n, _dp(A.view(np.float64)), tol, w, seed, k, _dp(U.view(np.float64)),
                                   _dp(W.view(np.float64)), fell):
        check()
        raise BfError("svd_batched_sketch failed")
    out, uo, wo = [], 0, 0
    for a, f, kk, fb in zip(ms, full, k, fell):
        out.append((U[uo:uo + a.shape[0] * kk].reshape(kk, a.shape[0]).T,
                    W[wo:wo + kk * a.shape[1]].reshape(a.shape[1], kk).T, kk, bool(fb)))
        uo += a.shape[0] * f
        wo += f * a.shape[1]
    return out


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def rank_features(col_box, row_box, k, level, tol):
    """bfhipRankFeatures: the rank model's six inputs for a (column node,
    row node) pair; boxes are (xmin, ymin, xmax, ymax)."""
    cb = np.ascontiguousarray(col_box, np.float64)
    rb = np.ascontiguousarray(row_box, np.float64)
    out = np.zeros(6, np.float32)
    lib().bfhipRankFeatures(_dp(cb), _dp(rb), float(k), int(level), float(tol), _fp(out))
    return out


class RankModel:
    """bfhipRankModel*: the FP32 rank MLP (6 -> 64 -> 64 -> 64 -> 1, GELU),
    trained and evaluated on the device."""

    def __init__(self, seed=0):
        self.ptr = lib().bfhipRankModelNew(seed)
        if not self.ptr:
            check()
            raise BfError("rank model allocation failed")

    def train(self, features, ranks, epochs=10, batch=4096, lr=1e-3):
        """Mini-batch Adam on the squared error; returns the mean squared
        error (ranks^2) of the last epoch's mini-batches."""
        X = np.ascontiguousarray(features, np.float32).reshape(-1, 6)
        y = np.ascontiguousarray(ranks, np.float32).ravel()
        loss = np.zeros(1, np.float32)
        if lib().bfhipRankModelTrain(self.ptr, len(y), _fp(X), _fp(y), epochs, batch, lr, _fp(loss)):
            check()
            raise BfError("rank model training failed")
        return float(loss[0])

    def predict(self, features):
        X = np.ascontiguousarray(features, np.float32).reshape(-1, 6)
        out = np.zeros(len(X), np.float32)
        if lib().bfhipRankModelPredict(self.ptr, len(X), _fp(X), _fp(out)):
            check()
            raise BfError("rank model prediction failed")
        return out

    def gradient(self, features, ranks):
        """(gradient, mean squared error) over all samples in standardized
        units, parameters unchanged."""
        X = np.ascontiguousarray(features, np.float32).reshape(-1, 6)
        y = np.ascontiguousarray(ranks, np.float32).ravel()
        g = np.zeros(lib().bfhipRankModelNumParams(), np.float32)
        loss = np.zeros(1, np.float32)
        if lib().bfhipRankModelGradient(self.ptr, len(y), _fp(X), _fp(y), _fp(g), _fp(loss)):
            check()
            raise BfError("rank model gradient failed")
        return g, float(loss[0])

    @property
    def params(self):
        """(parameters, normalization): W1 64x6, b1, W2 64x64, b2, W3 64x64,
        b3, W4 1x64, b4 (row-major W[out][in]); feature mean[6], deviation[6],
        target mean, deviation."""
        p = np.zeros(lib().bfhipRankModelNumParams(), np.float32)
        nm = np.zeros(14, np.float32)
        if lib().bfhipRankModelGetParams(self.ptr, _fp(p), _fp(nm)):
            check()
            raise BfError("rank model read failed")
        return p, nm

    def set_params(self, p, nm):
        p = np.ascontiguousarray(p, np.float32)
        nm = np.ascontiguousarray(nm, np.float32)
        if len(p) != lib().bfhipRankModelNumParams() or len(nm) != 14:
            raise BfError("wrong parameter count")
        if lib().bfhipRankModelSetParams(self.ptr, _fp(p), _fp(nm)):
            check()
            raise BfError("rank model write failed")

    def __del__(self):
        try:
            if self.ptr:
                lib().bfhipRankModelDelete(C.byref(C.c_void_p(self.ptr)))
                self.ptr = None
        except Exception:
            pass


class FacSpecStruct(C.Structure):
    _fields_ = [("rowTree", C.c_void_p), ("colTree", C.c_void_p), ("rowTreeInitDepth", C.c_size_t),
                ("tol", C.c_double), ("minNumRows", C.c_size_t), ("minNumCols", C.c_size_t),
                ("compareRelativeErrors", C.c_bool), ("bailIfBottomedOut", C.c_bool)]


class Fac:
    """A finished streamed factorization (BfFac, owned by its streamer)."""

    def __init__(self, ptr, owner):
        self.ptr, self._owner = ptr, owner

    @property
    def num_w(self):
        return lib().bfFacGetNumW(self.ptr)

    def row_nodes(self):
        n = lib().bfFacGetRowNodes(self.ptr, None, 0)
        out = (C.c_int32 * max(n, 1))()
        lib().bfFacGetRowNodes(self.ptr, out, n)
        return np.array(out[:n], dtype=np.int64)

    def factor_shape(self, which):
        r, c, b = C.c_size_t(), C.c_size_t(), C.c_size_t()
        lib().bfFacGetFactorShape(self.ptr, which, C.byref(r), C.byref(c), C.byref(b))
        return r.value, c.value, b.value

    def factor_dense(self, which):
        r, c, _ = self.factor_shape(which)
        out = np.zeros(2 * r * c)
        lib().bfFacGetFactorDense(self.ptr, which, _dp(out))
        return out.view(np.complex128).reshape(r, c)

    def num_bytes(self):
        return lib().bfFacGetNumBytes(self.ptr)

    def mat(self):
        """bfFacGetMat: the device product Psi W0 ... W_L."""
        return Mat(lib().bfFacGetMat(self.ptr, 1), keep=self._owner)

    def mat_product(self):
        """bfFacGetMatProduct: the same product as host block containers."""
        return Mat(lib().bfFacGetMatProduct(self.ptr, 1), keep=self._owner)


class FacStreamer:
    """bfFacStreamer over two Quadtrees (rows, columns)."""

    def __init__(self, row_tree, col_tree, tol, init_depth=1, min_rows=1, min_cols=1, bail=False):
        self.row_tree, self.col_tree = row_tree, col_tree
        self.spec = FacSpecStruct(row_tree.ptr, col_tree.ptr, init_depth, tol, min_rows, min_cols, False, bail)
        L = lib()
        self.ptr = L.bfFacStreamerNew()
        L.bfFacStreamerInit(self.ptr, C.byref(self.spec))
        check()

    @property
    def num_rows(self):
        return lib().bfFacStreamerGetNumRows(self.ptr)

    @property
    def done(self):
        return bool(lib().bfFacStreamerIsDone(self.ptr))

    def current_column_node(self):
        """Index of the current column node in the column tree's node array."""
        p = lib().bfFacStreamerGetCurrentColumnNode(self.ptr)
        if not p:
            return -1
        base = QuadtreeStruct.from_address(self.col_tree.ptr).super.root
        return (p - base) // C.sizeof(QuadtreeNodeStruct)

    def feed(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        L = lib()
        xm = L.bfMatDenseComplexNew()
        L.bfMatDenseComplexInitViewFromPtr(xm, phi.shape[0], phi.shape[1], phi.ctypes.data)
        ok = L.bfFacStreamerFeed(self.ptr, xm)
        L.bfMatDelete(C.byref(C.c_void_p(xm)))
        if not ok:
            err = L.bfGetError()
            if err:
                raise BfError(f"feed failed (error {err})")
        return bool(ok)

    def set_rank_model(self, model, k=0.0, oversample=8):
        """bfhipFacStreamerSetRankModel: randomized SVDs of width predicted
        rank + oversample (model None: plain SVDs); samples recorded either way."""
        self.rank_model = model
        if lib().bfhipFacStreamerSetRankModel(self.ptr, model.ptr if model is not None else None, float(k),
                                              int(oversample)):
            check()
            raise BfError("set_rank_model failed")

    def rank_samples(self):
        """(features n x 6, numerical ranks n) of the SVDs so far."""
        L = lib()
        n = L.bfhipFacStreamerGetRankSamples(self.ptr, None, None, 0)
        X = np.zeros((n, 6), np.float32)
        y = np.zeros(n, np.float32)
        if n:
            L.bfhipFacStreamerGetRankSamples(self.ptr, _fp(X), _fp(y), n)
        return X, y

    def sketch_stats(self):
        a, b, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
        lib().bfhipFacStreamerGetSketchStats(self.ptr, C.byref(a), C.byref(b), C.byref(c))
        return {"svds": a.value, "sketched": b.value, "fell_back": c.value}

    def fac(self):
        p = lib().bfFacStreamerGetFac(self.ptr)
        if not p:
            check()
            raise BfError("no factorization")
        return Fac(p, self)

    def span_mat(self):
        L = lib()
        sp = L.bfFacStreamerGetFacSpan(self.ptr)
        m = L.bfFacSpanGetMat(sp, 1)
        L.bfFacSpanDelete(C.byref(C.c_void_p(sp)))
        return Mat(m, keep=self)

    def __del__(self):
        try:
            if self.ptr:
                lib().bfFacStreamerDelete(C.byref(C.c_void_p(self.ptr)))
                self.ptr = None
        except Exception:
            pass


class DeviceArray:
    """A device allocation filled from / read back to numpy (bfhipMalloc)."""

    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.ptr = lib().bfhipMalloc(nbytes)
        if not self.ptr:
            check()
            raise BfError("bfhipMalloc failed")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if lib().bfhipMemcpyHtoD(d.ptr, a.ctypes.data, a.nbytes):
            check()
        return d

    def to_numpy(self, dtype, shape):
        out = np.empty(shape, dtype)
        if lib().bfhipMemcpyDtoH(out.ctypes.data, self.ptr, out.nbytes):
            check()
        return out

    def __del__(self):
        try:
            if self.ptr:
                lib().bfhipFree(self.ptr)
                self.ptr = None
        except Exception:
            pass


def mul_device(fac, dx, dy, nrhs, precision="fp64"):
    """bfhipMatMulDevice: Y = A X on resident device buffers."""
    rc = lib().bfhipMatMulDevice(fac.ptr, dx.ptr, dy.ptr, nrhs, PRECISION[precision])
    if rc:
        check()
        raise BfError(f"apply failed ({rc})")


def mul_device_no_alloc(fac, dx, dy, nrhs, precision="fp64"):
    """bfhipMatMulDeviceNoAlloc: the same apply with the intermediate vectors
    in the library's fixed device scratch (no allocation)."""
    rc = lib().bfhipMatMulDeviceNoAlloc(fac.ptr, dx.ptr, dy.ptr, nrhs, PRECISION[precision])
    if rc:
        check()
        raise BfError(f"apply failed ({rc})")


def kernel_times():
    """Last measured device ms: [sampling, compression, apply fp64, apply fp32]."""
    out = np.zeros(8)
    lib().bfhipGetKernelTimes(_dp(out))
    return out


def kernel_time_sums():
    """(summed ms, counts) per slot since reset_kernel_times()."""
    s, c = np.zeros(8), np.zeros(8)
    lib().bfhipGetKernelTimeSums(_dp(s), _dp(c))
    return s, c


def reset_kernel_times():
    lib().bfhipResetKernelTimes()


class BoundaryOperator(Mat):
    """bfhipHelm2BieNew: the second-kind operator alpha I + (S + C) W of a
    closed curve on the device (helm2_bie.c:109-121): S the device
    factorization `fac` (sources = targets = the curve), C the order-`kr_order`
    Kapur-Rokhlin corrections along the curve's point order, W the weights
    (original point order).  Vectors are in tree order; `tree.perm` maps."""

    def __init__(self, fac, helm, tree, weights, kr_order=6, alpha=0.5):
        self.weights = np.ascontiguousarray(weights, dtype=np.float64)
        q = QuadtreeStruct.from_address(tree.ptr)
        perm = C.addressof(q.super.perm.contents)
        p = lib().bfhipHelm2BieNew(fac.ptr, C.byref(helm.s), tree._pts, _dp(self.weights), perm,
                                   kr_order, float(alpha))
        super().__init__(p, keep=(fac, tree))

    def mul_device(self, dx, dy, nrhs, precision="fp64"):
        """bfhipBieMulDevice: Y = A X on resident device buffers."""
        rc = lib().bfhipBieMulDevice(self.ptr, dx.ptr, dy.ptr, nrhs, PRECISION[precision])
        if rc:
            check()
            raise BfError(f"boundary apply failed ({rc})")

    def solve(self, b, tol, max_iter, x0=None):
        """bfSolveGMRES (linalg.c:47-317): unrestarted GMRES per column of b
        (tree order); returns (x, iterations)."""
        b = np.ascontiguousarray(b, dtype=np.complex128)
        squeeze = b.ndim == 1
        if squeeze:
            b = b[:, None]
        L = lib()
        bm = L.bfMatDenseComplexNew()
        L.bfMatDenseComplexInitViewFromPtr(bm, b.shape[0], b.shape[1], b.ctypes.data)
        xm = None
        if x0 is not None:
            x0 = np.ascontiguousarray(x0, dtype=np.complex128).reshape(b.shape)
            xm = L.bfMatDenseComplexNew()
            L.bfMatDenseComplexInitViewFromPtr(xm, x0.shape[0], x0.shape[1], x0.ctypes.data)
        it = C.c_size_t(0)
        x = L.bfSolveGMRES(self.ptr, bm, xm, float(tol), int(max_iter), C.byref(it), None)
        L.bfMatDelete(C.byref(C.c_void_p(bm)))
        if xm:
            L.bfMatDelete(C.byref(C.c_void_p(xm)))
        check()
        out = Mat(x).to_numpy()
        return (out[:, 0] if squeeze else out), it.value

    def solve_mixed(self, b, tol, restart=50, max_iter=1000):
        """bfhipSolveGmresMixed: FP64 residuals around GMRES(restart) with the
        FP32 apply (prepare_mixed must have run on the factorization).
        Returns (x, {"converged", "iterations", "refinements", "rel_res"})."""
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.zeros_like(b)
        it, nref, rel = C.c_size_t(0), C.c_size_t(0), C.c_double(0)
        rc = lib().bfhipSolveGmresMixed(self.ptr, _dp(b.view(np.float64)), _dp(x.view(np.float64)), float(tol),
                                        int(restart), int(max_iter), C.byref(it), C.byref(nref), C.byref(rel))
        if rc < 0:
            check()
            raise BfError("mixed GMRES failed")
        return x, {"converged": rc == 0, "iterations": it.value, "refinements": nref.value, "rel_res": rel.value}
