"""ctypes front end of the oracle (bf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    vp, sz, d, dp, i64p = C.c_void_p, C.c_size_t, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int64)
    sig = {
        "orc_seed": (None, [C.c_uint64]),
        "orc_splitmix_seed": (None, [C.c_uint64]),
        "orc_splitmix_draw": (C.c_uint64, []),
        "orc_next_u64": (C.c_uint64, []),
        "orc_uniform": (d, []),
        "orc_uniform_fill": (None, [sz, dp]),
        "orc_randn_fill": (None, [sz, dp]),
        "orc_sample_box": (None, [sz, d, d, d, d, dp]),
        "orc_rank_est": (sz, [d, d, d, d, d, d, d, d]),
        "orc_h0": (None, [d, dp, dp]),
        "orc_kernel_matrix": (None, [d, sz, dp, sz, dp, dp]),
        "orc_circle_points": (None, [d, d, d, sz, dp]),
        "orc_lstsq": (sz, [sz, sz, dp, sz, dp, dp]),
        "orc_svd_values": (None, [sz, sz, dp, dp]),
        "orc_fac_make": (vp, [d, dp, sz, dp, sz, sz, d, sz, C.c_int]),
        "orc_set_plan_only": (None, [C.c_int]),
        "orc_h1": (None, [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "orc_dense_rows_layer": (None, [C.c_double, C.c_int, dp, dp, C.c_size_t, dp, C.c_size_t, dp, dp, dp, dp,
                                        dp]),
        "orc_kernel_matrix_layer": (None, [C.c_double, C.c_int, dp, dp, C.c_size_t, dp, C.c_size_t, dp, dp, dp,
                                           dp]),
        "orc_set_rank_const": (None, [d]),
        "orc_set_pinv_rcond": (None, [d]),
        "orc_set_impedance": (None, [C.c_int]),
        "orc_set_field_form": (None, [C.c_int]),
        "orc_set_test_oversample": (None, [d]),
        "orc_fac_free": (None, [vp]),
        "orc_fac_perms": (None, [vp, C.POINTER(sz), C.POINTER(sz)]),
        "orc_fac_plan_sizes": (None, [vp, C.POINTER(sz), C.POINTER(sz), C.POINTER(sz)]),
        "orc_fac_plan_export": (None, [vp, i64p, i64p, i64p]),
        "orc_fac_block_data": (C.c_int, [vp, sz, dp]),
        "orc_fac_dense_data": (C.c_int, [vp, sz, dp]),
        "orc_fac_apply": (None, [vp, sz, dp, dp]),
        "orc_dense_rows": (None, [d, sz, dp, sz, dp, dp, dp]),
        "orc_tree_build": (sz, [dp, sz, sz, C.POINTER(sz), i64p, dp, sz]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def seed(s):
    lib().orc_seed(s)


def sample_box(n, x0, x1, y0, y1):
    out = np.zeros((n, 2))
    lib().orc_sample_box(n, x0, x1, y0, y1, _dp(out))
    return out


def complex_randn(n):
    out = np.zeros(2 * n)
    lib().orc_randn_fill(2 * n, _dp(out))
    return out.view(np.complex128)


def h0(x):
    L = lib()
    out = np.zeros(len(x), np.complex128)
    re, im = C.c_double(), C.c_double()
    for i, v in enumerate(x):
        L.orc_h0(float(v), C.byref(re), C.byref(im))
        out[i] = complex(re.value, im.value)
    return out


def kernel_matrix(k, tgt, src):
    tgt = np.ascontiguousarray(tgt, np.float64)
    src = np.ascontiguousarray(src, np.float64)
    out = np.zeros(2 * len(tgt) * len(src))
    lib().orc_kernel_matrix(k, len(tgt), _dp(tgt), len(src), _dp(src), _dp(out))
    return out.view(np.complex128).reshape(len(tgt), len(src))


def circle_points(cx, cy, r, n):
    out = np.zeros((n, 2))
    lib().orc_circle_points(cx, cy, r, n, _dp(out))
    return out


def lstsq(A, B):
    A = np.ascontiguousarray(A, np.complex128)
    B = np.ascontiguousarray(B, np.complex128)
    m, n = A.shape
    X = np.zeros((n, B.shape[1]), np.complex128)
    kept = lib().orc_lstsq(m, n, _dp(A.view(np.float64)), B.shape[1], _dp(B.view(np.float64)),
                           _dp(X.view(np.float64)))
    return X, kept


def svd_values(A):
    A = np.ascontiguousarray(A, np.complex128)
    s = np.zeros(A.shape[1])
    lib().orc_svd_values(A.shape[0], A.shape[1], _dp(A.view(np.float64)), _dp(s))
    return np.sort(s)[::-1]


def tree(points, leaf_size):
    points = np.ascontiguousarray(points, np.float64)
    n = len(points)
    perm = np.zeros(n, np.uintp)
    cnt = lib().orc_tree_build(_dp(points), n, leaf_size, perm.ctypes.data_as(C.POINTER(C.c_size_t)),
                               None, None, 0)
    nodes = np.zeros((cnt, 8), np.int64)
    bbox = np.zeros((cnt, 4))
    lib().orc_tree_build(_dp(points), n, leaf_size, perm.ctypes.data_as(C.POINTER(C.c_size_t)),
                         nodes.ctypes.data_as(C.POINTER(C.c_int64)), _dp(bbox), cnt)
    return perm.astype(np.int64), nodes, bbox


def _layer_args(layer, alpha, beta):
    lay = {"single": 0, "double": 1, "combined": 2, "sprime": 3}[layer]
    a = np.array([complex(alpha).real, complex(alpha).imag])
    b = np.array([complex(beta).real, complex(beta).imag])
    return lay, a, b


def _normals(nrm):
    if nrm is None:
        return None, None
    nrm = np.ascontiguousarray(nrm, np.float64)
    return nrm, _dp(nrm)


def dense_rows_layer(k, tgt, src, src_normals, x, layer="double", alpha=1.0, beta=1.0, tgt_normals=None):
    """Rows of K x for the double layer, the combined field alpha S + beta D
    (source normals) or S' (target normals), helm2.c:38-90, Neumaier-summed."""
    tgt = np.ascontiguousarray(tgt, np.float64)
    src = np.ascontiguousarray(src, np.float64)
    sn, psn = _normals(src_normals)
    tn, ptn = _normals(tgt_normals)
    x = np.ascontiguousarray(x, np.complex128)
    lay, a, b = _layer_args(layer, alpha, beta)
    assert (lay in (1, 2)) <= (sn is not None) and (lay == 3) <= (tn is not None)
    out = np.zeros(2 * len(tgt))
    lib().orc_dense_rows_layer(k, lay, _dp(a), _dp(b), len(tgt), _dp(tgt), len(src), _dp(src), psn, ptn,
                               _dp(x.view(np.float64)), _dp(out))
    return out.view(np.complex128)


def kernel_matrix_layer(k, tgt, src, src_normals, layer="double", alpha=1.0, beta=1.0, tgt_normals=None):
    tgt = np.ascontiguousarray(tgt, np.float64)
    src = np.ascontiguousarray(src, np.float64)
    sn, psn = _normals(src_normals)
    tn, ptn = _normals(tgt_normals)
    lay, a, b = _layer_args(layer, alpha, beta)
    assert (lay in (1, 2)) <= (sn is not None) and (lay == 3) <= (tn is not None)
    out = np.zeros(2 * len(tgt) * len(src))
    lib().orc_kernel_matrix_layer(k, lay, _dp(a), _dp(b), len(tgt), _dp(tgt), len(src), _dp(src), psn, ptn,
                                  _dp(out))
    return out.view(np.complex128).reshape(len(tgt), len(src))


def h1(x):
    x = np.atleast_1d(np.asarray(x, np.float64))
    out = np.empty(len(x), complex)
    re, im = C.c_double(), C.c_double()
    for i, v in enumerate(x):
        lib().orc_h1(float(v), C.byref(re), C.byref(im))
        out[i] = complex(re.value, im.value)
    return out


def dense_rows(k, tgt, src, x):
    tgt = np.ascontiguousarray(tgt, np.float64)
    src = np.ascontiguousarray(src, np.float64)
    x = np.ascontiguousarray(x, np.complex128)
    out = np.zeros(2 * len(tgt))
    lib().orc_dense_rows(k, len(tgt), _dp(tgt), len(src), _dp(src), _dp(x.view(np.float64)), _dp(out))
    return out.view(np.complex128)


class Fac:
    """The oracle's factorization (reference rules, or pass-through)."""

    def __init__(self, k, src, tgt, leaf_size=1, tol=1e-15, max_dense=128 * 128, pass_through=False,
                 plan_only=False, rank_const=1.0, pinv_rcond=0.0, field_form=False):
        self.src = np.ascontiguousarray(src, np.float64)
        self.tgt = np.ascontiguousarray(tgt, np.float64)
        lib().orc_set_plan_only(int(plan_only))
        lib().orc_set_rank_const(rank_const)
        lib().orc_set_pinv_rcond(pinv_rcond)
        lib().orc_set_field_form(int(field_form))
        self.h = lib().orc_fac_make(k, _dp(self.src), len(self.src), _dp(self.tgt), len(self.tgt),
                                    leaf_size, tol, max_dense, int(pass_through))
        lib().orc_set_plan_only(0)
        lib().orc_set_rank_const(1.0)
        lib().orc_set_pinv_rcond(0.0)
        lib().orc_set_field_form(0)

    def perms(self):
        sp = np.zeros(len(self.src), np.uintp)
        tp = np.zeros(len(self.tgt), np.uintp)
        p = C.POINTER(C.c_size_t)
        lib().orc_fac_perms(self.h, sp.ctypes.data_as(p), tp.ctypes.data_as(p))
        return sp.astype(np.int64), tp.astype(np.int64)

    def plan(self):
        a, b, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
        lib().orc_fac_plan_sizes(self.h, C.byref(a), C.byref(b), C.byref(c))
        lt = np.zeros((a.value, 6), np.int64)
        ft = np.zeros((b.value, 4), np.int64)
        bt = np.zeros((c.value, 7), np.int64)
        p = C.POINTER(C.c_int64)
        lib().orc_fac_plan_export(self.h, lt.ctypes.data_as(p), ft.ctypes.data_as(p), bt.ctypes.data_as(p))
        return lt, ft, bt

    def block(self, idx, m, n):
        out = np.zeros(2 * m * n)
        assert lib().orc_fac_block_data(self.h, idx, _dp(out)) == 0
        return out.view(np.complex128).reshape(m, n)

    def dense_leaf(self, idx, m, n):
        out = np.zeros(2 * m * n)
        assert lib().orc_fac_dense_data(self.h, idx, _dp(out)) == 0
        return out.view(np.complex128).reshape(m, n)

    def apply(self, x_tree):
        x = np.ascontiguousarray(x_tree, np.complex128)
        if x.ndim == 1:
            x = x[:, None]
        y = np.zeros((len(self.tgt), x.shape[1]), np.complex128)
        lib().orc_fac_apply(self.h, x.shape[1], _dp(x.view(np.float64)), _dp(y.view(np.float64)))
        return y

    def __del__(self):
        try:
            if self.h:
                lib().orc_fac_free(self.h)
                self.h = None
        except Exception:
            pass
