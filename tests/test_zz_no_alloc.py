"""The allocation-free apply (bfhipMatMulDeviceNoAlloc) against the
allocating one (bfhipMatMulDevice), GPU.

Both run the same stage kernels on the same factors and inputs; only where
the intermediate vectors live differs (the library's fixed device scratch
array instead of buffers it allocates), so the outputs must be identical.
Two calls in a row, the second with a smaller problem, check that nothing
of the first call's scratch contents leaks into the second.
"""
import numpy as np
import pytest

from conftest import box_problem, curve_problem

pytestmark = pytest.mark.gpu


def _apply_both(bf, F, x, nrhs, prec, dtype):
    n_out = F.shape[0]
    dx = bf.DeviceArray.from_numpy(np.ascontiguousarray(x, dtype))
    esz = np.dtype(dtype).itemsize
    dy0, dy1 = bf.DeviceArray(esz * nrhs * n_out), bf.DeviceArray(esz * nrhs * n_out)
    bf.mul_device(F, dx, dy0, nrhs, prec)
    bf.mul_device_no_alloc(F, dx, dy1, nrhs, prec)
    shape = (n_out, nrhs) if nrhs > 1 else n_out
    return dy0.to_numpy(dtype, shape), dy1.to_numpy(dtype, shape)


def test_no_alloc_apply_equals_the_allocating_apply(gpu_lib):
    bf = gpu_lib
    rng = np.random.default_rng(21)
    # first call: a curve problem, 4 right-hand sides, FP64
    src, tgt, k = curve_problem(8192)
    ts, tt = bf.Quadtree(src, 64), bf.Quadtree(tgt, 64)
    F1 = bf.FacHelm2.make_multilevel(bf.Helm2(k), ts, tt, bf.options(tol=1e-8, pass_through=True, rank_const=1.25,
                                                                     pinv_rcond=1e-11))
    X = rng.standard_normal((len(src), 4)) + 1j * rng.standard_normal((len(src), 4))
    a, b = _apply_both(bf, F1, X, 4, "fp64", np.complex128)
    print("call 1 (N=8192, 4 RHS): max |diff|", np.abs(a - b).max())
    assert np.array_equal(a, b) and np.abs(a).max() > 0
    # second call: a smaller problem, 1 right-hand side, FP64 then FP32
    src, tgt, k = box_problem(2048)
    ts2, tt2 = bf.Quadtree(src, 1), bf.Quadtree(tgt, 1)
    F2 = bf.FacHelm2.make_multilevel(bf.Helm2(k), ts2, tt2)
    x = rng.standard_normal(len(src)) + 1j * rng.standard_normal(len(src))
    a, b = _apply_both(bf, F2, x, 1, "fp64", np.complex128)
    print("call 2 (N=2048, 1 RHS): max |diff|", np.abs(a - b).max())
    assert np.array_equal(a, b)
    assert np.array_equal(b, F2 @ x)
    F2.prepare_mixed("fp32")
    a, b = _apply_both(bf, F2, x, 1, "fp32", np.complex64)
    assert np.array_equal(a, b)


def test_no_alloc_apply_refuses_what_exceeds_the_scratch(gpu_lib):
    bf = gpu_lib
    src, tgt, k = box_problem(2048)
    ts, tt = bf.Quadtree(src, 1), bf.Quadtree(tgt, 1)
    F = bf.FacHelm2.make_multilevel(bf.Helm2(k), ts, tt)
    # one intermediate row of this many FP64 right-hand sides exceeds a
    # 32 MiB half of the scratch: refused on the host before any launch
    nrhs = (32 << 20) // 16 + 1
    dx = bf.DeviceArray(16)
    dy = bf.DeviceArray(16)
    with pytest.raises(bf.BfError, match="OUT_OF_RANGE"):
        bf.mul_device_no_alloc(F, dx, dy, nrhs, "fp64")
