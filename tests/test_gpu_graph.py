"""Graph replay of the apply's stage chain and the apply autotuner (GPU).

With graphs on, bfhipMatMulDevice captures the stage launches once and
replays them; every replay must equal the eager apply bit for bit (the same
kernels on the same buffers).  Replays are interleaved with eager work of
the program (another factorization's applies, copies, synchronizations).
The tuner only chooses launch shapes: the applies after tuning agree with
the untuned ones to FP64 rounding (a different lane map sums in another
order), and a second factorization of the same shape reuses the cached
choices.
"""
import numpy as np
import pytest

from conftest import box_problem, curve_problem

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _curve_fac(bf, n):
    src, tgt, k = curve_problem(n)
    ts, tt = bf.Quadtree(src, 64), bf.Quadtree(tgt, 64)
    F = bf.FacHelm2.make_multilevel(bf.Helm2(k), ts, tt, bf.options(tol=1e-8, pass_through=True, rank_const=1.25,
                                                                    pinv_rcond=1e-11))
    return F, len(src), len(tgt)


def test_graph_replays_equal_the_eager_apply(gpu_lib):
    bf = gpu_lib
    F, ns, nt = _curve_fac(bf, 8192)
    src, tgt, k = box_problem(2048)
    G = bf.FacHelm2.make_multilevel(bf.Helm2(k), bf.Quadtree(src, 1), bf.Quadtree(tgt, 1))
    rng = np.random.default_rng(40)
    for nrhs, prec, dt in ((1, "fp64", np.complex128), (16, "fp64", np.complex128), (16, "fp32", np.complex64)):
        if prec == "fp32":
            F.prepare_mixed("fp32")
        X = (rng.standard_normal((ns, nrhs)) + 1j * rng.standard_normal((ns, nrhs))).astype(dt)
        dx = bf.DeviceArray.from_numpy(X)
        dy = bf.DeviceArray(np.dtype(dt).itemsize * nt * nrhs)
        bf.set_apply_graphs(False)
        bf.mul_device(F, dx, dy, nrhs, prec)
        want = dy.to_numpy(dt, (nt, nrhs))
        xg = bf.DeviceArray.from_numpy(rng.standard_normal(len(src)) + 0j)
        yg = bf.DeviceArray(16 * len(tgt))
        try:
            for rep in range(4):
                bf.set_apply_graphs(True)
                bf.lib().bfhipMemcpyHtoD(dy.ptr, np.zeros(nt * nrhs, dt).ctypes.data, np.dtype(dt).itemsize * nt * nrhs)
                bf.mul_device(F, dx, dy, nrhs, prec)
                got = dy.to_numpy(dt, (nt, nrhs))
                assert np.array_equal(got, want), (nrhs, prec, rep)
                # the program's own eager work between the replays
                bf.set_apply_graphs(False)
                bf.mul_device(G, xg, yg, 1, "fp64")
                bf.lib().bfhipSynchronize()
        finally:
            bf.set_apply_graphs(False)


def test_tuner_keeps_the_operator_and_caches_its_choices(gpu_lib, tmp_path):
    bf = gpu_lib
    F, ns, nt = _curve_fac(bf, 8192)
    rng = np.random.default_rng(41)
    cache = tmp_path / "tune.txt"
    for nrhs in (1, 32):
        X = rng.standard_normal((ns, nrhs)) + 1j * rng.standard_normal((ns, nrhs))
        dx = bf.DeviceArray.from_numpy(X)
        dy = bf.DeviceArray(16 * nt * nrhs)
        bf.mul_device(F, dx, dy, nrhs, "fp64")
        before = dy.to_numpy(np.complex128, (nt, nrhs))
        assert all(c == -1 for c in F.apply_config(nrhs))
        F.tune_apply(nrhs, "fp64", cache)
        cfg = F.apply_config(nrhs)
        allowed = {1, 2, 4, 17, 18, 20} if nrhs == 1 else {0, 1, 2, 4, 8, 17, 18, 20}
        assert set(cfg) <= allowed, cfg
        bf.mul_device(F, dx, dy, nrhs, "fp64")
        after = dy.to_numpy(np.complex128, (nt, nrhs))
        assert rel(after, before) < 1e-13
    lines = cache.read_text().split("\n")
    assert len([ln for ln in lines if ln.strip()]) >= 2
    # a second factorization of the same shape finds every bucket cached
    H, _, _ = _curve_fac(bf, 8192)
    text = cache.read_text()
    H.tune_apply(32, "fp64", cache)
    assert cache.read_text() == text and H.apply_config(32) == F.apply_config(32)
