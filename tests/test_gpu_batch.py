"""Independent problems factored as one block-diagonal operator (GPU).

bfhipFacHelm2MakeBatched plans every problem on its own trees, with its own
wavenumber, leaf size and tolerance, and runs all of them through shared
launches.  The checker is the per-problem factorization: the batched plan is
the problems' plans stacked (rows and columns shifted), and the batched
apply equals the per-problem applies on the stacked vectors.  Each
per-problem factorization is itself checked against the oracle and dense
rows elsewhere (test_gpu_parity.py); here the batched result must match it
to rounding (1e-13: the same jobs and sums, in merged launches).
"""
import numpy as np
import pytest

from conftest import box_problem, curve_problem

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def batch(gpu_lib):
    bf = gpu_lib
    probs = []
    for n, leaf, kind in [(2048, 64, "curve"), (1024, 32, "box"), (4096, 128, "curve")]:
        src, tgt, k = curve_problem(n) if kind == "curve" else box_problem(n)
        probs.append((src, tgt, k, leaf))
    trees = [(bf.Quadtree(s, leaf), bf.Quadtree(t, leaf)) for s, t, _, leaf in probs]
    helms = [bf.Helm2(k) for _, _, k, _ in probs]
    opts = [bf.options(tol=1e-8, pass_through=True, rank_const=1.25, pinv_rcond=1e-11) for _ in probs]
    singles = [bf.FacHelm2.make_multilevel(h, ts, tt, o) for h, (ts, tt), o in zip(helms, trees, opts)]
    B = bf.FacHelm2.make_batched(helms, [t[0] for t in trees], [t[1] for t in trees], opts)
    return bf, probs, singles, B


def test_batched_plan_is_the_stacked_plans(batch):
    bf, probs, singles, B = batch
    lt, ft, bt = B.plan()
    parts = [F.plan() for F in singles]
    r0 = c0 = 0
    lts = []
    for (lti, _, _), (src, tgt, _, _) in zip(parts, probs):
        x = lti.copy()
        x[:, 1] += r0
        x[:, 2] += c0
        lts.append(x)
        r0 += len(tgt)
        c0 += len(src)
    assert np.array_equal(lt, np.concatenate(lts))
    assert np.array_equal(ft, np.concatenate([p[1] for p in parts]))
    assert np.array_equal(bt, np.concatenate([p[2] for p in parts]))
    assert B.shape == (r0, c0)


def test_batched_apply_equals_per_problem_applies(batch):
    bf, probs, singles, B = batch
    rng = np.random.default_rng(31)
    xs = [rng.standard_normal((len(s), 3)) + 1j * rng.standard_normal((len(s), 3)) for s, _, _, _ in probs]
    Y = B @ np.concatenate(xs)
    ys = [F @ x for F, x in zip(singles, xs)]
    assert rel(Y, np.concatenate(ys)) < 1e-13
    r0 = 0
    for y in ys:
        assert rel(Y[r0:r0 + len(y)], y) < 1e-13
        r0 += len(y)


def test_batched_fp32_apply_of_16_rhs(batch):
    bf, probs, singles, B = batch
    rng = np.random.default_rng(32)
    n = B.shape[1]
    X = rng.standard_normal((n, 16)) + 1j * rng.standard_normal((n, 16))
    Y64 = B @ X
    B.orthogonalize()
    B.prepare_mixed("fp32")
    dX = bf.DeviceArray.from_numpy(X.astype(np.complex64))
    dY = bf.DeviceArray(8 * 16 * B.shape[0])
    bf.mul_device(B, dX, dY, 16, "fp32")
    Y32 = dY.to_numpy(np.complex64, (B.shape[0], 16)).astype(complex)
    assert rel(Y32, Y64) < 1e-5


def test_batched_rejects_bad_arguments(gpu_lib):
    bf = gpu_lib
    with pytest.raises(ValueError):
        bf.FacHelm2.make_batched([bf.Helm2(1.0)], [], [], [])
