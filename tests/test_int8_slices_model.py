"""The arithmetic of the int8-sliced (Ozaki) products, modelled on the host
with exact integer matrix products: the slice range, the exactness of the
slicing, the int32 bound of a level sum and the error of the fifteen kept
slice products against the FP64 product.  The kernels in ozaki.hip compute the
same quantities on the int8 matrix cores; the GPU tests compare them with the
FP64 path."""
import numpy as np
import pytest

NS = 5  # planes per element


def slices(X):
    """row scale 2^ex (max|row| / 2^ex in [1/4, 1/2)) and NS rounded slices"""
    m = np.abs(X).max(axis=1)
    ex = np.frexp(m)[1] + 1
    scale = np.where(m > 0, np.ldexp(1.0, ex), 0.0)
    t = np.ldexp(X, (7 - ex)[:, None])
    out = []
    for _ in range(NS):
        a = np.rint(t)
        out.append(a.astype(np.int64))
        t = (t - a) * 128.0
    return scale, out


def product(A, Bt):
    """A @ Bt.T from the slice pairs with i + j <= NS - 1 (0-based), exact
    integer level sums, combined in FP64"""
    sa, a = slices(A)
    sb, b = slices(Bt)
    acc = np.zeros((A.shape[0], Bt.shape[0]))
    levels = []
    for lev in range(NS):
        I = sum(a[i] @ b[lev - i].T for i in range(lev + 1))
        levels.append(I)
        acc += I.astype(np.float64) * 2.0 ** (-7 * (lev + 2))
    return acc * sa[:, None] * sb[None, :], levels


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_slices_are_exact_and_in_range(seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (40, 300)) * np.exp(rng.normal(0, 3, (40, 1)))
    X[3] = 0.0  # a zero row
    sc, a = slices(X)
    assert all(np.abs(p).max() <= 64 for p in a)
    # the slices reproduce x to 2^-(7 NS + 1) of the row scale
    rec = sum(p.astype(np.float64) * 2.0 ** (-7 * (i + 1)) for i, p in enumerate(a)) * sc[:, None]
    err = np.abs(rec - X).max(axis=1)
    assert np.all(err <= sc * 2.0 ** (-7 * NS - 1))
    assert sc[3] == 0.0 and np.all(rec[3] == 0.0)


def test_product_error_and_int32_bound():
    rng = np.random.default_rng(7)
    K = 640
    for A, B in [(rng.normal(0, 1, (64, K)), rng.normal(0, 1, (64, K))),
                 (np.abs(rng.normal(0, 1, (64, K))), np.abs(rng.normal(0, 1, (64, K)))),
                 (rng.normal(0, 1, (64, K)) * (rng.random((64, K)) < 0.1), rng.normal(0, 1, (64, K)))]:
        got, levels = product(A, B)
        want = A @ B.T
        bound = np.abs(A).max(axis=1)[:, None] * np.abs(B).max(axis=1)[None, :] * K * 2.0 ** -33
        assert np.all(np.abs(got - want) <= bound)
        # every level sum fits int32 (|a_i b_j| <= 2^12, at most NS pairs per level)
        assert max(np.abs(I).max() for I in levels) < 2 ** 31
        assert NS * 64 * 64 * K < 2 ** 31


def test_rounding_keeps_same_sign_sums_unbiased():
    """with all-positive operands truncated slices would bias every product the
    same way; rounded slices leave the summed error near the per-element error"""
    rng = np.random.default_rng(3)
    A = np.abs(rng.normal(0, 1, (64, 512)))
    B = np.abs(rng.normal(0, 1, (64, 512)))
    got, _ = product(A, B)
    d = got - A @ B.T
    assert abs(d.sum()) < 64 * np.abs(d).max()
