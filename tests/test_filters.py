"""TO_convolve / TO_convolve_transpose (app/ffip.cpp:222-306).

The oracle restatement is pinned by what the mathematics fixes: a constant
array is left unchanged (the window normalisation), a centred delta filter is
the identity, and the transpose is the adjoint of the convolution,
<C x, y> = <x, C^T y>, on random data.  The reference's own check
(scripts/transpose test.py: a random 5x5 filter, d sum(C p) / dp = C^T 1) is
repeated as a finite difference.  The GPU kernels are compared with the oracle
bit for bit: same operations in the same order, no contraction.
"""
import numpy as np
import pytest

from oracle import oracle as O


def test_oracle_constant_and_delta():
    rng = np.random.default_rng(1)
    f = rng.random((5, 5))
    c = np.full((13, 11), 2.5)
    assert np.allclose(O.to_convolve(c, f), 2.5, rtol=1e-14)
    d = np.zeros((3, 3))
    d[1, 1] = 1.0
    x = rng.random((9, 7))
    assert np.array_equal(O.to_convolve(x, d), x)
    assert np.array_equal(O.to_convolve_transpose(x, d), x)


def test_oracle_adjoint_identity_and_reference_check():
    rng = np.random.default_rng(2)
    f = rng.random((5, 5))
    x, y = rng.standard_normal((14, 12)), rng.standard_normal((14, 12))
    lhs = np.sum(O.to_convolve(x, f) * y)
    rhs = np.sum(x * O.to_convolve_transpose(y, f))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # d sum(C p) / d p[i, j] = (C^T 1)[i, j] (linear map, so the difference is exact up to rounding)
    p = np.ones((10, 10))
    g = O.to_convolve_transpose(np.ones((10, 10)), f)
    dp = np.zeros_like(p)
    dp[-1, -1] = 100.0
    fd = (np.sum(O.to_convolve(p + dp, f)) - np.sum(O.to_convolve(p, f))) / 100.0
    assert abs(fd - g[-1, -1]) < 1e-12


def test_argument_errors():
    import ffip_amd as ffip
    with pytest.raises(KeyError):
        ffip.TO_convolve(np.ones(5), np.ones((3, 3)))
    with pytest.raises(KeyError):
        ffip.TO_convolve(np.ones((5, 5)), np.ones((4, 3)))
    with pytest.raises(KeyError):
        ffip.TO_convolve_transpose(np.ones((5, 5)), np.ones((3, 5)))


@pytest.mark.gpu
@pytest.mark.parametrize("shape,fw", [((37, 29), 5), ((16, 40), 7), ((3, 4), 5), ((1, 1), 3)])
def test_gpu_matches_oracle_bitwise(shape, fw):
    import ffip_amd as ffip
    rng = np.random.default_rng(sum(shape) + fw)
    x = rng.standard_normal(shape)
    f = rng.random((fw, fw))
    assert np.array_equal(ffip.TO_convolve(x, f), O.to_convolve(x, f))
    assert np.array_equal(ffip.TO_convolve_transpose(x, f), O.to_convolve_transpose(x, f))
