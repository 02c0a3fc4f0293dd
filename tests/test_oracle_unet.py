"""Pins of the numpy U-Net oracle (oracle/unet.py): gradient against central
finite differences, linearity, the zero-weight network is the Jacobi
preconditioner, and the transposed convolution is the adjoint of the
stride-2 convolution."""
import numpy as np


def test_unet_oracle_gradient_matches_finite_differences():
    from oracle import unet as U
    rng = np.random.default_rng(0)
    C, n = 4, 8
    w = rng.standard_normal(U.n_params(C)) * 0.2
    R = [rng.standard_normal((n, n, n)) for _ in range(2)]
    E = [rng.standard_normal((n, n, n)) for _ in range(2)]
    d0 = 3.7
    _, G = U.loss_and_grad(w, R, E, d0, C)
    for idx in (0, 7, 131, 600, U.n_params(C) // 2, U.n_params(C) - 40, U.n_params(C) - 1):
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[idx] += h
        wm[idx] -= h
        fd = (U.loss_and_grad(wp, R, E, d0, C)[0] - U.loss_and_grad(wm, R, E, d0, C)[0]) / (2 * h)
        assert abs(G[idx] - fd) <= 1e-6 * max(1.0, abs(fd))


def test_unet_oracle_linear_and_jacobi():
    from oracle import unet as U
    rng = np.random.default_rng(1)
    C, n = 4, 8
    w = rng.standard_normal(U.n_params(C))
    r = rng.standard_normal((n, n, n))
    f1, _ = U.forward(w, r, 2.5, C)
    f2, _ = U.forward(w, -3.0 * r, 2.5, C)
    assert np.max(np.abs(f2 + 3.0 * f1)) < 1e-10 * np.max(np.abs(f1))
    f0, _ = U.forward(np.zeros_like(w), r, 2.5, C)
    assert np.array_equal(f0, r / 2.5)


def test_unet_oracle_transposed_conv_is_adjoint():
    from oracle import unet as U
    rng = np.random.default_rng(2)
    W = rng.standard_normal((3, 2, 3, 3, 3))
    x = rng.standard_normal((2, 4, 4, 4))
    y = rng.standard_normal((3, 8, 8, 8))
    lhs = (U.convT(x, W) * y).sum()
    rhs = (x * U.conv(y, W.transpose(1, 0, 2, 3, 4), 2)).sum()
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_unet_oracle_resampled_gradient_matches_finite_differences():
    """A grid whose dimensions are not multiples of 4 (10 x 9 x 6, network on
    12 x 12 x 8): the gradient through the interpolations against central
    finite differences."""
    from oracle import unet as U
    rng = np.random.default_rng(5)
    C, shape = 4, (6, 9, 10)
    assert U.internal_shape(shape) == (8, 12, 12)
    w = rng.standard_normal(U.n_params(C)) * 0.2
    R = [rng.standard_normal(shape) for _ in range(2)]
    E = [rng.standard_normal(shape) for _ in range(2)]
    d0 = 2.9
    _, G = U.loss_and_grad(w, R, E, d0, C)
    for idx in (0, 13, 108, 700, U.n_params(C) // 3, U.n_params(C) - 27, U.n_params(C) - 2):
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[idx] += h
        wm[idx] -= h
        fd = (U.loss_and_grad(wp, R, E, d0, C)[0] - U.loss_and_grad(wm, R, E, d0, C)[0]) / (2 * h)
        assert abs(G[idx] - fd) <= 1e-6 * max(1.0, abs(fd))


def test_unet_oracle_interpolation():
    """Periodic linear interpolation: rows sum to one (constants kept), equal
    sizes give the identity, a point that falls on a source point copies it,
    and a periodic linear ramp sampled at the midpoints of a twice-finer
    grid is reproduced away from the wrap."""
    from oracle import unet as U
    for ns, nd in ((26, 28), (28, 26), (78, 80), (9, 12), (5, 5)):
        A = U.interp_matrix(ns, nd)
        assert np.allclose(A.sum(axis=1), 1.0)
        assert np.all(A >= 0)
    assert np.array_equal(U.interp_matrix(7, 7), np.eye(7))
    A = U.interp_matrix(13, 26)   # dst 2k lands on src k
    assert np.array_equal(A[::2], np.eye(13))
    x = np.arange(13, dtype=float)
    y = U.interp_matrix(13, 26) @ x
    assert np.allclose(y[1:-1:2], x[:-1] + 0.5)
    rng = np.random.default_rng(3)
    f = rng.standard_normal((6, 9, 10))
    assert np.allclose(U.resample(np.full((6, 9, 10), 2.5), (8, 12, 12)), 2.5)
    # transpose: <I x, g> = <x, I^T g>
    g = rng.standard_normal((8, 12, 12))
    lhs = (U.resample(f, (8, 12, 12)) * g).sum()
    rhs = (f * U.resample(g, (6, 9, 10), transpose=True)).sum()
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
