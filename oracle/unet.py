"""FP64 numpy restatement of the learned AAR preconditioner (a linear 3-level
3-D convolutional U-Net) with its backward pass.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the device
forward/backward kernels (sparc_amd/csrc/unet.hip).

Network, C channels, periodic 3x3x3 convolutions, levels L0 = N^3, L1 =
(N/2)^3, L2 = (N/4)^3, no activations (a fixed linear preconditioner, as
AAR's theory assumes):
    a1 = R1(r)          conv stride 2, 1 -> C      L0 -> L1
    b1 = B1(a1)         conv,          C -> C      L1
    a2 = R2(b1)         conv stride 2, C -> C      L1 -> L2
    b2 = B2(a2)         conv,          C -> C      L2
    c2 = M2(b2)         conv,          C -> C      L2
    s1 = P2(c2) + b1    transposed conv stride 2   L2 -> L1
    c1 = C1(s1)         conv,          C -> C      L1
    U  = P1(c1)         transposed conv stride 2, C -> 1   L1 -> L0
    f  = (r + U) / |D0|,  D0 = D2x[0] + D2y[0] + D2z[0]
so that U = 0 is the reference's Jacobi preconditioner.

Tensors are [C, Z, Y, X] (x fastest); weights W[co, ci, kz, ky, kx].
conv stride s:   out[co, y] = sum W[co, ci, k] in[ci, s y + k - 1]
transposed s=2:  out[co, x] = sum over (y, k) with 2 y + k - 1 = x of W[co, ci, k] in[ci, y]
(indices periodic).

Grids whose dimensions are not multiples of 4: the network runs on the
internal grid M (each dimension rounded up to a multiple of 4) over the same
periodic cell.  The residual is interpolated onto M and the network output
back onto the grid, both by periodic trilinear interpolation (target point j
of an axis sits at s = j n_src / n_dst source points and takes (1 - t, t) of
source points floor(s), floor(s) + 1, t = s - floor(s)):
    f = (r + I_{M->N} U(I_{N->M} r)) / |D0|,
a fixed linear operator again; the backward pass takes dU through the
transpose of I_{M->N}.
"""
from __future__ import annotations

import numpy as np

LAYERS = [  # name, kind, cin, cout
    ("R1", "conv2", 1, None), ("B1", "conv1", None, None), ("R2", "conv2", None, None),
    ("B2", "conv1", None, None), ("M2", "conv1", None, None), ("P2", "convT", None, None),
    ("C1", "conv1", None, None), ("P1", "convT", None, 1),
]


def layer_shapes(C: int = 16):
    out = []
    for name, kind, cin, cout in LAYERS:
        out.append((name, kind, cin or C, cout or C))
    return out


def n_params(C: int = 16) -> int:
    return sum(ci * co * 27 for _, _, ci, co in layer_shapes(C))


def split_weights(w: np.ndarray, C: int = 16):
    ws, o = [], 0
    for _, _, ci, co in layer_shapes(C):
        ws.append(w[o:o + co * ci * 27].reshape(co, ci, 3, 3, 3))
        o += co * ci * 27
    return ws


def internal_shape(shape):
    """The grid the network runs on: every dimension rounded up to a multiple of 4."""
    return tuple(-(-n // 4) * 4 for n in shape)


def interp_matrix(n_src: int, n_dst: int) -> np.ndarray:
    """[n_dst, n_src] periodic linear interpolation along one axis."""
    A = np.zeros((n_dst, n_src))
    for j in range(n_dst):
        i0, rem = divmod(j * n_src, n_dst)
        t = rem / n_dst
        A[j, i0] += 1.0 - t
        A[j, (i0 + 1) % n_src] += t
    return A


def resample(x, shape_dst, transpose=False):
    """Trilinear periodic interpolation of x [Z, Y, X] onto shape_dst; with
    transpose, the transpose of the interpolation from shape_dst to x.shape."""
    out = x
    for ax in range(3):
        n_src, n_dst = x.shape[ax], shape_dst[ax]
        A = interp_matrix(n_dst, n_src).T if transpose else interp_matrix(n_src, n_dst)
        out = np.moveaxis(np.tensordot(A, out, axes=([1], [ax])), 0, ax)
    return out


def _shift(a, dz, dy, dx):
    return np.roll(a, shift=(-dz, -dy, -dx), axis=(-3, -2, -1))


def conv(x, W, stride):
    """x [ci, nz, ny, nx] -> [co, nz/s, ny/s, nx/s]."""
    co = W.shape[0]
    out = np.zeros((co,) + tuple(n // stride for n in x.shape[1:]))
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                xs = _shift(x, kz - 1, ky - 1, kx - 1)
                if stride == 2:
                    xs = xs[:, ::2, ::2, ::2]
                out += np.einsum("oc,czyx->ozyx", W[:, :, kz, ky, kx], xs)
    return out


def convT(x, W):
    """Transposed stride-2 conv: x [ci, nz, ny, nx] -> [co, 2nz, 2ny, 2nx]."""
    co = W.shape[0]
    fine = tuple(2 * n for n in x.shape[1:])
    out = np.zeros((co,) + fine)
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                up = np.zeros((co,) + fine)
                up[:, ::2, ::2, ::2] = np.einsum("oc,czyx->ozyx", W[:, :, kz, ky, kx], x)
                # value at fine 2y lands at fine 2y + k - 1
                out += np.roll(up, shift=(kz - 1, ky - 1, kx - 1), axis=(1, 2, 3))
    return out


def conv_wgrad(x, g, stride, shape):
    """dL/dW for out = conv(x, W, stride) given g = dL/dout."""
    dW = np.zeros(shape)
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                xs = _shift(x, kz - 1, ky - 1, kx - 1)
                if stride == 2:
                    xs = xs[:, ::2, ::2, ::2]
                dW[:, :, kz, ky, kx] = np.einsum("ozyx,czyx->oc", g, xs)
    return dW


def conv_dgrad(g, W, stride):
    """dL/dx for out = conv(x, W, stride)."""
    ci = W.shape[1]
    shape_in = tuple(stride * n for n in g.shape[1:])
    dx = np.zeros((ci,) + shape_in)
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                t = np.einsum("oc,ozyx->czyx", W[:, :, kz, ky, kx], g)
                if stride == 2:
                    up = np.zeros((ci,) + shape_in)
                    up[:, ::2, ::2, ::2] = t
                    t = up
                dx += np.roll(t, shift=(kz - 1, ky - 1, kx - 1), axis=(1, 2, 3))
    return dx


def convT_wgrad(x, g, shape):
    """dL/dW for out = convT(x, W): out[co, 2y + k - 1] += W[co, ci, k] x[ci, y]."""
    dW = np.zeros(shape)
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                gs = _shift(g, kz - 1, ky - 1, kx - 1)[:, ::2, ::2, ::2]
                dW[:, :, kz, ky, kx] = np.einsum("ozyx,czyx->oc", gs, x)
    return dW


def convT_dgrad(g, W):
    ci = W.shape[1]
    dx = np.zeros((ci,) + tuple(s // 2 for s in g.shape[1:]))
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                gs = _shift(g, kz - 1, ky - 1, kx - 1)[:, ::2, ::2, ::2]
                dx += np.einsum("oc,ozyx->czyx", W[:, :, kz, ky, kx], gs)
    return dx


def forward(w, r, d0abs, C=16):
    """r [nz, ny, nx] -> (f, cache)."""
    W = split_weights(w, C)
    m = internal_shape(r.shape)
    x0 = (r if m == r.shape else resample(r, m))[None]
    a1 = conv(x0, W[0], 2)
    b1 = conv(a1, W[1], 1)
    a2 = conv(b1, W[2], 2)
    b2 = conv(a2, W[3], 1)
    c2 = conv(b2, W[4], 1)
    s1 = convT(c2, W[5]) + b1
    c1 = conv(s1, W[6], 1)
    Um = convT(c1, W[7])[0]
    U = Um if m == r.shape else resample(Um, r.shape)
    f = (r + U) / d0abs
    return f, dict(x0=x0, a1=a1, b1=b1, a2=a2, b2=b2, c2=c2, s1=s1, c1=c1, U=U, shape=r.shape)


def backward(w, cache, df, d0abs, C=16):
    """Gradient of a scalar loss w.r.t. the flat weights, given df = dL/df."""
    W = split_weights(w, C)
    m = cache["x0"].shape[1:]
    dUn = df / d0abs
    dU = (dUn if m == dUn.shape else resample(dUn, m, transpose=True))[None]
    g = [None] * 8
    g[7] = convT_wgrad(cache["c1"], dU, W[7].shape)
    dc1 = convT_dgrad(dU, W[7])
    g[6] = conv_wgrad(cache["s1"], dc1, 1, W[6].shape)
    ds1 = conv_dgrad(dc1, W[6], 1)
    g[5] = convT_wgrad(cache["c2"], ds1, W[5].shape)
    dc2 = convT_dgrad(ds1, W[5])
    db1 = ds1.copy()
    g[4] = conv_wgrad(cache["b2"], dc2, 1, W[4].shape)
    db2 = conv_dgrad(dc2, W[4], 1)
    g[3] = conv_wgrad(cache["a2"], db2, 1, W[3].shape)
    da2 = conv_dgrad(db2, W[3], 1)
    g[2] = conv_wgrad(cache["b1"], da2, 2, W[2].shape)
    db1 += conv_dgrad(da2, W[2], 2)
    g[1] = conv_wgrad(cache["a1"], db1, 1, W[1].shape)
    da1 = conv_dgrad(db1, W[1], 1)
    g[0] = conv_wgrad(cache["x0"], da1, 2, W[0].shape)
    return np.concatenate([x.ravel() for x in g])


def loss_and_grad(w, R, E, d0abs, C=16):
    """L = mean_i |f(r_i) - e_i|^2 over the batch (pairs pre-normalised)."""
    L = 0.0
    G = np.zeros_like(w)
    for r, e in zip(R, E):
        f, cache = forward(w, r, d0abs, C)
        d = f - e
        L += float((d * d).sum())
        G += backward(w, cache, 2.0 * d, d0abs, C)
    return L / len(R), G / len(R)
