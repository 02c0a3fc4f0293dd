"""Learned AAR preconditioner on the device: forward pass and weight gradient
against the FP64 oracle (oracle/unet.py) on the same weights (tolerances for
BF16 activations: 2e-2 relative), training on harvested Jacobi-AAR pairs,
and AAR with the trained network reaching the Jacobi-AAR solution in fewer
iterations."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bf16(a):
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _d0abs(grid):
    from oracle import oracle as orc
    dx, dy, dz = orc.Oracle(grid).coeffs
    return abs(dx[0] + dy[0] + dz[0])


@pytest.fixture
def grid16():
    from sparc_amd.workloads import Grid
    return Grid(16, 16, 16, 0.4, 0.4, 0.4)


def test_unet_rejects_dirichlet_grid():
    from sparc_amd import gpu
    from sparc_amd.workloads import Grid
    with gpu.Context(Grid(16, 16, 16, 0.4, 0.4, 0.4, bcz=1)) as ctx:
        with pytest.raises(gpu.SparcGPUError, match="periodic"):
            ctx.unet()


def test_unet_initial_is_jacobi_on_resampled_grid():
    """On a grid that is not a multiple of 4 the network runs on the grid
    rounded up (interpolating in and out); the zero network is still exactly
    the Jacobi preconditioner."""
    from sparc_amd import gpu
    from sparc_amd.workloads import Grid
    g = Grid(26, 22, 13, 0.4, 0.4, 0.4)
    rng = np.random.default_rng(3)
    r = rng.standard_normal(g.nd)
    with gpu.Context(g) as ctx:
        f = ctx.unet(seed=5).apply(r)
    assert np.max(np.abs(f - r / _d0abs(g))) < 1e-14 * np.max(np.abs(r))


def test_unet_initial_is_jacobi(grid16):
    from sparc_amd import gpu
    rng = np.random.default_rng(3)
    r = rng.standard_normal(grid16.nd)
    with gpu.Context(grid16) as ctx:
        f = ctx.unet(seed=5).apply(r)
    assert np.max(np.abs(f - r / _d0abs(grid16))) < 1e-14 * np.max(np.abs(r))


def test_unet_matrix_core_and_fma_layers_agree(grid16, monkeypatch):
    """The MFMA layer kernels and the FP32-FMA layer kernels (selected with
    SPARC_AMD_UNET_SCALAR=1 when the network is created) give the same
    preconditioner up to FP32 summation order."""
    from sparc_amd import gpu
    rng = np.random.default_rng(6)
    r = rng.standard_normal(grid16.nd)
    R = rng.standard_normal((1, grid16.nd))
    E = rng.standard_normal((1, grid16.nd)) * 0.05
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("SPARC_AMD_UNET_SCALAR", flag)
        with gpu.Context(grid16) as ctx:
            net = ctx.unet(seed=9)
            net.set_weights((np.random.default_rng(8).standard_normal(net.nparams) * 0.15).astype(np.float32))
            out[flag] = (net.apply(r), net.loss_grad(R, E))
    (f1, (l1, g1)), (f0, (l0, g0)) = out["1"], out["0"]
    assert np.linalg.norm(f1 - f0) < 1e-2 * np.linalg.norm(f1 - r / _d0abs(grid16))
    assert abs(l1 - l0) < 1e-2 * l1
    assert np.linalg.norm(g1 - g0) < 2e-2 * np.linalg.norm(g1)


@pytest.mark.parametrize("shape", [(16, 16, 16), (14, 13, 10)])
def test_unet_forward_and_gradient_match_oracle(shape):
    """Forward and weight gradient against the oracle; (14, 13, 10) runs the
    network on 16 x 16 x 12 with the interpolations in and out (and the
    transpose of the output interpolation in the backward pass)."""
    from oracle import unet as U
    from sparc_amd import gpu
    from sparc_amd.workloads import Grid
    grid16 = Grid(*shape, 0.4, 0.4, 0.4)
    rng = np.random.default_rng(4)
    d0 = _d0abs(grid16)
    zyx = (grid16.nz, grid16.ny, grid16.nx)
    with gpu.Context(grid16) as ctx:
        net = ctx.unet(seed=9)
        w = (rng.standard_normal(net.nparams) * 0.15).astype(np.float32)
        net.set_weights(w)
        r = rng.standard_normal(grid16.nd)
        f = net.apply(r)
        R = rng.standard_normal((2, grid16.nd))
        E = rng.standard_normal((2, grid16.nd)) * 0.05
        loss, g = net.loss_grad(R, E)
    wq = _bf16(w)
    fo, _ = U.forward(wq, _bf16(r).reshape(zyx), d0)
    # the Jacobi part r/|D0| is exact; compare the network part U
    u_gpu = f * d0 - r
    u_ref = fo.ravel() * d0 - _bf16(r)
    assert np.linalg.norm(u_gpu - u_ref) < 2e-2 * np.linalg.norm(u_ref)
    Ro = [_bf16(x).reshape(zyx) for x in R]
    Eo = [x.reshape(zyx) for x in E.astype(np.float32).astype(np.float64)]
    lo, go = U.loss_and_grad(wq, Ro, Eo, d0)
    assert abs(loss - lo) < 2e-2 * lo
    cos = float(g @ go) / (np.linalg.norm(g) * np.linalg.norm(go))
    assert cos > 0.999
    assert np.linalg.norm(g - go) < 5e-2 * np.linalg.norm(go)
    # per layer too (no layer's gradient may be lost in the others' norm)
    off = 0
    for _, _, ci, co in U.layer_shapes():
        sl = slice(off, off + ci * co * 27)
        assert np.linalg.norm(g[sl] - go[sl]) < 6e-2 * np.linalg.norm(go[sl]) + 1e-12
        off += ci * co * 27


@pytest.mark.parametrize("n", [32, 30])
def test_unet_trained_preconditioner_beats_jacobi(n):
    """Train on pairs harvested from Jacobi-AAR Poisson solves on a 32^3
    cell, then solve a different right-hand side on the same grid: the same
    solution (tolerance 1e-8 relative residual for both) in fewer AAR
    iterations.  n = 30 is not a multiple of 4: the network runs on 32^3 and
    is trained through the interpolations."""
    from sparc_amd import gpu
    from sparc_amd.model import gaussian_density
    from sparc_amd.workloads import Grid
    g = Grid(n, n, n, 12.8 / n, 12.8 / n, 12.8 / n)
    rng = np.random.default_rng(11)

    def rhs(seed):
        rr = np.random.default_rng(seed)
        pos = rr.uniform(0, 12.8, (12, 3))
        q = gaussian_density(g, pos, 1.0, 0.9)
        return 4 * np.pi * (q - q.mean())

    with gpu.Context(g) as ctx:
        net = ctx.unet(seed=2)
        for s in (21, 22, 23):
            net.harvest(rhs(s), np.zeros(g.nd), tol=1e-8, max_iter=3000, stride=3, max_pairs=60)
        assert net.npairs > 50
        hist = net.train(steps=300, batch=4, lr=2e-3, seed=3)
        assert hist[-20:].mean() < 0.5 * hist[:5].mean()
        b = rhs(99)
        x0 = rng.standard_normal(g.nd) * 0.0
        xj, itj, rrj, okj = ctx.AAR(x0, b, tol=1e-8, max_iter=3000)
        xu, itu, rru, st = ctx.AAR_unet(net, x0, b, tol=1e-8, max_iter=3000)
    print(f"\nAAR iterations: Jacobi {itj}, U-Net {itu}; final loss {hist[-1]:.3e}")
    assert okj and st == 0 and rru <= 1e-8
    xj -= xj.mean()
    xu -= xu.mean()
    assert np.linalg.norm(xu - xj) < 1e-6 * np.linalg.norm(xj)
    assert itu < itj



def test_unet_lds_staged_layers_match_direct_taps(monkeypatch):
    """The stride-1 layers that stage their input box in LDS give bitwise the
    output of the kernel that fetches every tap from memory (the same MFMAs on
    the same operands), on a grid whose levels leave partial x, y and z
    tiles (24 x 20 x 12 -> 12 x 10 x 6 -> 6 x 5 x 3)."""
    from sparc_amd import gpu
    from sparc_amd.workloads import Grid
    g = Grid(24, 20, 12, 0.4, 0.4, 0.4)
    r = np.random.default_rng(21).standard_normal(g.nd)
    out = {}
    for lds in ("1", "0"):
        monkeypatch.setenv("SPARC_AMD_UNET_LDS", lds)
        with gpu.Context(g) as ctx:
            net = ctx.unet(seed=5)
            w = net.get_weights()
            net.set_weights(w + np.random.default_rng(6).standard_normal(w.size).astype(np.float32)
                            * 0.05)
            out[lds] = net.apply(r)
    assert np.isfinite(out["1"]).all()
    assert np.array_equal(out["1"], out["0"])


def test_context_destroyed_before_unet_and_block(grid16):
    """sgpu_destroy before sgpu_unet_destroy / sgpu_block_destroy (the order a
    C caller with static handles may use): the context releases the network's
    and the block's device memory and orphans both handles; every other call
    on the orphaned network reports SGPU_ERR_STATE, and destroying the handles
    afterwards is safe.  Calls go straight through the C ABI."""
    import ctypes as C
    from sparc_amd import gpu
    lib = gpu.load()
    g = grid16
    ctx = C.c_void_p()
    assert lib.sgpu_create(C.byref(ctx), 0, g.nx, g.ny, g.nz, g.hx, g.hy, g.hz, 0, 0, 0, 12) == 0
    net, blk = C.c_void_p(), C.c_void_p()
    assert lib.sgpu_unet_create(ctx, 16, 3, C.byref(net)) == 0
    assert lib.sgpu_block_create(ctx, 2, gpu.F64, C.byref(blk)) == 0
    n = lib.sgpu_unet_nparams(net)
    w = np.zeros(n, np.float32)
    wp = w.ctypes.data_as(C.POINTER(C.c_float))
    assert lib.sgpu_unet_get_weights(net, wp) == 0
    assert lib.sgpu_destroy(ctx) == 0
    assert lib.sgpu_unet_get_weights(net, wp) == -4          # SGPU_ERR_STATE
    assert b"destroyed" in lib.sgpu_last_error()
    r = np.ones(g.nd)
    f = np.zeros(g.nd)
    assert lib.sgpu_unet_apply(net, r.ctypes.data_as(C.POINTER(C.c_double)),
                               f.ctypes.data_as(C.POINTER(C.c_double))) == -4
    lib.sgpu_unet_destroy(net)
    assert lib.sgpu_block_destroy(blk) == 0
    # a fresh context and network work after the orphaned handles are gone
    with gpu.Context(g) as c2:
        f2 = c2.unet(seed=5).apply(r)
    assert np.allclose(f2, r / _d0abs(g), rtol=0, atol=1e-14)
