"""GPU parity of the operator kernels against the FP64 oracle (through the C ABI)."""
import numpy as np
import pytest

from tests.helpers import rel_err, small_case, split_images

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from sparc_amd import gpu
    gpu.load()
    return gpu


CASES = [
    ((14, 13, 12), (0, 0, 0)),   # periodic, non-cubic
    ((14, 13, 12), (1, 1, 1)),   # Dirichlet
    ((13, 17, 9), (0, 1, 0)),    # mixed
    ((40, 19, 8), (0, 0, 0)),    # several x tiles, partial y tile, nz < 2*radius+...
    ((6, 6, 6), (0, 0, 0)),      # smallest periodic grid allowed (n = radius)
]


@pytest.mark.parametrize("n,bc", CASES)
def test_lap_plus_diag_fp64(gpu, oracle_mod, n, bc):
    g, veff, nl, rng = small_case(n=n, bc=bc, seed=1)
    o = oracle_mod.Oracle(g, veff)
    x = rng.standard_normal((g.nd, 3))
    with gpu.Context(g) as ctx:
        for a, b, c in ((1.0, 0.0, 0.0), (-0.5, 1.0, -0.7), (2.0, 0.5, 3.0)):
            y = ctx.Lap_plus_diag_vec_mult(x, a, b, c, veff)
            yo = o.lap_plus_diag(x, a, b, c, veff)
            assert rel_err(y, yo) < 1e-13, (a, b, c)


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4, 7, 8, 9, 10, 11, 12])
@pytest.mark.parametrize("n,bc", [((40, 35, 13), (0, 0, 0)), ((70, 33, 9), (0, 1, 1)),
                                  ((7, 6, 30), (0, 0, 0))])
def test_stencil_tile_shapes(gpu, oracle_mod, monkeypatch, cfg, n, bc):
    """Every radius-6 tile shape against the oracle (partial tiles, wrap, walls)."""
    monkeypatch.setenv("SPARC_AMD_STENCIL_TILE", str(cfg))
    g, veff, nl, rng = small_case(n=n, bc=bc, seed=30 + cfg)
    o = oracle_mod.Oracle(g, veff)
    x = rng.standard_normal((g.nd, 2))
    with gpu.Context(g) as ctx:
        y = ctx.Lap_plus_diag_vec_mult(x, -0.5, 1.0, 0.4, veff)
    assert rel_err(y, o.lap_plus_diag(x, -0.5, 1.0, 0.4, veff)) < 1e-13


@pytest.mark.parametrize("cfg,n,bc", [(13, (96, 21, 40), (0, 0, 0)), (13, (96, 21, 10), (1, 0, 1)),
                                      (14, (80, 19, 33), (0, 1, 0)), (14, (80, 37, 17), (1, 1, 0))])
def test_stencil_whole_row_tiles(gpu, oracle_mod, monkeypatch, cfg, n, bc):
    """Whole-row FP32 tiles (x halo from the block's own rows, y halo queued
    kR planes ahead; z-chunks of 20 / 17 planes at nz 40 / 33): bitwise the
    32x16x2 tile, and the FP32 result against the FP64 oracle."""
    g, veff, nl, rng = small_case(n=n, bc=bc, seed=70 + cfg)
    o = oracle_mod.Oracle(g, veff)
    x = rng.uniform(-0.5, 0.5, (g.nd, 2))
    z = rng.uniform(-0.5, 0.5, (g.nd, 2))
    outs = {}
    for c in (cfg, 0):
        monkeypatch.setenv("SPARC_AMD_STENCIL_TILE", str(c))
        with gpu.Context(g) as ctx:
            ctx.set_veff(veff)
            bx = ctx.block(2, gpu.F32).upload(x)
            bz = ctx.block(2, gpu.F32).upload(z)
            ctx.chebyshev_filter_dev(bx, bz, 2, 2.0, 9.0, -1.0)
            outs[c] = (bx.download(), bz.download())
            name = ctx.stencil_kernel(gpu.F32, 2)
        assert name == ("stencil_rows<%d,16,2>" % n[0] if c == cfg else "stencil_tiled<float,32,16,2>")
    assert np.array_equal(outs[cfg][0], outs[0][0]) and np.array_equal(outs[cfg][1], outs[0][1])
    monkeypatch.setenv("SPARC_AMD_STENCIL_TILE", str(cfg))
    with gpu.Context(g) as ctx:
        y = ctx.Lap_plus_diag_vec_mult(x, -0.5, 1.0, 0.4, veff)
        assert ctx.stencil_kernel(gpu.F64, 2) == "stencil_tiled<double,%d,16,2>" % n[0]
    assert rel_err(y, o.lap_plus_diag(x, -0.5, 1.0, 0.4, veff)) < 1e-13


@pytest.mark.parametrize("dtype", [0, 1])
def test_stencil_tile_tuning_is_bitwise_neutral(gpu, monkeypatch, dtype):
    """The default tile choice (timed once per shape) gives bitwise the output
    of the fixed 32x16x2 tile, including the in-place recurrence form
    (out aliasing z) that the tuning must not apply twice."""
    g, veff, nl, rng = small_case(n=(40, 35, 13), seed=50)
    x = rng.uniform(-0.5, 0.5, (g.nd, 6))
    z = rng.uniform(-0.5, 0.5, (g.nd, 6))
    outs = {}
    for tile in (None, "0"):
        if tile is None:
            monkeypatch.delenv("SPARC_AMD_STENCIL_TILE", raising=False)
        else:
            monkeypatch.setenv("SPARC_AMD_STENCIL_TILE", tile)
        with gpu.Context(g) as ctx:
            ctx.set_veff(veff)
            bx = ctx.block(6, dtype).upload(x)
            bz = ctx.block(6, dtype).upload(z)
            ctx.chebyshev_filter_dev(bx, bz, 3, 2.0, 9.0, -1.0)
            outs[tile] = (bx.download(), bz.download())
            shape = ctx.stencil_tile(dtype, 6)
        if tile is None:
            assert shape in ((32, 16, 2), (32, 32, 2), (64, 16, 2))
        else:
            assert shape == (32, 16, 2)
    assert np.array_equal(outs[None][0], outs["0"][0])
    assert np.array_equal(outs[None][1], outs["0"][1])


@pytest.mark.parametrize("cfg", [5, 6])
@pytest.mark.parametrize("n,bc,ncol", [((40, 35, 13), (0, 0, 0), 3), ((70, 33, 9), (0, 1, 1), 2),
                                       ((7, 6, 30), (0, 0, 0), 1), ((33, 17, 40), (1, 0, 0), 1)])
def test_stencil_ring_fp32(gpu, oracle_mod, monkeypatch, cfg, n, bc, ncol):
    """The LDS-ring stencil (FP32 blocks) against the oracle: partial tiles,
    wrap, walls, and z-chunked single columns."""
    monkeypatch.setenv("SPARC_AMD_STENCIL_TILE", str(cfg))
    g, veff, nl, rng = small_case(n=n, bc=bc, seed=40 + cfg)
    o = oracle_mod.Oracle(g, veff)
    x = rng.uniform(-0.5, 0.5, (g.nd, ncol))
    with gpu.Context(g) as ctx:
        ctx.set_veff(veff)
        bx = ctx.block(ncol, gpu.F32).upload(x)
        by = ctx.block(ncol, gpu.F32)
        ctx.lap_plus_diag_dev(-0.5, 1.0, 0.4, bx, by)
        y = by.download()
        y64 = ctx.Lap_plus_diag_vec_mult(x, -0.5, 1.0, 0.4, veff)
    x32 = x.astype(np.float32).astype(np.float64)
    assert rel_err(y, o.lap_plus_diag(x32, -0.5, 1.0, 0.4, veff)) < 3e-6
    assert rel_err(y64, o.lap_plus_diag(x, -0.5, 1.0, 0.4, veff)) < 1e-13


@pytest.mark.parametrize("order", [2, 4, 8, 16])
def test_lap_other_orders(gpu, oracle_mod, order):
    g, veff, nl, rng = small_case(n=(12, 11, 10), order=order, seed=2)
    o = oracle_mod.Oracle(g, veff)
    x = rng.standard_normal((g.nd, 2))
    with gpu.Context(g) as ctx:
        y = ctx.Lap_vec_mult(x, 0.25)
        assert rel_err(y, o.lap_vec_mult(x, 0.25)) < 1e-13


@pytest.mark.parametrize("images", [False, True])
def test_vnl_fp64(gpu, oracle_mod, images):
    g, veff, nl, rng = small_case(n=(16, 15, 14), n_sites=5, rc=2.6, seed=3, lmax=2)
    if images:
        nl = split_images(nl, 1)
    o = oracle_mod.Oracle(g, veff, nl)
    x = rng.standard_normal((g.nd, 11))
    h0 = rng.standard_normal((g.nd, 11))
    with gpu.Context(g) as ctx:
        ctx.set_nonlocal(nl)
        y = ctx.Vnl_vec_mult(x, h0)
    yo = o.vnl_vec_mult(x, h0)
    assert rel_err(y - h0, yo - h0) < 1e-12


def test_vnl_matches_dense_operator(gpu):
    g, veff, nl, rng = small_case(n=(9, 8, 7), n_sites=2, rc=1.6, seed=4)
    V = nl.dense(g.nd)
    assert np.allclose(V, V.T)
    x = rng.standard_normal((g.nd, 4))
    with gpu.Context(g) as ctx:
        ctx.set_nonlocal(nl)
        y = ctx.Vnl_vec_mult(x)
    assert rel_err(y, V @ x) < 1e-12


@pytest.mark.parametrize("bc", [(0, 0, 0), (1, 1, 1)])
def test_hamiltonian_fp64(gpu, oracle_mod, bc):
    g, veff, nl, rng = small_case(n=(18, 17, 16), bc=bc, n_sites=4, seed=5)
    o = oracle_mod.Oracle(g, veff, nl)
    x = rng.standard_normal((g.nd, 7))
    with gpu.Context(g) as ctx:
        ctx.set_veff(veff)
        ctx.set_nonlocal(nl)
        y = ctx.Hamiltonian_vectors_mult(x, -0.3)
    assert rel_err(y, o.hamiltonian(x, -0.3)) < 1e-12


def test_hamiltonian_fp32_block(gpu, oracle_mod):
    g, veff, nl, rng = small_case(n=(20, 18, 16), n_sites=4, seed=6)
    o = oracle_mod.Oracle(g, veff, nl)
    x = rng.uniform(-0.5, 0.5, (g.nd, 9))
    with gpu.Context(g) as ctx:
        ctx.set_veff(veff)
        ctx.set_nonlocal(nl)
        bx = ctx.block(9, gpu.F32).upload(x)
        by = ctx.block(9, gpu.F32)
        ctx.hamiltonian_dev(0.1, bx, by)
        y = by.download()
    assert rel_err(y, o.hamiltonian(x.astype(np.float32).astype(np.float64), 0.1)) < 2e-6


def test_uniform_fill_matches_host_hash(gpu):
    from sparc_amd.workloads import Grid, uniform_block
    g = Grid(10, 9, 8, 0.5, 0.5, 0.5)
    with gpu.Context(g) as ctx:
        b = ctx.block(5, gpu.F64).fill_uniform(1234, -0.5, 0.5)
        d = b.download()
        b32 = ctx.block(5, gpu.F32).fill_uniform(1234, -0.5, 0.5)
        d32 = b32.download()
    h = uniform_block(g.nd, 5, 1234, -0.5, 0.5)
    assert np.array_equal(d, h)
    assert np.array_equal(d32, h.astype(np.float32).astype(np.float64))


@pytest.fixture
def nl_mode(monkeypatch):
    """Force the nonlocal layout (read when the projectors are set)."""
    def set_mode(mode):
        monkeypatch.setenv("SPARC_AMD_NL_MODE", mode)
    return set_mode


@pytest.mark.parametrize("mode", ["brick", "sphere"])
@pytest.mark.parametrize("images", [False, True])
def test_vnl_layouts_fp64(gpu, oracle_mod, nl_mode, mode, images):
    """Both projector layouts against the oracle, including atoms whose
    periodic images share alpha rows and overlapping spheres."""
    nl_mode(mode)
    g, veff, nl, rng = small_case(n=(16, 15, 14), n_sites=6, rc=2.6, seed=13, lmax=2)
    if images:
        nl = split_images(nl, 1)
    o = oracle_mod.Oracle(g, veff, nl)
    x = rng.standard_normal((g.nd, 37))
    h0 = rng.standard_normal((g.nd, 37))
    with gpu.Context(g) as ctx:
        ctx.set_nonlocal(nl)
        y = ctx.Vnl_vec_mult(x, h0)
    yo = o.vnl_vec_mult(x, h0)
    assert rel_err(y - h0, yo - h0) < 1e-12


@pytest.mark.parametrize("mode", ["brick", "sphere"])
def test_vnl_psp_projectors(gpu, oracle_mod, nl_mode, mode):
    """Real psp8 projectors (SiC cell: 18 / 8 projectors per atom, images
    across the cell faces) in both layouts, FP64 and FP32 blocks."""
    from sparc_amd.systems import sic8
    nl_mode(mode)
    st, _ = sic8()
    g = st.grid
    rng = np.random.default_rng(17)
    veff = np.zeros(g.nd)
    o = oracle_mod.Oracle(g, veff, st.nloc)
    x = rng.uniform(-0.5, 0.5, (g.nd, 40))
    yo = o.vnl_vec_mult(x)
    with gpu.Context(g) as ctx:
        ctx.set_veff(veff)
        ctx.set_nonlocal(st.nloc)
        y = ctx.Vnl_vec_mult(x)
        bx = ctx.block(40, gpu.F32).upload(x)
        by = ctx.block(40, gpu.F32)
        ctx.hamiltonian_dev(0.0, bx, by)
        y32 = by.download()
    assert rel_err(y, yo) < 1e-12
    x32 = x.astype(np.float32).astype(np.float64)
    assert rel_err(y32, o.hamiltonian(x32)) < 2e-6
