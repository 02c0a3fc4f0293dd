"""sgpu_hamiltonian_dev_noalloc: the Hamiltonian apply that allocates no device
memory (its nonlocal partials live in the library's fixed per-GPU array).

Through the C ABI, in both projector layouts and both precisions, two calls in
a row (the second with fewer columns, so stale partials of the first are left
in the array) give bit for bit the output of sgpu_hamiltonian_dev on the same
block, and the FP64 result matches the oracle's Hamiltonian to 1e-12 relative
(FP32 to 1e-5).  A block whose partials would not fit in the array is refused
with an error before anything runs."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(nsites, rng, n=(24, 22, 20)):
    from sparc_amd.workloads import Grid, gaussian_projectors, gaussian_veff
    g = Grid(*n, 0.4, 0.4, 0.4)
    veff = gaussian_veff(g, 6, rng)
    sites = rng.uniform(0, 1, size=(nsites, 3)) * np.asarray(g.lengths)
    nl = gaussian_projectors(g, sites, 2.5, sigmas=(1.0, 0.7), gammas=(1.0, -1.0), lmax=1)
    return g, veff, nl


class _nl_mode:
    """SPARC_AMD_NL_MODE for the sgpu_set_nonlocal calls inside the block."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        self.old = os.environ.get("SPARC_AMD_NL_MODE")
        os.environ["SPARC_AMD_NL_MODE"] = self.mode

    def __exit__(self, *exc):
        if self.old is None:
            os.environ.pop("SPARC_AMD_NL_MODE", None)
        else:
            os.environ["SPARC_AMD_NL_MODE"] = self.old


@pytest.mark.parametrize("mode", ["brick", "sphere"])
def test_noalloc_hamiltonian_matches_allocating_entry(mode):
    from oracle import oracle as O
    from sparc_amd import gpu
    g, veff, nl = _system(5, np.random.default_rng(11))
    o = O.Oracle(g, veff, nl)
    c = -0.3
    with _nl_mode(mode), gpu.Context(g) as ctx:
        ctx.set_veff(veff)
        ctx.set_nonlocal(nl)
        assert ctx.nonlocal_layout()[0] == (0 if mode == "brick" else 1)
        for dt, tol in ((gpu.F64, 1e-12), (gpu.F32, 1e-5)):
            for ncol, seed in ((24, 3), (9, 4)):
                print(f"\n[noalloc] layout {mode} dtype {dt} ncol {ncol}", flush=True)
                x = ctx.block(ncol, dt).fill_uniform(seed, -0.5, 0.5)
                h_na = ctx.block(ncol, dt).fill_uniform(99, -1.0, 1.0)  # stale output
                h_al = ctx.block(ncol, dt)
                ctx.hamiltonian_dev_noalloc(c, x, h_na)
                ctx.hamiltonian_dev(c, x, h_al)
                a, b = h_na.download(), h_al.download()
                assert np.array_equal(a, b), np.abs(a - b).max()
                ho = o.hamiltonian(x.download(), c)
                err = np.abs(a - ho).max() / np.abs(ho).max()
                assert err < tol, err
                for blk in (x, h_na, h_al):
                    blk.free()


def test_noalloc_hamiltonian_refuses_oversized_block():
    from sparc_amd import gpu
    # 40 sites x 8 projectors in the sphere layout: at least 40 entries x 64
    # rows x 8 bytes = 20 KiB of partials per FP64 column, so 3,400 columns
    # need more than the 64 MiB array
    g, veff, nl = _system(40, np.random.default_rng(5), n=(16, 16, 16))
    with _nl_mode("sphere"), gpu.Context(g) as ctx:
        ctx.set_veff(veff)
        ctx.set_nonlocal(nl)
        assert ctx.nonlocal_layout()[0] == 1
        x = ctx.block(3400, gpu.F64)
        h = ctx.block(3400, gpu.F64)
        with pytest.raises(gpu.SparcGPUError, match="scratch bytes"):
            ctx.hamiltonian_dev_noalloc(0.0, x, h)
        # a block that fits still runs after the refusal
        x2 = ctx.block(8, gpu.F64).fill_uniform(1, -0.5, 0.5)
        h_na, h_al = ctx.block(8, gpu.F64), ctx.block(8, gpu.F64)
        ctx.hamiltonian_dev_noalloc(0.1, x2, h_na)
        ctx.hamiltonian_dev(0.1, x2, h_al)
        assert np.array_equal(h_na.download(), h_al.download())
