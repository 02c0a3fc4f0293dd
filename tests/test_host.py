"""Host-side checks that need no GPU: the C-ABI library loads and exports
what include/sparc_gpu.h declares, workload builders, bench contract."""
import ctypes
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def header_symbols():
    text = (ROOT / "include" / "sparc_gpu.h").read_text()
    return sorted(set(re.findall(r"\b(sgpu_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from sparc_amd import gpu
    lib_path = gpu.LIB_PATH
    if not lib_path.exists():
        pytest.fail("libsparcgpu.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(lib_path))  # loading makes no GPU call
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    from sparc_amd import gpu
    assert set(header_symbols()) == set(gpu.exported_symbols())


def test_uniform_block_hash_is_deterministic_and_uniform():
    from sparc_amd.workloads import uniform_block
    a = uniform_block(1000, 3, 42, -0.5, 0.5)
    b = uniform_block(1000, 3, 42, -0.5, 0.5)
    assert np.array_equal(a, b) and a.flags.f_contiguous
    assert -0.5 <= a.min() and a.max() < 0.5 and abs(a.mean()) < 0.03
    assert not np.array_equal(a, uniform_block(1000, 3, 43, -0.5, 0.5))


def test_sphere_points_and_projectors():
    from sparc_amd.workloads import Grid, gaussian_projectors, sphere_points
    g = Grid(40, 40, 40, 0.3, 0.3, 0.3)
    lin, dx, dy, dz = sphere_points(g, (0.1, 11.9, 5.0), 3.0)
    assert np.all(np.diff(lin) > 0)
    assert np.all(dx * dx + dy * dy + dz * dz <= 9.0 + 1e-12)
    assert abs(lin.size - 4 / 3 * np.pi * 27 / 0.027) / lin.size < 0.05
    nl = gaussian_projectors(g, np.array([[1.0, 2.0, 3.0], [6, 6, 6]]), 3.0, (1.0, 0.7), (1.0, -1.0))
    assert nl.n_alpha == 16 and list(nl.nproj) == [8, 8]
    assert list(nl.gamma[:8]) == [1.0, -1.0] + [1.0] * 3 + [-1.0] * 3


def test_sphere_rejects_self_overlap():
    from sparc_amd.workloads import Grid, sphere_points
    g = Grid(10, 10, 10, 0.5, 0.5, 0.5)
    with pytest.raises(ValueError):
        sphere_points(g, (1, 1, 1), 3.0)


def test_algorithmic_bytes_formula():
    import bench
    # operator case: 96^3 x 512, degree 24, 64 x 8 x 33,510 projector values
    b32 = bench.algorithmic_bytes(96 ** 3, 512, 24, 4, 64 * 8 * 33510)
    assert abs(b32 / 1e9 - 132.0) < 1.0
    assert bench.algorithmic_bytes(10, 2, 3, 8, 0) == 8 * 10 * 2 * (2 + 3 * 2) + 3 * 8 * 10


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stdout
