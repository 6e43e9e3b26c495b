"""The allocation-free entry point wap_atom_potential_static against wap_atom_potential.

Both are called through the C ABI on the same inputs, twice in a row, the second
call with fewer atoms on a smaller grid (the static staging array keeps the first
call's records beyond the second's n_atoms, which the second call must not read).
The results must be bitwise equal: the same kernel reads the same records.  Sizes
beyond the static maximum are rejected on the host before any device work.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _call(fn, atoms, shape, extent, dtype):
    from wideanglepropagation_amd._lib import DT_F32, DT_F64
    nz, ny, nx = shape
    recs = np.ascontiguousarray(atoms, dtype=np.float64)
    ext = np.ascontiguousarray(extent, dtype=np.float64)
    out = torch.empty(shape, dtype=dtype, device="cuda")
    code = DT_F64 if dtype == torch.float64 else DT_F32
    rc = fn(ctypes.c_void_p(recs.ctypes.data), len(recs), nx, ny, nz,
            ctypes.c_void_p(ext.ctypes.data), 0.05, 6.0, ctypes.c_void_p(out.data_ptr()), code,
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc, out


def _atoms(rng, n, extent):
    x0, x1, y0, y1, z0, z1 = extent
    return np.column_stack([rng.uniform(x0, x1, n), rng.uniform(y0, y1, n),
                            rng.uniform(z0, z1, n), rng.choice([14.0, 29.0, 79.0], n),
                            rng.uniform(0.3, 0.5, n)])


def test_static_atom_potential_equals_the_allocating_entry_point():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from wideanglepropagation_amd._lib import lib
    L = lib()
    rng = np.random.default_rng(21)
    cases = [((64, 48, 40), (-1.0, 0.95, -1.2, 1.15, -0.8, 0.8), 900, torch.float32),
             ((24, 20, 18), (-0.5, 0.35, -0.5, 0.45, -0.3, 0.3), 37, torch.float64)]
    for shape, ext, n, dt in cases:
        print(f"case static atom potential {shape} {n} atoms {dt}", flush=True)
        atoms = _atoms(rng, n, ext)
        rc_a, want = _call(L.wap_atom_potential, atoms, shape, ext, dt)
        rc_b, got = _call(L.wap_atom_potential_static, atoms, shape, ext, dt)
        assert rc_a == 0 and rc_b == 0, L.wap_last_error()
        assert float(want.abs().max()) > 0
        assert torch.equal(got, want)


def test_static_atom_potential_rejects_too_many_atoms():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from wideanglepropagation_amd._lib import lib
    L = lib()
    ext = (-0.5, 0.5, -0.5, 0.5, -0.5, 0.5)
    atoms = _atoms(np.random.default_rng(22), 4097, ext)
    rc, _ = _call(L.wap_atom_potential_static, atoms, (8, 8, 8), ext, torch.float32)
    assert rc != 0 and b"exceed the static staging" in L.wap_last_error()
