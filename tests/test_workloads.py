"""Potential builder and workload parameters (CPU)."""
import math

import numpy as np
import pytest

from wideanglepropagation_amd.workloads import (KE2, atom_potential, centred_extent, kg_slab,
                                                random_slab, schrodinger_gaussian_64)


def periodic_sphere_sum(shape, ext, pos, Z, scr, rsoft, cut):
    """Brute force: every atom image within cut * s of every voxel."""
    nz, ny, nx = shape
    x = np.linspace(ext[0], ext[1], nx)
    y = np.linspace(ext[2], ext[3], ny)
    z = np.linspace(ext[4], ext[5], nz)
    Lx, Ly = nx * (x[1] - x[0]), ny * (y[1] - y[0])
    W = np.zeros(shape)
    for (px, py, pz), zz, a in zip(pos, Z, scr):
        rc = cut * a
        for sx in range(-3, 4):
            for sy in range(-3, 4):
                X = x[None, None, :] - px + sx * Lx
                Y = y[None, :, None] - py + sy * Ly
                Zc = z[:, None, None] - pz
                r2 = X ** 2 + Y ** 2 + Zc ** 2
                v = -KE2 * zz * np.exp(-np.sqrt(r2) / a) / np.sqrt(r2 + rsoft ** 2)
                W += np.where(r2 <= rc * rc, v, 0.0)
    return W


@pytest.mark.parametrize("shape,sp,spz,cut", [((12, 9, 8), 0.1, 0.05, 1.0),
                                              ((7, 16, 21), 0.05, 0.05, 1.5)])
def test_atom_potential_equals_periodic_sphere_sum(shape, sp, spz, cut):
    ext = centred_extent(shape, sp, spz)
    rng = np.random.default_rng(0)
    n = 4
    pos = np.stack([rng.uniform(ext[0], ext[1], n), rng.uniform(ext[2], ext[3], n),
                    rng.uniform(ext[4], ext[5], n)], 1)
    Z = rng.choice([14., 29., 79.], n)
    scr = rng.uniform(0.1, 0.25, n)
    V = atom_potential(shape, ext, pos, Z, scr, 0.05, cutoff=cut).numpy()
    W = periodic_sphere_sum(shape, ext, pos, Z, scr, 0.05, cut)
    assert np.abs(V - W).max() <= 1e-12 * np.abs(W).max()


def _unused_box_sum(shape, sp, spz, cut):
    ext = centred_extent(shape, sp, spz)
    rng = np.random.default_rng(0)
    n = 4
    pos = np.stack([rng.uniform(ext[0], ext[1], n), rng.uniform(ext[2], ext[3], n),
                    rng.uniform(ext[4], ext[5], n)], 1)
    Z = rng.choice([14., 29., 79.], n)
    scr = rng.uniform(0.1, 0.4, n)
    V = atom_potential(shape, ext, pos, Z, scr, 0.05, cutoff=cut).numpy()
    nz, ny, nx = shape
    dx = (ext[1] - ext[0]) / (nx - 1)
    dy = (ext[3] - ext[2]) / (ny - 1)
    dz = (ext[5] - ext[4]) / (nz - 1)
    W = np.zeros(shape)
    mx = min(math.ceil(cut * scr.max() / dx) + 1, nx)
    my = min(math.ceil(cut * scr.max() / dy) + 1, ny)
    for (px, py, pz), zz, a in zip(pos, Z, scr):
        ci = ((px - ext[0]) / dx) % nx
        cj = ((py - ext[2]) / dy) % ny
        ck = (pz - ext[4]) / dz
        ri = min(math.ceil(cut * a / dx), mx - 1)
        rj = min(math.ceil(cut * a / dy), my - 1)
        rk = math.ceil(cut * a / dz)
        for k in range(max(0, math.floor(ck) - rk), min(nz, math.floor(ck) + rk + 2)):
            for j in range(math.floor(cj) - rj, math.floor(cj) + rj + 1):
                for i in range(math.floor(ci) - ri, math.floor(ci) + ri + 1):
                    r2 = ((i - ci) * dx) ** 2 + ((j - cj) * dy) ** 2 + (k * dz + ext[4] - pz) ** 2
                    W[k, j % ny, i % nx] += (-KE2 * zz * math.exp(-math.sqrt(r2) / a)
                                             / math.sqrt(r2 + 0.05 ** 2))
    assert np.abs(V - W).max() <= 1e-12 * np.abs(W).max()


def test_single_atom_closed_form_at_centre():
    shape = (9, 9, 9)
    ext = centred_extent(shape, 0.05)
    # cutoff 0.6 a = 0.18 A keeps the box inside one period (no periodic images)
    V = atom_potential(shape, ext, [(0.0, 0.0, 0.0)], [14.0], [0.3], r_soft=0.05,
                       cutoff=0.6).numpy()
    assert V[4, 4, 4] == pytest.approx(-KE2 * 14.0 / 0.05, rel=1e-14)
    r = 0.05
    assert V[4, 4, 5] == pytest.approx(-KE2 * 14 * math.exp(-r / 0.3) / math.sqrt(2 * r * r),
                                       rel=1e-14)


def test_random_slab_is_seeded_and_bounded():
    a = random_slab(50, ((-1, 1), (-2, 2)), (-0.5, 0.5), 7)
    b = random_slab(50, ((-1, 1), (-2, 2)), (-0.5, 0.5), 7)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    pos, Z, scr = a
    assert np.all(np.abs(pos[:, 2]) <= 0.5) and set(np.unique(Z)) <= {14.0, 29.0, 79.0}
    assert np.all((scr >= 0.3) & (scr <= 0.5))


def test_standard_workload_parameters():
    p, V, psi0, n = schrodinger_gaussian_64()
    assert p.shape == (64, 64, 64) and n == 200 and not p.src_enabled
    assert p.spacing()[0] == pytest.approx(0.05) and p.pml_thick == pytest.approx(0.4)
    q, Vq = kg_slab((40, 8, 8), n_atoms=3, device="cpu")
    assert q.equation == "klein_gordon" and q.src_k == 18
    dz = q.spacing()[2]
    assert dz == pytest.approx(0.005)
    # the 200 keV carrier propagates on this grid: (k dz / 2) < pi / 2
    assert q.k_beam * dz / 2 < math.pi / 2
