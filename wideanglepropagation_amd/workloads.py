"""Synthetic inputs of the solver's standard workloads (potentials, initial states).

The reference builds potentials with abTEM in its notebooks
(``FDTDmethod/fdtd_3D_baseline_comparison.ipynb:129-180, 210-221``), which is
not available here; these builders make screened-Coulomb atom sums with the
regularised ``1/sqrt(r^2 + r_s^2)`` form the reference itself uses in 2D
(``FDTDmethod/fdtd_1D_baseline_comparison.ipynb:162-165``):

    V(r) = - KE2 * Z * exp(-r / a) / sqrt(r^2 + r_s^2)     [eV],  KE2 = e^2/(4 pi eps0) = 14.399645 eV A

summed over atoms with minimum-image distances along the periodic x and y
axes (z is not periodic).  Atoms contribute only inside a cutoff box
(default 8 screening lengths, where exp(-r/a) < 3.4e-4).

Standard workloads (all lengths in A, energies in eV, time in hbar/eV):
  schrodinger_gaussian_64    64^3 Schrodinger, complex128, dx = 0.05, PML 8 cells,
                             Gaussian packet (sigma 1.5 A, k along z for 200 keV),
                             one Z = 14 atom (screening 0.3 A) at the centre,
                             source off, 200 steps.
  kg_slab                    Klein-Gordon plane wave (soft source at PML + 2),
                             random atom slab, PML 16 cells (the 256x256x512 and
                             512x512x256 grids).
"""
import math

import numpy as np

from .physics import (schrodinger_params, klein_gordon_params, beam_constants,
                      wavelength_from_kinetic_energy)

KE2 = 14.399645  # eV A


def centred_extent(shape, spacing):
    nz, ny, nx = shape
    hx = (nx - 1) * spacing / 2.0
    hy = (ny - 1) * spacing / 2.0
    hz = (nz - 1) * spacing / 2.0
    return (-hx, hx, -hy, hy, -hz, hz)


def atom_potential(shape, extent, positions, Z, screening, r_soft, cutoff=8.0, device="cpu",
                   dtype=None):
    """Sum of screened-Coulomb atoms on the grid; returns a torch tensor (nz, ny, nx).

    positions (n,3) as (x, y, z) in A; Z, screening (n,) arrays.  x, y periodic
    with period n*d (the grid has n points, spacing d = extent/(n-1)).
    """
    import torch
    dtype = dtype or torch.float64
    nz, ny, nx = shape
    x0, x1, y0, y1, z0, z1 = extent
    dx = (x1 - x0) / (nx - 1)
    dy = (y1 - y0) / (ny - 1)
    dz = (z1 - z0) / (nz - 1)
    V = torch.zeros(shape, dtype=dtype, device=device)
    positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    Z = np.broadcast_to(np.asarray(Z, dtype=np.float64), (len(positions),))
    screening = np.broadcast_to(np.asarray(screening, dtype=np.float64), (len(positions),))
    for (px, py, pz), z_at, a in zip(positions, Z, screening):
        rc = cutoff * a
        ci, cj, ck = (px - x0) / dx, (py - y0) / dy, (pz - z0) / dz
        ri, rj, rk = int(math.ceil(rc / dx)), int(math.ceil(rc / dy)), int(math.ceil(rc / dz))
        ii = np.arange(int(math.floor(ci)) - ri, int(math.ceil(ci)) + ri + 1)
        jj = np.arange(int(math.floor(cj)) - rj, int(math.ceil(cj)) + rj + 1)
        kk = np.arange(max(0, int(math.floor(ck)) - rk), min(nz, int(math.ceil(ck)) + rk + 1))
        if len(kk) == 0:
            continue
        if len(ii) >= nx:
            ii = np.arange(nx) + int(math.floor(ci)) - nx // 2
        if len(jj) >= ny:
            jj = np.arange(ny) + int(math.floor(cj)) - ny // 2
        # minimum-image offsets of the (unwrapped) window indices
        xs = torch.as_tensor((ii - ci) * dx, dtype=dtype, device=device)
        ys = torch.as_tensor((jj - cj) * dy, dtype=dtype, device=device)
        zs = torch.as_tensor(kk * dz + z0 - pz, dtype=dtype, device=device)
        r2 = zs[:, None, None] ** 2 + ys[None, :, None] ** 2 + xs[None, None, :] ** 2
        r = torch.sqrt(r2)
        v = -KE2 * z_at * torch.exp(-r / a) / torch.sqrt(r2 + r_soft * r_soft)
        iw = torch.as_tensor(np.mod(ii, nx), device=device)
        jw = torch.as_tensor(np.mod(jj, ny), device=device)
        kw = torch.as_tensor(kk, device=device)
        V.index_put_((kw[:, None, None], jw[None, :, None], iw[None, None, :]), v,
                     accumulate=True)
    return V


def random_slab(n_atoms, cell_xy, slab_z, seed, z_choices=(14, 29, 79), screening=(0.3, 0.5)):
    """Seeded random atoms: x, y uniform over the cell, z uniform in slab_z=(z0, z1)."""
    rng = np.random.default_rng(seed)
    (xa, xb), (ya, yb) = cell_xy
    pos = np.stack([rng.uniform(xa, xb, n_atoms), rng.uniform(ya, yb, n_atoms),
                    rng.uniform(slab_z[0], slab_z[1], n_atoms)], axis=1)
    Z = rng.choice(np.asarray(z_choices, dtype=np.float64), n_atoms)
    scr = rng.uniform(screening[0], screening[1], n_atoms)
    return pos, Z, scr


def schrodinger_gaussian_64(precision="f64", nsteps=200, n=64, spacing=0.05):
    """The 64^3 Schrodinger wavepacket case; returns (params, V (numpy), psi0 (numpy), nsteps)."""
    lam = wavelength_from_kinetic_energy(200e3)
    shape = (n, n, n)
    ext = centred_extent(shape, spacing)
    p = schrodinger_params(lam, ext, shape, pml_thick=8 * spacing, k_source=0,
                           precision=precision, src_enabled=False)
    V = atom_potential(shape, ext, [(0.0, 0.0, 0.0)], [14.0], [0.3], r_soft=spacing).numpy()
    x = np.linspace(ext[0], ext[1], n)
    X = x[None, None, :]
    Y = x[None, :, None]
    Zc = x[:, None, None]
    k0 = beam_constants(lam)["P_INF"]
    psi0 = np.exp(-(X ** 2 + Y ** 2 + Zc ** 2) / (2 * 1.5 ** 2)) * np.exp(1j * k0 * Zc)
    return p, V, psi0, nsteps


def kg_slab(shape, precision="f32", n_atoms=400, seed=0, spacing=0.05, pml_cells=16,
            slab=20.0, device="cpu", kinetic_ev=200e3):
    """Klein-Gordon plane wave through a random atom slab; returns (params, V torch)."""
    lam = wavelength_from_kinetic_energy(kinetic_ev)
    ext = centred_extent(shape, spacing)
    nz, ny, nx = shape
    half = min(slab / 2.0, ext[5] - (pml_cells + 4) * spacing)
    pos, Z, scr = random_slab(n_atoms, ((ext[0], ext[1] + spacing), (ext[2], ext[3] + spacing)),
                              (-half, half), seed)
    p = klein_gordon_params(lam, ext, shape, pml_thick=pml_cells * spacing,
                            k_source=pml_cells + 2, precision=precision)
    V = atom_potential(shape, ext, pos, Z, scr, r_soft=spacing, device=device)
    return p, V
