"""Synthetic inputs of the solver's standard workloads (potentials, initial states).

The reference builds potentials with abTEM in its notebooks
(``FDTDmethod/fdtd_3D_baseline_comparison.ipynb:129-180, 210-221``), which is
not available here; these builders make screened-Coulomb atom sums with the
regularised ``1/sqrt(r^2 + r_s^2)`` form the reference itself uses in 2D
(``FDTDmethod/fdtd_1D_baseline_comparison.ipynb:162-165``):

    V(r) = - KE2 * Z * exp(-r / a) / sqrt(r^2 + r_s^2)     [eV],  KE2 = e^2/(4 pi eps0) = 14.399645 eV A

summed over atoms and their periodic images along x and y (z is not
periodic) within a cutoff sphere of 8 screening lengths (exp(-r/a) < 3.4e-4
there).  On the GPU the sum is one HIP launch (``wap_atom_potential``); the
CPU builder adds each atom over its cutoff box and needs the cutoff below
half the lateral cell.

Standard workloads (all lengths in A, energies in eV, time in hbar/eV):
  schrodinger_gaussian_64    64^3 Schrodinger, complex128, dx = 0.05, PML 8 cells,
                             Gaussian packet (sigma 1.5 A, k along z for 200 keV),
                             one Z = 14 atom (screening 0.3 A) at the centre,
                             source off, 200 steps.
  kg_slab                    Klein-Gordon plane wave (soft source at PML + 2),
                             random atom slab, PML 16 planes, dx = dy = 0.05 A,
                             dz = 0.005 A (the 256x256x512 and 512x512x256 grids).
"""
import math

import numpy as np

from .physics import (schrodinger_params, klein_gordon_params, beam_constants,
                      wavelength_from_kinetic_energy)

KE2 = 14.399645  # eV A


def centred_extent(shape, spacing, spacing_z=None):
    nz, ny, nx = shape
    dz = spacing if spacing_z is None else spacing_z
    hx = (nx - 1) * spacing / 2.0
    hy = (ny - 1) * spacing / 2.0
    hz = (nz - 1) * dz / 2.0
    return (-hx, hx, -hy, hy, -hz, hz)


def atom_potential(shape, extent, positions, Z, screening, r_soft, cutoff=8.0, device="cpu",
                   dtype=None):
    """Sum of screened-Coulomb atoms on the grid; returns a torch tensor (nz, ny, nx).

    positions (n,3) as (x, y, z) in A; Z, screening (n,) arrays.  x, y periodic
    with period n*d (the grid has n points, spacing d = extent/(n-1)).  Each
    atom is added over its cutoff box on a laterally padded grid (plain slice
    adds, no scatter); the padding margins are folded back periodically.
    """
    import torch
    dtype = dtype or torch.float64
    nz, ny, nx = shape
    x0, x1, y0, y1, z0, z1 = extent
    dx = (x1 - x0) / (nx - 1)
    dy = (y1 - y0) / (ny - 1)
    dz = (z1 - z0) / (nz - 1)
    positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    Z = np.broadcast_to(np.asarray(Z, dtype=np.float64), (len(positions),))
    screening = np.broadcast_to(np.asarray(screening, dtype=np.float64), (len(positions),))
    if len(positions) == 0:
        return torch.zeros(shape, dtype=dtype, device=device)
    if str(device).startswith("cuda"):
        return _atom_potential_hip(shape, extent, positions, Z, screening, r_soft, cutoff, device,
                                   dtype)
    if cutoff * float(np.max(screening)) * 2 >= min(nx * dx, ny * dy):
        return _atom_potential_images(shape, extent, positions, Z, screening, r_soft, cutoff,
                                      device, dtype)
    rmax = cutoff * float(np.max(screening))
    mx = min(int(math.ceil(rmax / dx)) + 1, nx)   # lateral padding (at most one period)
    my = min(int(math.ceil(rmax / dy)) + 1, ny)
    Vp = torch.zeros((nz, ny + 2 * my, nx + 2 * mx), dtype=dtype, device=device)
    for (px, py, pz), z_at, a in zip(positions, Z, screening):
        rc = cutoff * a
        ci, cj, ck = (px - x0) / dx, (py - y0) / dy, (pz - z0) / dz
        # wrap the centre into the primary cell, then take its box on the padded grid
        ci = ci % nx
        cj = cj % ny
        ri = min(int(math.ceil(rc / dx)), mx - 1)
        rj = min(int(math.ceil(rc / dy)), my - 1)
        rk = int(math.ceil(rc / dz))
        i0, i1 = int(math.floor(ci)) - ri, int(math.floor(ci)) + ri + 1
        j0, j1 = int(math.floor(cj)) - rj, int(math.floor(cj)) + rj + 1
        k0, k1 = max(0, int(math.floor(ck)) - rk), min(nz, int(math.floor(ck)) + rk + 2)
        if k1 <= k0:
            continue
        xs = torch.arange(i0, i1, dtype=dtype, device=device) - ci
        ys = torch.arange(j0, j1, dtype=dtype, device=device) - cj
        zs = torch.arange(k0, k1, dtype=dtype, device=device) * dz + z0 - pz
        r2 = (zs[:, None, None] ** 2 + (ys[None, :, None] * dy) ** 2
              + (xs[None, None, :] * dx) ** 2)
        r = torch.sqrt(r2)
        v = -KE2 * z_at * torch.exp(-r / a) / torch.sqrt(r2 + r_soft * r_soft)
        v = torch.where(r2 <= rc * rc, v, torch.zeros_like(v))
        Vp[k0:k1, j0 + my:j1 + my, i0 + mx:i1 + mx] += v
    # fold the margins back into the periodic cell
    V = Vp[:, my:my + ny, mx:mx + nx].clone()
    V[:, :, nx - mx:] += Vp[:, my:my + ny, :mx]
    V[:, :, :mx] += Vp[:, my:my + ny, mx + nx:]
    top = Vp[:, :my, :]
    bot = Vp[:, my + ny:, :]
    for band, dst in ((top, slice(ny - my, ny)), (bot, slice(0, my))):
        row = band[:, :, mx:mx + nx].clone()
        row[:, :, nx - mx:] += band[:, :, :mx]
        row[:, :, :mx] += band[:, :, mx + nx:]
        V[:, dst, :] += row
    return V


def _atom_potential_images(shape, extent, positions, Z, screening, r_soft, cutoff, device, dtype):
    """CPU sum over every periodic image within the cutoff (cutoffs beyond half the cell)."""
    import torch
    nz, ny, nx = shape
    x0, x1, y0, y1, z0, z1 = extent
    x = torch.linspace(x0, x1, nx, dtype=dtype)
    y = torch.linspace(y0, y1, ny, dtype=dtype)
    z = torch.linspace(z0, z1, nz, dtype=dtype)
    Lx, Ly = nx * (x1 - x0) / (nx - 1), ny * (y1 - y0) / (ny - 1)
    V = torch.zeros(shape, dtype=dtype)
    for (px, py, pz), z_at, a in zip(positions, Z, screening):
        rc = cutoff * a
        mx, my = int(math.ceil(rc / Lx)) + 1, int(math.ceil(rc / Ly)) + 1
        ez = (z - pz)[:, None, None]
        for sy in range(-my, my + 1):
            ey = (y - py + sy * Ly)[None, :, None]
            for sx in range(-mx, mx + 1):
                ex = (x - px + sx * Lx)[None, None, :]
                r2 = ex * ex + ey * ey + ez * ez
                v = -KE2 * z_at * torch.exp(-torch.sqrt(r2) / a) / torch.sqrt(r2 + r_soft ** 2)
                V += torch.where(r2 <= rc * rc, v, torch.zeros_like(v))
    return V.to(device)


def _atom_potential_hip(shape, extent, positions, Z, screening, r_soft, cutoff, device, dtype):
    """The same sum on the GPU in one launch (``wap_atom_potential``)."""
    import ctypes
    import torch
    from ._lib import check, lib, DT_F32, DT_F64
    nz, ny, nx = shape
    recs = np.ascontiguousarray(np.column_stack([positions, Z, screening]).astype(np.float64))
    ext = np.ascontiguousarray(np.asarray(extent, dtype=np.float64))
    out = torch.empty(shape, dtype=dtype, device=device)
    code = DT_F64 if dtype == torch.float64 else DT_F32
    check(lib().wap_atom_potential(ctypes.c_void_p(recs.ctypes.data), len(recs), nx, ny, nz,
                                   ctypes.c_void_p(ext.ctypes.data), float(r_soft), float(cutoff),
                                   ctypes.c_void_p(out.data_ptr()), code,
                                   ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
          "wap_atom_potential")
    return out


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


def kg_slab(shape, precision="f32", n_atoms=400, seed=0, spacing=0.05, spacing_z=0.005,
            pml_cells=16, slab=20.0, device="cpu", kinetic_ev=200e3, schedule="auto",
            positions=None, cutoff=8.0):
    """Klein-Gordon plane wave through a random atom slab; returns (params, V torch).

    Lateral spacing 0.05 A; along z 0.005 A so that the 200 keV carrier
    (k = 250.5 / A) propagates on the grid (k dz = 1.25 rad; at dz = 0.05 A it
    is evanescent).  The slab is clipped to the interior between the PML
    layers; ``positions`` = (pos, Z, screening) overrides the random atoms.
    """
    lam = wavelength_from_kinetic_energy(kinetic_ev)
    ext = centred_extent(shape, spacing, spacing_z)
    half = max(0.0, min(slab / 2.0, ext[5] - (pml_cells + 4) * spacing_z))
    if positions is None:
        pos, Z, scr = random_slab(n_atoms, ((ext[0], ext[1] + spacing),
                                            (ext[2], ext[3] + spacing)), (-half, half), seed)
    else:
        pos, Z, scr = positions
    p = klein_gordon_params(lam, ext, shape, pml_thick=pml_cells * spacing_z,
                            k_source=pml_cells + 2, precision=precision, schedule=schedule)
    V = atom_potential(shape, ext, pos, Z, scr, r_soft=spacing, cutoff=cutoff, device=device)
    return p, V
