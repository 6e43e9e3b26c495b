"""CPU oracle for the 3D RK4 FDTD electron-wave path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
package.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker
(or, in ``bench.py``, as the timed CPU baseline), never as a code path of the
solver.

Contents
--------
physics   closed forms of the reference's constants
          (``fdtd_solver.py:11-29, 306-342``) and its macro quantisation.
fdtd3d    FP64 NumPy restatement of ``fdtd_3d.cu`` (race-free classical RK4,
          same stencil, PML, soft source and frozen end planes) plus the
          first-order Klein-Gordon extension.
extract   the notebooks' exit-wave / far-field recipe
          (``FDTDmethod/fdtd_3D_baseline_comparison.ipynb:267-290``,
          ``propagation.py:346-350``).
spectral  the Fresnel / angular-spectrum multislice the notebooks compare FDTD
          exit waves against (``propagation.py:14-177, 305-352``), JAX-free.
fdtd3d_c  ctypes front of ``fdtd3d_omp.c``: the same FP64 step restated in C +
          OpenMP (one streaming sweep per RK4 stage), bench.py's multi-threaded
          CPU baseline; coefficients come from ``fdtd3d``.

Pinning (what checks each function, in ``tests/``):
  physics.*            reference-held golden scalars (notebook outputs and the
                       checked-in macros of ``fdtd_3d.cu:18-67``).
  fdtd3d.rk4_step      brute force: one step equals P4(-i dt H) applied to the
                       injected state, with H assembled independently as a
                       scipy.sparse matrix from the ``fdtd_3d.cu:192-248``
                       formulae; closed-form discrete dispersion in a periodic
                       box; bitwise determinism; the reference's five physics
                       invariants (``tests/test_fdtd_3d.py:203-323``).
  fdtd3d (KG)          closed-form KG dispersion in a periodic box
                       (omega = V +- sqrt(c^2 k_d^2 + m^2), ``tests/test_oracle.py``);
                       and in ``tests/test_oracle_kg.py``: the incident wave
                       (pi = -i w psi) passing ``inject`` unchanged and, on the
                       discrete dispersion relation, stepping by P4(-i w dt); the
                       sponge's closed form and a step against P4(dt M) assembled
                       from an independent sponge and source; the KG charge
                       conserved to RK4 order without sponge and drained at the
                       damping term's rate with it; the source's wave rising in
                       phase along +z with the fully discrete wavenumber.  Each
                       check fails on a mistake planted in the oracle (+i w, a
                       dropped profile, r^(p+1), a dropped or flipped 2 i V pi,
                       gamma pi for 2 gamma pi).
  fdtd3d.ref_time      the int * float product of ``fdtd_3d.cu:365`` rounded to
                       binary32 in exact rational arithmetic.
  fdtd3d_c.run         fdtd3d.run to 1e-12 (both equations, source profile, 2D
                       zero-end rule, float32 clock, 1 and 3 threads), and a
                       planted dropped sponge term that the check catches
                       (``tests/test_oracle_c.py``).
  extract.*            synthetic snapshots with a known answer.
  spectral.*           textbook electron wavelengths, the refractive index in
                       60-digit decimal arithmetic, single Fourier modes
                       propagated analytically, the paraxial expansion, kernel
                       composition, vacuum / constant-potential multislices,
                       norm conservation (``tests/test_spectral_oracle.py``).
"""
