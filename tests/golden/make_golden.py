"""Write the golden fixtures of tests/golden/ from the FP64 oracle.

Run from the repository root:  python tests/golden/make_golden.py
Each .npz holds the inputs (parameters as JSON, V, initial state) and the
oracle's output after `nsteps` steps.  The oracle is pinned by
tests/test_oracle.py; these fixtures freeze its output so later changes to
the oracle or the HIP stepper are caught against stored vectors.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

from oracle import fdtd3d  # noqa: E402
from wideanglepropagation_amd.physics import schrodinger_params, klein_gordon_params  # noqa: E402
from wideanglepropagation_amd.workloads import centred_extent  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def save(name, p, V, st0, nsteps, f32_time=False):
    d = p.oracle_dict()
    out = fdtd3d.run(st0, V, nsteps, d, f32_time=f32_time)
    kw = {"params": json.dumps(p.as_dict()), "V": V, "nsteps": nsteps,
          "f32_time": f32_time}
    if isinstance(st0, tuple):
        kw.update(psi0=st0[0], pi0=st0[1], psi=out[0], pi=out[1])
    else:
        kw.update(psi0=st0, psi=out)
    np.savez_compressed(os.path.join(OUT, name), **kw)
    print(name, nsteps, "steps")


def main():
    rng = np.random.default_rng(2024)
    # the reference's physics-test case (tests/test_fdtd_3d.py:166-182), quantised like the
    # compiled reference, from a zero field: 20 steps of the soft source
    p = schrodinger_params(0.1, (-1, 1, -1, 1, -1.5, 1.5), (24, 16, 16), pml_thick=0.3,
                           k_source=4, reference_quantize=True, precision="f64")
    save("schr_ref_case_16x16x24.npz", p, np.zeros(p.shape), fdtd3d.zero_state(p.oracle_dict()),
         20, f32_time=True)
    # random potential and state, PML and source on
    shape = (12, 8, 8)
    ext = centred_extent(shape, 0.05)
    p = schrodinger_params(0.05, ext, shape, pml_thick=0.15, k_source=3, precision="f64")
    V = 2e3 * rng.standard_normal(shape)
    psi0 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    save("schr_rand_8x8x12.npz", p, V, psi0, 10)
    p = klein_gordon_params(0.05, ext, shape, pml_thick=0.15, k_source=3, precision="f64")
    psi0 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    pi0 = -1j * p.e_sim * psi0
    save("kg_rand_8x8x12.npz", p, 2e3 * rng.standard_normal(shape), (psi0, pi0), 10)


if __name__ == "__main__":
    main()
