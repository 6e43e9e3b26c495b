"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle still reproduces its stored vectors.  GPU: the HIP stepper
(f64, every schedule of the equation) reproduces them to 1e-12 relative L2 per state
component.
"""
import glob
import json
import os

import numpy as np
import pytest

from oracle import fdtd3d
from wideanglepropagation_amd.physics import FDTDParams

FIX = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def _load(path):
    z = np.load(path, allow_pickle=False)
    p = FDTDParams(**json.loads(str(z["params"])))
    st0 = (z["psi0"], z["pi0"]) if "pi0" in z else z["psi0"]
    want = (z["psi"], z["pi"]) if "pi" in z else z["psi"]
    return p, z["V"], st0, int(z["nsteps"]), bool(z["f32_time"]), want


def _rel(a, b):
    """Largest relative L2 over the components (psi, and pi for Klein-Gordon)."""
    a = a if isinstance(a, tuple) else (a,)
    b = b if isinstance(b, tuple) else (b,)
    return max(np.linalg.norm(np.ravel(x) - np.ravel(y)) / np.linalg.norm(np.ravel(y))
               for x, y in zip(a, b))


def test_fixtures_exist():
    assert len(FIX) >= 3


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(f) for f in FIX])
def test_oracle_reproduces_golden(path):
    p, V, st0, n, f32t, want = _load(path)
    got = fdtd3d.run(st0, V, n, p.oracle_dict(), f32_time=f32t)
    assert _rel(got, want) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(f) for f in FIX])
def test_hip_reproduces_golden(path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from wideanglepropagation_amd.engine import FDTD3D
    p, V, st0, n, f32t, want = _load(path)
    scheds = (["stages", "kg_2pass", "kg_fused"] if p.equation == "klein_gordon"
              else ["stages", "auto"])
    for sched in scheds:
        print(f"case golden {os.path.basename(path)} {sched}", flush=True)
        sim = FDTD3D(p.with_(schedule=sched, precision="f64"), V=V)
        if isinstance(st0, tuple):
            sim.set_state(st0[0], st0[1])
        else:
            sim.set_state(st0)
        sim.step(n)
        got = sim.get_state(np.complex128)
        sim.close()
        assert _rel(got, want) < 1e-12
