"""GPU atom-potential generator against the CPU builder (same sphere-cutoff sum)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("shape,spz,cut", [((40, 48, 40), 0.01, 3.0),    # box path on the CPU
                                           ((10, 12, 14), 0.05, 8.0)])   # all images on the CPU
def test_hip_atom_potential_matches_cpu_builder(shape, spz, cut):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from wideanglepropagation_amd.workloads import atom_potential, centred_extent, random_slab
    print(f"case atom potential {shape} cutoff {cut}", flush=True)
    ext = centred_extent(shape, 0.05, spz)
    pos, Z, scr = random_slab(30, ((ext[0], ext[1]), (ext[2], ext[3])), (ext[4], ext[5]), 5)
    cpu = atom_potential(shape, ext, pos, Z, scr, 0.05, cutoff=cut, device="cpu").numpy()
    gpu = atom_potential(shape, ext, pos, Z, scr, 0.05, cutoff=cut, device="cuda").cpu().numpy()
    assert np.abs(gpu - cpu).max() <= 1e-11 * np.abs(cpu).max()
    g32 = atom_potential(shape, ext, pos, Z, scr, 0.05, cutoff=cut, device="cuda",
                         dtype=torch.float32).cpu().numpy()
    assert np.array_equal(g32, gpu.astype(np.float32))
