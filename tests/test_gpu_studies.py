"""The adjoint and API studies in their quick sizes: they run, report finite numbers, and
the adjoint study's f64 gradient meets its finite-difference target."""
import math

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_adjoint_study_quick():
    from wideanglepropagation_amd.studies import adjoint_study
    print("case adjoint study quick", flush=True)
    r = adjoint_study(quick=True)
    assert r["fd_voxels"] == 6
    assert r["max_rel_err_f64_adjoint_vs_fd"] < 1e-4
    s = math.ceil(math.sqrt(r["steps"]))
    sq = r["cost"]["sqrt_t"]["segment_plan"]
    assert sq["segment"] == s and sq["recomputed_steps"] == r["steps"] - s
    plan = r["cost"]["memory_budget"]["segment_plan"]
    assert plan["segment"] >= s
    assert plan["recomputed_steps"] == r["steps"] - plan["segment"] or plan["segment"] == r["steps"]
    for c in r["cost"].values():
        assert c["t_forward_s"] > 0 and c["t_forward_backward_s"] > c["t_forward_s"]


def test_api_study_quick():
    from wideanglepropagation_amd.studies import api_study
    print("case api study quick", flush=True)
    r = api_study(quick=True)
    assert r["atom_potential"]["ms"] > 0
    for k in ("solver3d_run", "solver2d_run"):
        v = r[k]["stepping_Gvoxel_steps_per_s"]
        assert math.isfinite(v) and v > 0
    c = r["cli_vs_inprocess"]
    assert c["cli_s"] > 0 and c["inprocess_s"] > 0
