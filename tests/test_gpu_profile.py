"""Live timing of the fine-level Poisson operator (hoc_profile_*): every launch, or every n-th,
is bracketed by HIP events on the handle's stream; the counts, the per-variant algorithmic bytes
and the durations are reported per fused-epilogue variant."""
import numpy as np
import pytest

import hoc
from hoc import meshes

pytestmark = pytest.mark.gpu


def test_profile_counts_and_sampling(gpu_available):
    m = meshes.cavity(32, 2)
    g = hoc.HO2D(K=4, problem_type=1, Re=100.0, dt=2e-4, time_integration_type=4, advection_type=1, tol=1e-10,
                 maxiter=300)
    g.load_mesh(m)
    g.setup()
    g.step(1)
    counts = {}
    for every in (1, 3):
        g.set_vorticity(np.zeros(g.D))
        g.set_streamfunction(np.zeros(g.D))
        g.profile(True, every=every)
        g.step(2)
        p = g.profile_read()
        g.profile(False)
        # launches predicated off after a solve converged are sampled too, then left out of the timings
        counts[every] = sum(v["launches"] for v in p.values()) + g.profile_noop_launches
        for name, v in p.items():
            if v["launches"]:
                assert v["total_ms"] > 0 and v["bytes_per_launch"] > 0, name
        assert p["lap_cheb"]["launches"] > 0
    print("launches timed: every launch", counts[1], "every 3rd", counts[3])
    # the same two steps from the same state: one launch in three is timed
    assert abs(counts[3] - counts[1] / 3) <= 1
    g.close()
