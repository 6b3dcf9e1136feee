"""The reference's MISMIP+ Ice0 verification experiment
(test/verification_tests/MISMIP_PLUS_Ice0.jl, checked in
test/verification_tests.jl:13-49): 80 x 10 cells of 8 km, nsigma 4, one
Picard iteration per step, dt 0.5 a for 10,000 years, then the grounding
line's symmetry across the channel and its position."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mismip_plus_ice0_grounding_line(W):
    print("[gpu] mismip+ Ice0: 20000 steps of 0.5 a", flush=True)
    sys.stdout.flush()
    m = W.workloads.mismip_plus_fixture()
    sim = W.Simulation(model=m, timestepping_params=W.TimesteppingParams(dt=0.5, end_time=10000.0))
    W.run_simulation(sim)
    ny = m.grid.ny
    glx = W.get_glx(m)
    assert len(glx) == ny  # the grounding line crosses every row once
    g = glx[[0, ny // 2 - 1, ny // 2, ny - 1]]
    # symmetric about the channel's centre line (the reference tests the
    # signed ratios; their magnitudes are tested here)
    assert abs(g[3] - g[0]) / (g[3] + g[0]) < 1e-4
    assert abs(g[1] - g[2]) / (g[1] + g[2]) < 1e-4
    # the reference's loose position bounds
    assert 480000 < g[0] < 550000 and 480000 < g[3] < 550000
    assert 430000 < g[1] < 490000 and 430000 < g[2] < 490000
    # the reference's check of the Ice1r melt function on the final state
    h = m.fields.gh.h
    draft = -(918.0 / 1028.0) * h
    cavity = np.maximum(draft - h, 0.0)
    melt = 0.2 * np.tanh(cavity / 75.0) * np.maximum(-100.0 - draft, 0.0)
    assert np.all(melt >= 0) and melt.max() < 100
    print("[gpu] mismip+ Ice0 grounding line (m): %s" % np.array2string(glx, precision=1), flush=True)
