"""The reference's verification experiments (test/verification_tests.jl),
run whole on the GPU: the freely floating, spinning iceberg against its
closed form, and MISMIP+ Ice0 (80 x 10 cells of 8 km, nsigma 4, one Picard
iteration per step, dt 0.5 a for 10,000 years) against the grounding line's
symmetry across the channel and its position."""
import sys

import numpy as np
import pytest

from helpers import iceberg

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


def test_iceberg_closed_form(W):
    """test/verification_tests.jl:3-10 with iceberg_test.jl:85-93: 10,000
    steps of 0.1 a; steady thickness h0 and linear spreading plus the solid
    body motion."""
    print("[gpu] iceberg: 10000 steps of 0.1 a", flush=True)
    model, solid_u, solid_v = iceberg(W)
    sim = W.Simulation(model=model, timestepping_params=W.TimesteppingParams(dt=0.1, end_time=1000.0))
    W.run_simulation(sim)
    f, g = model.fields, model.grid
    h0 = ((36.0 * 0.3 / (2.0e-17)) * (1.0 / (9.81 * 918.0 * (1 - 918.0 / 1028.0))) ** 3) ** (1.0 / 4.0)
    u0 = g.xxu * 0.3 / (2.0 * h0) + solid_u
    v0 = g.yyv * 0.3 / (2.0 * h0) + solid_v
    hm, um, vm = f.gh.mask, f.gu.mask, f.gv.mask
    relerr_h = np.linalg.norm(f.gh.h[hm] - h0) / np.linalg.norm(h0 * np.ones(hm.sum()))
    relerr_u = np.linalg.norm(f.gu.u[um] - u0[um]) / np.linalg.norm(u0[um])
    relerr_v = np.linalg.norm(f.gv.v[vm] - v0[vm]) / np.linalg.norm(v0[vm])
    print("[gpu] iceberg relative errors: h %.2e u %.2e v %.2e" % (relerr_h, relerr_u, relerr_v), flush=True)
    assert relerr_h < 1e-4
    assert relerr_u < 3e-4
    assert relerr_v < 3e-4
