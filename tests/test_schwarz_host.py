"""Host-side pieces of the restricted additive Schwarz spec
(src/SharedMemorySpec.jl): partition of unity, subdomain bounds and models,
restriction and prolongation (no GPU needed)."""
import numpy as np
import pytest


def test_partition_of_unity_known_values(W):
    # ramps (i-1)/overlap and (m-i)/overlap, capped at 1, on the edges not left at 1
    p = W.schwarz_partition_of_unity(5, 3, False, False, True, True, 2, 1)
    np.testing.assert_array_equal(p[:, 0], [0.0, 0.5, 1.0, 0.5, 0.0])
    np.testing.assert_array_equal(p[:, 0], p[:, 2])
    q = W.schwarz_partition_of_unity(4, 4, True, False, False, True, 3, 1)
    np.testing.assert_array_equal(q[:, 1], [1.0, 2 / 3, 1 / 3, 0.0])
    np.testing.assert_array_equal(q[0, :], [0.0, 1.0, 1.0, 1.0])
    with pytest.raises(W.ArgumentError):
        W.schwarz_partition_of_unity(4, 4, False, False, True, True, 2, 1)


def test_partitions_sum_to_one_away_from_fixed_halos(W):
    """With damping 0, prolonging the restriction of a field gives the field
    back wherever the overlapping ramps add up to one: the u points of the
    interior of every subdomain and of the overlaps."""
    m = W.workloads.mismip_plus_fixture()
    spec = dict(ngridsx=8, ngridsy=2, overlap=2)
    f = m.fields
    rng = np.random.default_rng(0)
    u0 = rng.standard_normal(f.gu.u.shape)
    f.gu.u[:, :] = u0
    f.gv.v[:, :] = rng.standard_normal(f.gv.v.shape)
    subs = {}
    for ig in range(1, 9):
        for jg in range(1, 3):
            mg = W.schwarz_model(m, ig, jg, overlap=2, ngridsx=8, ngridsy=2)
            W.schwarz_restrict_velocities(mg, m, ig, jg, overlap=2, ngridsx=8, ngridsy=2)
            subs[ig, jg] = mg
    f.gu.u[:, :] = 0.0
    f.gv.v[:, :] = 0.0
    for (ig, jg), mg in subs.items():
        W.schwarz_prolong_velocities(m, mg, ig, jg, overlap=2, ngridsx=8, ngridsy=2, damping=0.0)
    # the u ramps have length 2*overlap over 2*overlap + 1 shared points: they sum to one
    np.testing.assert_allclose(f.gu.u[:, 2:-2], u0[:, 2:-2], rtol=0, atol=1e-14)


def test_schwarz_model_halos_and_bounds(W):
    m = W.workloads.mismip_plus_fixture()
    mg = W.schwarz_model(m, 2, 1, ngridsx=16, ngridsy=2, overlap=2)
    # x cells 3..11 (0-based: 5 per subdomain, 2 of overlap each side) and y cells 0..6
    assert (mg.grid.nx, mg.grid.ny) == (9, 7)
    assert mg.grid.x0 == m.grid.x0 + 3 * m.grid.dx
    assert mg.grid.u_isfixed[0, :].all() and mg.grid.u_isfixed[-1, :].all()
    assert mg.grid.u_isfixed[:, -1].all() and not mg.grid.u_isfixed[1:-1, 0].any()
    np.testing.assert_array_equal(mg.fields.gh.h, m.fields.gh.h[3:12, 0:7])
    with pytest.raises(W.BoundsError):
        W.schwarz_model(m, 17, 1, ngridsx=16, ngridsy=2)
    with pytest.raises(W.ArgumentError):
        W.schwarz_model(m, 1, 1, ngridsx=3, ngridsy=2)


def test_timestepping_checkpoint_steps_and_output_params(W, tmp_path):
    tp = W.TimesteppingParams(dt=0.5, n_iter_total=40, chkpt_freq=2.0, pchkpt_freq=10.0, chkpt_path=str(tmp_path))
    assert (tp.n_iter_chkpt, tp.n_iter_pchkpt) == (4, 20)
    assert tp.chkpt_path.endswith("/")
    with pytest.warns(UserWarning):
        assert W.TimesteppingParams(dt=1.0, n_iter_total=2, chkpt_path=str(tmp_path / "missing")).chkpt_path == "./"
    with pytest.raises(W.ArgumentError):
        W.OutputParams(outputs={"h": None})


def test_pickup_restores_fields_clock_and_saved_dt(W, tmp_path, monkeypatch):
    """pickup! from PChkpt_<niter0>: the fields and the clock come back, and
    the model's dt is the saved one (the reference replaces the model with
    the saved model, Simulation.jl:77-79), while the new TimesteppingParams
    keeps its own dt for the thickness steps."""
    monkeypatch.chdir(tmp_path)
    m = W.workloads.mismip_plus_fixture()
    sim = W.Simulation(model=m, timestepping_params=W.TimesteppingParams(dt=0.1, n_iter_total=20))
    m.fields.gh.h[:] = m.fields.gh.h + 1.25
    m.fields.gu.u[:] = 3.0
    sim.clock = W.Clock(10, 1.0)
    W.save_checkpoint(sim, "PChkpt_%010d.npz" % 10)
    m2 = W.workloads.mismip_plus_fixture()
    sim2 = W.Simulation(model=m2, timestepping_params=W.TimesteppingParams(niter0=10, dt=0.2, n_iter_total=20))
    assert (sim2.clock.n_iter, sim2.clock.time) == (10, 1.0)
    assert m2.params.dt == 0.1 and sim2.timestepping_params.dt == 0.2
    np.testing.assert_array_equal(m2.fields.gh.h, m.fields.gh.h)
    np.testing.assert_array_equal(m2.fields.gu.u, m.fields.gu.u)
