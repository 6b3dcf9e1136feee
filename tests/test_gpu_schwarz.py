"""Restricted additive Schwarz (SharedMemorySpec) on the GPU: the
reference's own test (test/domain_decomposition_tests/
test_domains_shared_memory.jl): MISMIP+ 80 x 10, 16 x 2 subdomains with an
overlap of 2, one Schwarz iteration per Picard iteration, 1000 steps of
0.1 a; the thickness within rtol 0.01 (isapprox) of the v1.8.3 golden of
the plain solver."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _step(name):
    print("[gpu] %s" % name, flush=True)
    sys.stdout.flush()


def _mismip_ras(W, precision="fp64", nsteps=None):
    grid = W.Grid(nx=80, ny=10, nσ=4, x0=0.0, y0=-40000.0, dx=8000.0, dy=8000.0,
                  u_iszero=["north"], v_iszero=["west", "east"])
    spec = W.SharedMemorySpec(ngridsx=16, ngridsy=2, overlap=2, niterations=1, precision=precision)
    params = W.Params(default_thickness=100.0, accumulation_rate=0.3, default_temperature=265.700709)
    model = W.Model(grid=grid, bed_elevation=W.mismip_plus_bed, params=params,
                    solver_params=W.SolverParams(maxiter_picard=1), parallel_spec=spec)
    tp = W.TimesteppingParams(dt=0.1, end_time=100.0) if nsteps is None else \
        W.TimesteppingParams(dt=0.1, n_iter_total=nsteps)
    sim = W.Simulation(model=model, timestepping_params=tp)
    W.run_simulation(sim)
    return sim


@pytest.mark.timeout(600)
def test_shared_memory_spec_against_golden(W):
    _step("schwarz mismip+ 1000 steps")
    sim = _mismip_ras(W)
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "mismip_plus_v1_8_3.npz"))
    h = sim.model.fields.gh.h
    err = np.linalg.norm(h - d["h"]) / max(np.linalg.norm(h), np.linalg.norm(d["h"]))
    print("[gpu] schwarz thickness rel %.3e" % err, flush=True)
    assert err <= 0.01
    assert np.all(np.isfinite(sim.model.fields.gu.u))


def test_shared_memory_spec_few_steps_close_to_plain_solver(W):
    """Ten steps of RAS against ten of the plain V-cycle solver, both on the
    GPU: the Schwarz preconditioner gives another approximation of the same
    velocities."""
    _step("schwarz mismip+ 10 steps vs plain")
    ras = _mismip_ras(W, nsteps=10)
    m = W.workloads.mismip_plus_fixture()
    W.run_simulation(W.Simulation(model=m, timestepping_params=W.TimesteppingParams(dt=0.1, n_iter_total=10)))
    h1, h0 = ras.model.fields.gh.h, m.fields.gh.h
    err = np.linalg.norm(h1 - h0) / np.linalg.norm(h0)
    print("[gpu] 10 steps: thickness rel %.3e" % err, flush=True)
    assert err <= 0.01


def test_device_schwarz_round_equals_host_restatement(W):
    """One precondition!(model, ::SharedMemorySpec) on the device
    (wavi_schwarz_precondition: two Schwarz rounds over 4 x 2 subdomains,
    overlap 2, damping 0.3) against the same rounds restated on the host
    with the reference's steps (schwarz_restrict_velocities, update_state! of
    every subdomain through wavi_update_state_batch,
    schwarz_prolong_velocities in subdomain order): the velocities and the
    residual fields bit for bit, the relative residual to rounding."""
    import copy
    import ctypes
    _step("schwarz: device round vs host restatement")

    def make():
        grid = W.Grid(nx=80, ny=10, nσ=4, x0=0.0, y0=-40000.0, dx=8000.0, dy=8000.0,
                      u_iszero=["north"], v_iszero=["west", "east"])
        spec = W.SharedMemorySpec(ngridsx=4, ngridsy=2, overlap=2, niterations=2, damping=0.3)
        params = W.Params(default_thickness=100.0, accumulation_rate=0.3, default_temperature=265.700709)
        m = W.Model(grid=grid, bed_elevation=W.mismip_plus_bed, params=params,
                    solver_params=W.SolverParams(maxiter_picard=1), parallel_spec=spec)
        m.params.dt = 0.1
        W.update_geometry(m)
        W.inner_update(m)
        return m

    S = W.schwarz
    dev, host = make(), make()
    for m in (dev, host):
        S.update_preconditioner_schwarz(m)
        W.push_fields(m)
    conv, rel = S.precondition_schwarz(dev)
    W.pull_fields(dev)
    # the host restatement
    spec = host.parallel_spec
    lib = spec._lib
    bu, bv = W.get_rhs(host)
    f = host.fields
    xu = np.where(f.gu.mask_inner, f.gu.u, 0.0)
    xv = np.where(f.gv.mask_inner, f.gv.v, 0.0)
    au, av = W.op_apply(host, xu, xv)
    ru, rv = bu - au, bv - av
    rel_h = np.linalg.norm(np.concatenate([ru[f.gu.mask_inner], rv[f.gv.mask_inner]])) / \
        np.linalg.norm(np.concatenate([bu[f.gu.mask_inner], bv[f.gv.mask_inner]]))
    assert not conv and abs(rel - rel_h) <= 1e-13 * rel_h
    np.testing.assert_array_equal(dev.fields.gu.residual[f.gu.mask_inner], ru[f.gu.mask_inner])
    np.testing.assert_array_equal(dev.fields.gv.residual[f.gv.mask_inner], rv[f.gv.mask_inner])
    members = [(ig, jg, spec.schwarz_model_array[ig][jg]) for ig in range(4) for jg in range(2)]
    for _ in range(spec.niterations):
        for ig, jg, mg in members:
            S.schwarz_restrict_velocities(mg, host, ig + 1, jg + 1, 4, 2, 2)
            W.push_fields(mg)
        ctxs = (ctypes.c_void_p * 8)(*[mg.parallel_spec.ctx for _, _, mg in members])
        assert lib.wavi_update_state_batch(ctxs, 8, 4, None, None) == 0
        for _, _, mg in members:
            W.pull_fields(mg)
        f.gu.u[:, :] = spec.damping * f.gu.u
        f.gv.v[:, :] = spec.damping * f.gv.v
        for ig, jg, mg in members:
            S.schwarz_prolong_velocities(host, mg, ig + 1, jg + 1, 4, 2, 2, spec.damping)
    np.testing.assert_array_equal(dev.fields.gu.u, host.fields.gu.u)
    np.testing.assert_array_equal(dev.fields.gv.v, host.fields.gv.v)


def test_device_subdomain_refresh_equals_host_build(W):
    """update_preconditioner!(model, ::SharedMemorySpec) on the device
    (wavi_schwarz_refresh) gives every subdomain context the fields that
    schwarzModel and setup_fields build on the host, bit for bit: the host
    build is pushed, the subdomains are then stepped (update_state!) so that
    every field changes, and the device refresh from the same full state
    brings each field back."""
    import ctypes
    _step("schwarz: device refresh vs host build")
    S = W.schwarz
    grid = W.Grid(nx=80, ny=10, nσ=4, x0=0.0, y0=-40000.0, dx=8000.0, dy=8000.0,
                  u_iszero=["north"], v_iszero=["west", "east"])
    spec = W.SharedMemorySpec(ngridsx=4, ngridsy=2, overlap=2, niterations=1)
    params = W.Params(default_thickness=100.0, accumulation_rate=0.3, default_temperature=265.700709)
    m = W.Model(grid=grid, bed_elevation=W.mismip_plus_bed, params=params,
                solver_params=W.SolverParams(maxiter_picard=2), parallel_spec=spec)
    m.params.dt = 0.1
    W.update_geometry(m)
    W.update_velocities(m)  # builds the subdomains on the host and solves once
    W.update_geometry(m)
    W.inner_update(m)
    spec.__dict__.pop("_subdomains_built")
    S.update_preconditioner_schwarz(m)  # host build from the current state
    members = [spec.schwarz_model_array[ig][jg] for ig in range(4) for jg in range(2)]
    names = [fid for _, _, fid in W.model.FIELD_MAP]
    lib = spec._lib

    def device_fields(mg):
        out = {}
        for fid in W._lib.FIELD_NAMES:
            n = ctypes.c_size_t(0)
            assert lib.wavi_field_size(mg.parallel_spec.ctx, W._lib.FIELD[fid], ctypes.byref(n)) == 0
            a = np.empty(n.value)
            assert lib.wavi_get_field(mg.parallel_spec.ctx, W._lib.FIELD[fid], W._lib.dptr(a), a.size) == 0
            out[fid] = a
        return out

    built = [device_fields(mg) for mg in members]
    ctxs = (ctypes.c_void_p * 8)(*[mg.parallel_spec.ctx for mg in members])
    assert lib.wavi_update_state_batch(ctxs, 8, 4, None, None) == 0
    stepped = [device_fields(mg) for mg in members]
    assert any(not np.array_equal(stepped[0][f], built[0][f]) for f in names)
    W.push_fields(m)
    assert lib.wavi_schwarz_refresh(spec.ctx, ctxs, 4, 2, 2) == 0
    for k, mg in enumerate(members):
        got = device_fields(mg)
        for fid in W._lib.FIELD_NAMES:
            assert np.array_equal(got[fid], built[k][fid]), (k, fid)
    assert lib.wavi_schwarz_refresh(spec.ctx, ctxs, 3, 2, 2) == W._lib.WAVI_EARG
