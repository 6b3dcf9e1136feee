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
