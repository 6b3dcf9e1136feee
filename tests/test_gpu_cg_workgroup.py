"""The coarse CG run whole in one workgroup (wavi_set_coarse_cg with
WAVI_CG_WORKGROUP: k_cg_solve1 in FP64, k_cgm_solve1 in mixed precision)
against the same CG as launches over all tiles (the default).  The
workgroup runs each part of an iteration on the tiles in turn, with the
same per-tile code, so the bar is bit equality: with the oracle in FP64,
between the two modes in mixed precision, and between a lockstep batch
(whose CG is one launch per iteration over all members) and lone solves in
the workgroup."""
import ctypes
import sys

import numpy as np
import pytest

from helpers import channel, ragged

pytestmark = pytest.mark.gpu


def _step(name):
    print("[gpu] %s" % name, flush=True)
    sys.stdout.flush()


def _mode(spec):
    got = ctypes.c_int()
    assert spec._lib.wavi_coarse_cg(spec.ctx, ctypes.byref(got)) == 0
    return got.value


def _solve(W, m, mode):
    """update_velocities! with the coarse CG in the given mode; the
    velocities, the report and the mode the library reports."""
    W.update_preconditioner(m)
    W.push_fields(m)
    spec = m.parallel_spec
    lib = spec._lib
    auto = W._lib.WAVI_CG_LAUNCHES
    assert _mode(spec) == auto
    assert lib.wavi_set_coarse_cg(spec.ctx, mode) == 0
    assert lib.wavi_update_geometry(spec.ctx) == 0
    rep = W._lib.SolveReport()
    assert lib.wavi_update_velocities(spec.ctx, ctypes.byref(rep)) == 0
    W.pull_fields(m)
    assert _mode(spec) == (auto if mode == W._lib.WAVI_CG_AUTO else mode)
    vel = np.concatenate([m.fields.gu.u.ravel(order="F"), m.fields.gv.v.ravel(order="F")])
    return vel, rep.as_dict()


@pytest.mark.parametrize("nx,ny,maxiter_coarse", [(80, 10, 1000), (40, 24, 1000), (70, 40, 1000), (40, 24, 7)])
def test_fp64_coarse_cg_in_one_workgroup_equals_oracle(W, oracle, nx, ny, maxiter_coarse):
    """FP64, 3 Picard iterations on grids of 6 (80 x 10, 40 x 24) and 15
    (70 x 40) tiles; with maxiter_coarse = 7 every coarse CG stops at its
    cap.  Both modes give the oracle's velocities, CG counts and residual
    history."""
    ref = None
    for mode in (W._lib.WAVI_CG_LAUNCHES, W._lib.WAVI_CG_WORKGROUP):
        _step("fp64 coarse CG mode %d at %dx%d, maxiter_coarse %d" % (mode, nx, ny, maxiter_coarse))
        m = channel(W, nx=nx, ny=ny, maxiter_picard=3, maxiter_coarse=maxiter_coarse)
        if ref is None:
            o = oracle.OracleModel(m)
            o.update_geometry()
            orep = o.update_velocities()
            ref = np.concatenate([o.field("U_U").ravel(order="F"), o.field("V_V").ravel(order="F")])
        vel, rep = _solve(W, m, mode)
        assert rep == orep
        assert np.array_equal(vel, ref)
    if maxiter_coarse == 7:
        assert all(k == 7 for k in orep["coarse_cg_iters"][:orep["picard_iters"]])


def test_fp64_coarse_cg_in_one_workgroup_ragged_domain(W, oracle):
    """The irregular domain (moving calving front, holes, interior Dirichlet
    points; 6 tiles): the workgroup's CG gives the oracle's bits."""
    _step("fp64 coarse CG in one workgroup, ragged 40x24")
    m = ragged(W, nx=40, ny=24, maxiter_picard=3)
    o = oracle.OracleModel(m)
    o.update_geometry()
    orep = o.update_velocities()
    ref = np.concatenate([o.field("U_U").ravel(order="F"), o.field("V_V").ravel(order="F")])
    vel, rep = _solve(W, m, W._lib.WAVI_CG_WORKGROUP)
    assert rep == orep
    assert np.array_equal(vel, ref)


@pytest.mark.parametrize("levels", [0, 1, 2, 3])
@pytest.mark.parametrize("nx,ny,maxiter_coarse", [(80, 10, 1000), (96, 24, 1000), (80, 10, 9)])
def test_mixed_coarse_cg_modes_bit_equal(W, levels, nx, ny, maxiter_coarse):
    """Mixed precision, 6 Picard iterations, Haar levels 0..3: the one-
    workgroup CG (each k_cgm_solve1 instantiation) and one launch per
    iteration give the same bits, also when the CG stops at its cap."""
    runs = []
    for mode in (W._lib.WAVI_CG_LAUNCHES, W._lib.WAVI_CG_WORKGROUP, W._lib.WAVI_CG_AUTO):
        _step("mixed coarse CG mode %d, levels %d, %dx%d, maxiter_coarse %d" % (mode, levels, nx, ny, maxiter_coarse))
        m = channel(W, nx=nx, ny=ny, precision="mixed", maxiter_picard=6, maxiter_coarse=maxiter_coarse)
        m.solver_params.levels = levels
        runs.append(_solve(W, m, mode))
    for vel, rep in runs[1:]:
        assert rep == runs[0][1]
        assert np.array_equal(vel, runs[0][0])
    if maxiter_coarse == 9:
        rep = runs[0][1]
        assert all(k == 9 for k in rep["coarse_cg_iters"][:rep["picard_iters"]])


def test_lockstep_batch_equals_lone_workgroup_solves(W):
    """Small mixed members solved alone with the one-workgroup CG, and in a
    lockstep batch, whose CGs share one launch per iteration: every member
    is bit-identical either way."""
    shapes = [(80, 10), (48, 12), (40, 24), (80, 10)]

    def members():
        ms = [channel(W, nx=nx, ny=ny, precision="mixed", maxiter_picard=5, seed=11 + k)
              for k, (nx, ny) in enumerate(shapes)]
        for m in ms:
            W.update_geometry(m)
        return ms

    _step("lockstep vs lone one-workgroup solves: %d members" % len(shapes))
    alone = members()
    for m in alone:
        spec = m.parallel_spec
        assert spec._lib.wavi_set_coarse_cg(spec.ctx, W._lib.WAVI_CG_WORKGROUP) == 0
        W.update_velocities(m)
        assert _mode(spec) == W._lib.WAVI_CG_WORKGROUP
        assert np.isfinite(m.last_report["rel_resid"])
    lock = members()
    W.update_velocities_batch(lock, lockstep=True)
    for a, b in zip(alone, lock):
        assert a.last_report == b.last_report
        np.testing.assert_array_equal(a.fields.gu.u, b.fields.gu.u)
        np.testing.assert_array_equal(a.fields.gv.v, b.fields.gv.v)


def test_coarse_cg_mode_rejects_unknown(W):
    _step("coarse CG mode: unknown mode rejected")
    m = channel(W, nx=40, ny=24)
    W.update_preconditioner(m)
    spec = m.parallel_spec
    assert spec._lib.wavi_set_coarse_cg(spec.ctx, 3) == W._lib.WAVI_EARG
    assert spec._lib.wavi_set_coarse_cg(spec.ctx, -1) == W._lib.WAVI_EARG
    assert spec._lib.wavi_set_coarse_cg(None, 0) == W._lib.WAVI_EARG
