"""Batched update_velocities over an ensemble of independent models
(wavi_update_velocities_batch): every member's velocities and report must be
bit-identical to solving that member alone with wavi_update_velocities, for
both precisions, whatever the number of host threads driving the batch."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _step(name):
    print("[gpu] %s" % name, flush=True)
    sys.stdout.flush()


def _small_specs(W, n, seed):
    specs = [sp for sp in W.workloads.ensemble_specs(64, seed) if sp["nx"] <= 320]
    return specs[:n]


def _members(W, specs, precision):
    return [W.workloads.ensemble_member(sp, precision, maxiter_picard=6) for sp in specs]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_individual_solves(W, precision):
    specs = _small_specs(W, 6, seed=5)
    assert len({(s["nx"], s["nσ"]) for s in specs}) > 1  # a ragged batch
    _step("ensemble %s: %d individual solves" % (precision, len(specs)))
    alone = _members(W, specs, precision)
    for m in alone:
        W.update_velocities(m)
    _step("ensemble %s: one batched call, 4 threads" % precision)
    batch = _members(W, specs, precision)
    W.update_velocities_batch(batch, nthreads=4)
    for a, b in zip(alone, batch):
        assert a.last_report["picard_iters"] == b.last_report["picard_iters"]
        assert a.last_report["coarse_cg_iters"] == b.last_report["coarse_cg_iters"]
        np.testing.assert_array_equal(a.fields.gu.u, b.fields.gu.u)
        np.testing.assert_array_equal(a.fields.gv.v, b.fields.gv.v)
        np.testing.assert_array_equal(a.fields.gh.av_speed, b.fields.gh.av_speed)


def test_batch_edge_cases(W):
    _step("ensemble: empty batch, one member, more threads than members")
    assert W.update_velocities_batch([], nthreads=4) == []
    specs = _small_specs(W, 2, seed=7)
    alone = _members(W, specs, "mixed")
    for m in alone:
        W.update_velocities(m)
    one = _members(W, specs[:1], "mixed")
    W.update_velocities_batch(one, nthreads=16)
    np.testing.assert_array_equal(alone[0].fields.gu.u, one[0].fields.gu.u)
    two = _members(W, specs, "mixed")
    W.update_velocities_batch(two, nthreads=1)
    for a, b in zip(alone, two):
        np.testing.assert_array_equal(a.fields.gv.v, b.fields.gv.v)
    _step("ensemble: a context twice is rejected before any member runs")
    with pytest.raises(W.ArgumentError):
        W.update_velocities_batch([two[0], two[0]])
    import ctypes
    lib = W._lib.load()
    c = two[0].parallel_spec.ctx
    ctxs = (ctypes.c_void_p * 3)(two[1].parallel_spec.ctx, c, c)
    assert lib.wavi_update_velocities_batch(ctxs, 3, 2, None, None) == W._lib.WAVI_EARG
    assert b"twice" in lib.wavi_last_error(None)


def test_gradient_batch_matches_individual_calls(W):
    specs = _small_specs(W, 4, seed=9)
    _step("ensemble gradients: %d individual calls" % len(specs))
    obs = []
    for sp in specs:
        m = W.workloads.ensemble_member(sp, "mixed", maxiter_picard=3)
        W.update_geometry(m)
        nx, ny = sp["nx"], sp["ny"]
        W.misfit_gradient(m, 3, np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1)), gradient=False)
        obs.append(W.workloads.inversion_observations(m.fields.gu.u, m.fields.gv.v, seed=sp["seed"]))

    def members():
        out = []
        for sp in specs:
            m = W.workloads.ensemble_member(sp, "mixed", maxiter_picard=3)
            W.update_geometry(m)
            out.append(m)
        return out

    alone = [W.misfit_gradient(m, 3, uo, vo) for m, (uo, vo) in zip(members(), obs)]
    _step("ensemble gradients: one batched call, 3 threads")
    batch = W.misfit_gradient_batch(members(), 3, obs, nthreads=3)
    for (Ja, ga, ra), (Jb, gb, rb) in zip(alone, batch):
        assert Ja == Jb
        np.testing.assert_array_equal(ga, gb)
        assert ra == rb


def test_batch_while_other_contexts_are_created(W):
    """Contexts created, filled, read and destroyed on the calling thread
    while a mixed-precision batch (whose coarse CG captures a HIP graph per
    member) runs on another thread: the library serialises its allocations,
    frees and synchronous copies with the captures, so the batch completes
    and equals the members solved alone."""
    import threading
    specs = _small_specs(W, 4, seed=11)
    _step("ensemble under concurrent context churn: individual solves")
    alone = _members(W, specs, "mixed")
    for m in alone:
        W.update_velocities(m)
    batch = _members(W, specs, "mixed")
    errors = []

    def run_batch():
        try:
            W.update_velocities_batch(batch, nthreads=4)
        except Exception as ex:  # reported below, on the test's thread
            errors.append(ex)

    _step("ensemble under concurrent context churn: batch on a thread, contexts created here")
    t = threading.Thread(target=run_batch)
    t.start()
    churn = 0
    while t.is_alive() or churn < 3:
        m = _members(W, specs[:1], "fp64")[0]
        W.update_geometry(m)  # creates the context, sets and reads fields
        m.parallel_spec.close()
        churn += 1
    t.join()
    assert not errors, errors
    for a, b in zip(alone, batch):
        np.testing.assert_array_equal(a.fields.gu.u, b.fields.gu.u)
        np.testing.assert_array_equal(a.fields.gv.v, b.fields.gv.v)


def test_lockstep_matches_individual_solves(W):
    """wavi_update_velocities_lockstep: the members' coarse CGs share one
    launch per iteration; every member must still be bit-identical to its
    own solve, also when members stop their Picard loops at different
    iterations."""
    specs = _small_specs(W, 6, seed=5)
    _step("lockstep: %d individual solves" % len(specs))

    def members():
        ms = _members(W, specs, "mixed")
        ms[1].solver_params.maxiter_picard = 3  # leaves the lockstep early
        return ms

    alone = members()
    for m in alone:
        W.update_velocities(m)
    _step("lockstep: one call")
    lock = members()
    W.update_velocities_batch(lock, lockstep=True)
    for a, b in zip(alone, lock):
        assert a.last_report["picard_iters"] == b.last_report["picard_iters"]
        assert a.last_report["coarse_cg_iters"] == b.last_report["coarse_cg_iters"]
        np.testing.assert_array_equal(a.fields.gu.u, b.fields.gu.u)
        np.testing.assert_array_equal(a.fields.gv.v, b.fields.gv.v)
    assert lock[1].last_report["picard_iters"] == 3
    _step("lockstep: one member; FP64 members rejected")
    one = members()[:1]
    W.update_velocities_batch(one, lockstep=True)
    np.testing.assert_array_equal(alone[0].fields.gu.u, one[0].fields.gu.u)
    with pytest.raises(W.ArgumentError):
        W.update_velocities_batch(_members(W, specs[:2], "fp64"), lockstep=True)


def test_gradient_batch_two_pools(W):
    """Small mixed members meet at their coarse CGs on threads of their own
    while an FP64 member (not eligible to meet) runs on the other pool: all
    bit-identical to their own calls."""
    specs = _small_specs(W, 4, seed=13)
    precs = ["mixed", "fp64", "mixed", "mixed"]
    _step("ensemble gradients, two pools: individual calls")

    def members():
        out = []
        for sp, p in zip(specs, precs):
            m = W.workloads.ensemble_member(sp, p, maxiter_picard=2)
            W.update_geometry(m)
            out.append(m)
        return out

    obs = []
    for m, sp in zip(members(), specs):
        nx, ny = sp["nx"], sp["ny"]
        W.misfit_gradient(m, 2, np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1)), gradient=False)
        obs.append(W.workloads.inversion_observations(m.fields.gu.u, m.fields.gv.v, seed=sp["seed"]))
    alone = [W.misfit_gradient(m, 2, uo, vo) for m, (uo, vo) in zip(members(), obs)]
    _step("ensemble gradients, two pools: one batched call")
    batch = W.misfit_gradient_batch(members(), 2, obs, nthreads=2)
    for (Ja, ga, ra), (Jb, gb, rb) in zip(alone, batch):
        assert Ja == Jb
        np.testing.assert_array_equal(ga, gb)
        assert ra == rb


def test_gradient_batch_members_with_either_coarse_cg(W):
    """Small mixed members, two of them with the plain coarse CG in their
    linear solves (wavi_set_linear_coarse_jacobi(ctx, 0)) and two with the
    Jacobi-preconditioned one: only members of one kind share coarse-CG
    launches, and every member is bit-identical to its own call."""
    specs = _small_specs(W, 4, seed=19)
    jacobi = [1, 0, 1, 0]
    _step("ensemble gradients, plain and Jacobi coarse CGs: individual calls")

    def members():
        out = []
        for sp, on in zip(specs, jacobi):
            m = W.workloads.ensemble_member(sp, "mixed", maxiter_picard=2)
            W.update_geometry(m)
            spec = m.parallel_spec
            assert spec._lib.wavi_set_linear_coarse_jacobi(spec.ctx, on) == 0
            out.append(m)
        return out

    obs = []
    for m, sp in zip(members(), specs):
        nx, ny = sp["nx"], sp["ny"]
        W.misfit_gradient(m, 2, np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1)), gradient=False)
        obs.append(W.workloads.inversion_observations(m.fields.gu.u, m.fields.gv.v, seed=sp["seed"]))
    alone = [W.misfit_gradient(m, 2, uo, vo) for m, (uo, vo) in zip(members(), obs)]
    _step("ensemble gradients, plain and Jacobi coarse CGs: one batched call")
    batch = W.misfit_gradient_batch(members(), 2, obs, nthreads=2)
    for (Ja, ga, ra), (Jb, gb, rb) in zip(alone, batch):
        assert Ja == Jb
        np.testing.assert_array_equal(ga, gb)
        assert ra == rb


def test_gradient_batch_more_members_than_rendezvous_threads(W):
    """36 small members: more than the 32 threads the meeting members get, so
    threads leave the rendezvous with one member and join with the next."""
    specs = [sp for sp in W.workloads.ensemble_specs(512, 17) if sp["nx"] == 160][:36]
    assert len(specs) == 36
    _step("ensemble gradients: 36 small members, individual calls")

    def members():
        out = []
        for sp in specs:
            m = W.workloads.ensemble_member(sp, "mixed", maxiter_picard=2)
            W.update_geometry(m)
            out.append(m)
        return out

    obs = []
    for m, sp in zip(members(), specs):
        nx, ny = sp["nx"], sp["ny"]
        W.misfit_gradient(m, 2, np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1)), gradient=False)
        obs.append(W.workloads.inversion_observations(m.fields.gu.u, m.fields.gv.v, seed=sp["seed"]))
    alone = [W.misfit_gradient(m, 2, uo, vo) for m, (uo, vo) in zip(members(), obs)]
    _step("ensemble gradients: 36 small members, one batched call")
    batch = W.misfit_gradient_batch(members(), 2, obs, nthreads=4)
    for (Ja, ga, ra), (Jb, gb, rb) in zip(alone, batch):
        assert Ja == Jb
        np.testing.assert_array_equal(ga, gb)
        assert ra == rb


def test_lockstep_with_irregular_members(W):
    """Lockstep members on irregular domains (calving fronts that move from
    row to row, interior holes, interior Dirichlet points) beside channel
    members: each bit-identical to its own solve."""
    from helpers import ragged
    specs = _small_specs(W, 2, seed=9)
    _step("lockstep with irregular members")

    def members():
        ms = _members(W, specs, "mixed")
        ms += [ragged(W, nx=40, ny=24, precision="mixed", seed=s, maxiter_picard=6) for s in (7, 11)]
        ms.append(ragged(W, nx=24, ny=12, precision="mixed", seed=3, maxiter_picard=4, dirichlet=False))
        for m in ms:
            W.update_geometry(m)
        return ms

    alone = members()
    for m in alone:
        W.update_velocities(m)
    lock = members()
    W.update_velocities_batch(lock, lockstep=True)
    for a, b in zip(alone, lock):
        assert a.last_report["coarse_cg_iters"] == b.last_report["coarse_cg_iters"]
        np.testing.assert_array_equal(a.fields.gu.u, b.fields.gu.u)
        np.testing.assert_array_equal(a.fields.gv.v, b.fields.gv.v)


@pytest.mark.parametrize("levels", [0, 1, 2])
def test_lockstep_other_haar_levels(W, levels):
    """k_cgm_iter_batch at Haar levels 0..2 (blocks of 2, 4, 8): lockstep
    members bit-identical to their own solves."""
    specs = _small_specs(W, 3, seed=21)
    _step("lockstep levels %d" % levels)

    def members():
        ms = _members(W, specs, "mixed")
        for m in ms:
            m.solver_params.levels = levels
            m.solver_params.maxiter_picard = 3
        return ms

    alone = members()
    for m in alone:
        W.update_velocities(m)
    lock = members()
    W.update_velocities_batch(lock, lockstep=True)
    for a, b in zip(alone, lock):
        assert a.last_report["coarse_cg_iters"] == b.last_report["coarse_cg_iters"]
        np.testing.assert_array_equal(a.fields.gu.u, b.fields.gu.u)
        np.testing.assert_array_equal(a.fields.gv.v, b.fields.gv.v)
