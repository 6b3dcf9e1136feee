"""Parity of the HIP library (FP64 path) with the CPU oracle.

The FP64 path evaluates every expression in the oracle's order, reduces with
the same blocked tree and uses the same deterministic power function, so the
bar here is bit equality (an ulp-level difference would show up as a failure
and point at the kernel that broke the order)."""
import ctypes
import sys

import numpy as np
import pytest

from helpers import channel, iceberg, ragged, rel_l2

pytestmark = pytest.mark.gpu


def _step(name):
    print("[gpu] %s" % name, flush=True)
    sys.stdout.flush()


def _gpu_model(W, model):
    W.update_preconditioner(model)
    W.push_fields(model)
    return model.parallel_spec


def _assert_fields_equal(W, model, o, names=None, rtol=0.0):
    W.pull_fields(model)
    bad = []
    for cont, attr, fid in W.model.FIELD_MAP:
        if names is not None and fid not in names:
            continue
        a = np.asarray(model._array(cont, attr), dtype=float)
        b = o.field(fid).reshape(a.shape, order="F")
        if rtol == 0.0:
            ok = np.array_equal(a, b, equal_nan=True)
        else:
            ok = rel_l2(a, b) <= rtol
        if not ok:
            d = np.abs(a - b)
            bad.append("%s max|diff| %.3e rel %.3e" % (fid, np.nanmax(d), rel_l2(a, b)))
    assert not bad, bad


def _rand_velocities(model, seed, scale=100.0):
    rng = np.random.default_rng(seed)
    f = model.fields
    f.gu.u[:] = scale * (1.0 + rng.standard_normal(f.gu.u.shape))
    f.gv.v[:] = 0.2 * scale * rng.standard_normal(f.gv.v.shape)


def test_device_math_matches_cpu(W, oracle):
    _step("math probe")
    lib = W._lib.load()
    rng = np.random.default_rng(0)
    n = 50000
    x = np.exp(rng.uniform(-30, 30, n))
    y = rng.choice([-2.0 / 3.0, -0.5, 1.0 / 3.0 - 1.0, 0.0, 2.0], n)
    out = np.empty(n)
    assert lib.wavi_math_probe(0, W._lib.dptr(x), W._lib.dptr(y), W._lib.dptr(out), n) == 0
    ref = np.array([oracle.det_pow(a, b) for a, b in zip(x, y)])
    assert np.array_equal(out, ref)
    assert lib.wavi_math_probe(1, W._lib.dptr(x), None, W._lib.dptr(out), n) == 0
    assert np.array_equal(out, np.sqrt(x))
    z = rng.standard_normal(n)
    assert lib.wavi_math_probe(2, W._lib.dptr(x), W._lib.dptr(z), W._lib.dptr(out), n) == 0
    assert np.array_equal(out, x / z)


@pytest.mark.parametrize("shape", [(37, 21), (48, 12), (64, 32)])
def test_haar_transforms(W, oracle, shape):
    nx, ny = shape
    _step("haar %s" % (shape,))
    m = channel(W, nx=nx, ny=ny)
    o = oracle.OracleModel(m)
    _gpu_model(W, m)
    rng = np.random.default_rng(5)
    for comp, sh in (("u", (nx + 1, ny)), ("v", (nx, ny + 1))):
        x = rng.standard_normal(sh)
        for d in ("forward", "inverse", "inverse_t"):
            assert np.array_equal(W.haar(m, comp, d, x), o.haar(comp, d, x)), (comp, d)


@pytest.mark.parametrize("shape", [(48, 12), (33, 17)])
def test_geometry_and_inner_update(W, oracle, shape):
    nx, ny = shape
    _step("inner update %s" % (shape,))
    m = channel(W, nx=nx, ny=ny)
    _rand_velocities(m, 7)
    o = oracle.OracleModel(m)
    spec = _gpu_model(W, m)
    lib = spec._lib
    assert lib.wavi_update_geometry(spec.ctx) == 0
    o.update_geometry()
    _assert_fields_equal(W, m, o)
    assert lib.wavi_inner_update(spec.ctx) == 0
    o.inner_update()
    _assert_fields_equal(W, m, o)


def test_operator_rhs_diag(W, oracle):
    _step("operator/rhs/diag")
    m, _, _ = iceberg(W)
    m.params.dt = 0.1
    _rand_velocities(m, 8, scale=5.0)
    o = oracle.OracleModel(m)
    spec = _gpu_model(W, m)
    lib = spec._lib
    assert lib.wavi_update_geometry(spec.ctx) == 0
    assert lib.wavi_inner_update(spec.ctx) == 0
    o.update_geometry()
    o.inner_update()
    rng = np.random.default_rng(9)
    nx, ny = m.grid.nx, m.grid.ny
    for sampled in (True, False):
        xu = rng.standard_normal((nx + 1, ny))
        xv = rng.standard_normal((nx, ny + 1))
        yu, yv = W.op_apply(m, xu, xv, sampled=sampled)
        if sampled:
            mk = o.masks()
            xu = np.where(mk["u_inner"], xu, 0.0)
            xv = np.where(mk["v_inner"], xv, 0.0)
        ou, ov = o.op_apply(xu, xv)
        assert np.array_equal(yu, ou) and np.array_equal(yv, ov)
    bu, bv = W.get_rhs(m)
    obu, obv = o.rhs()
    assert np.array_equal(bu, obu) and np.array_equal(bv, obv)
    du, dv = W.get_op_diag(m)
    odu, odv = o.op_diag()
    assert np.array_equal(du, odu) and np.array_equal(dv, odv)


def test_vcycle(W, oracle):
    _step("vcycle")
    m = channel(W, nx=40, ny=16)
    _rand_velocities(m, 10)
    o = oracle.OracleModel(m)
    spec = _gpu_model(W, m)
    lib = spec._lib
    assert lib.wavi_update_geometry(spec.ctx) == 0
    assert lib.wavi_inner_update(spec.ctx) == 0
    o.update_geometry()
    o.inner_update()
    bu, bv = o.rhs()
    for rep in range(2):  # the second V-cycle starts from the cached warm start
        eu, ev, it = W.vcycle(m, bu, bv)
        oeu, oev, oit = o.vcycle(bu, bv)
        assert it == oit
        assert np.array_equal(eu, oeu) and np.array_equal(ev, oev)
    _assert_fields_equal(W, m, o, names={"WU_CORRECTION_COARSE", "WV_CORRECTION_COARSE"})


@pytest.mark.parametrize("shape,picard", [((48, 12), 4), ((40, 24), 3), ((47, 15), 3)])
def test_update_velocities(W, oracle, shape, picard):
    nx, ny = shape
    _step("update_velocities %s" % (shape,))
    m = channel(W, nx=nx, ny=ny, maxiter_picard=picard)
    o = oracle.OracleModel(m)
    spec = _gpu_model(W, m)
    lib = spec._lib
    assert lib.wavi_update_geometry(spec.ctx) == 0
    o.update_geometry()
    rep = W._lib.SolveReport()
    assert lib.wavi_update_velocities(spec.ctx, ctypes.byref(rep)) == 0
    orep = o.update_velocities()
    assert rep.as_dict()["coarse_cg_iters"] == orep["coarse_cg_iters"]
    assert rep.as_dict()["rel_resid_hist"] == orep["rel_resid_hist"]
    _assert_fields_equal(W, m, o)


def test_ragged_domain_update_velocities(W, oracle):
    """An irregular domain (a calving front that moves from row to row,
    interior holes, interior Dirichlet points on u, grounded and floating
    parts), solved to convergence in 30 Picard iterations: bit-equal to the
    oracle, with the same coarse-CG counts."""
    _step("update_velocities ragged 40x24")
    m = ragged(W, maxiter_picard=30)
    o = oracle.OracleModel(m)
    spec = _gpu_model(W, m)
    lib = spec._lib
    assert lib.wavi_update_geometry(spec.ctx) == 0
    o.update_geometry()
    rep = W._lib.SolveReport()
    assert lib.wavi_update_velocities(spec.ctx, ctypes.byref(rep)) == 0
    orep = o.update_velocities()
    assert orep["converged"] and rep.as_dict() == orep
    _assert_fields_equal(W, m, o)


def test_ragged_domain_transient(W, oracle):
    """Ten coupled steps on the irregular domain (thickness, geometry,
    wavelet masks and solves in turn): bit-equal to the oracle."""
    _step("ragged transient 10 steps")
    m = ragged(W, maxiter_picard=3)
    m.params.dt = 0.5
    o = oracle.OracleModel(m)
    o.set_dt(0.5)
    sim = W.Simulation(model=m, timestepping_params=W.TimesteppingParams(dt=0.5, n_iter_total=10))
    W.run_simulation(sim)
    o.run_steps(10)
    _assert_fields_equal(W, m, o)


def test_iceberg_steps(W, oracle):
    _step("iceberg steps")
    m, _, _ = iceberg(W)
    m.params.dt = 0.1
    o = oracle.OracleModel(m)
    sim = W.Simulation(model=m, timestepping_params=W.TimesteppingParams(dt=0.1, n_iter_total=50))
    o.set_dt(0.1)
    W.run_simulation(sim)
    o.run_steps(50)
    _assert_fields_equal(W, m, o)


def test_mismip_plus_fixture_on_gpu(W, oracle):
    """The reference's regression run (1000 steps) on the GPU: bit-equal to
    the oracle and within 1e-6 of the reference's v1.8.3 golden."""
    _step("mismip+ 1000 steps")
    import os
    m = W.workloads.mismip_plus_fixture()
    sim = W.Simulation(model=m, timestepping_params=W.TimesteppingParams(dt=0.1, end_time=100.0))
    o = oracle.OracleModel(m)
    W.run_simulation(sim)
    o.run_steps(1000)
    _assert_fields_equal(W, m, o)
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "mismip_plus_v1_8_3.npz"))
    f = m.fields
    for name, arr in (("h", f.gh.h), ("u", f.gu.u), ("v", f.gv.v), ("viscosity", f.gh.ηav),
                      ("grounded_fraction", f.gh.grounded_fraction), ("bed_speed", f.gh.bed_speed)):
        assert rel_l2(arr, d[name]) < 1e-6, name


def test_channel_transient_200_steps(W, oracle):
    """80x10 channel, dt 0.5, 200 steps: thickness within 1e-9 of the oracle
    (the FP64 path is bit-equal, so this is the looser of the two bars)."""
    _step("channel transient 200 steps")
    m = W.workloads.channel_transient()
    sim = W.Simulation(model=m, timestepping_params=W.TimesteppingParams(dt=0.5, n_iter_total=200))
    o = oracle.OracleModel(m)
    W.run_simulation(sim)
    o.run_steps(200)
    assert rel_l2(m.fields.gh.h, o.field("H_H")) <= 1e-9
    _assert_fields_equal(W, m, o)


def test_transient_with_fixed_thickness(W, oracle):
    """update_thickness! with h_isfixed (run_simulation.jl:24-42): 30 steps
    of the channel with the upstream column and an interior patch fixed:
    bit-equal to the oracle, and the fixed cells keep their initial 100 m."""
    _step("channel transient with fixed thickness")
    m = W.workloads.channel_transient()
    m.grid.h_isfixed[0, :] = True
    m.grid.h_isfixed[30:34, 4:6] = True
    sim = W.Simulation(model=m, timestepping_params=W.TimesteppingParams(dt=0.5, n_iter_total=30))
    o = oracle.OracleModel(m)
    W.run_simulation(sim)
    o.run_steps(30)
    _assert_fields_equal(W, m, o)
    fixed = m.grid.h_isfixed & m.fields.gh.mask
    assert fixed.sum() == 18
    np.testing.assert_array_equal(m.fields.gh.h[fixed], 100.0)
    assert np.any(m.fields.gh.h[~fixed & m.fields.gh.mask] != 100.0)


def test_transient_other_implicitness(W, oracle):
    """super_implicitness 0.6 (the theta of the implicit thickness term in
    the operator and the right-hand side): 20 channel steps bit-equal to the
    oracle."""
    _step("channel transient, super_implicitness 0.6")
    m = W.workloads.channel_transient()
    m.solver_params.super_implicitness = 0.6
    sim = W.Simulation(model=m, timestepping_params=W.TimesteppingParams(dt=0.5, n_iter_total=20))
    o = oracle.OracleModel(m)
    W.run_simulation(sim)
    o.run_steps(20)
    _assert_fields_equal(W, m, o)


@pytest.mark.parametrize("levels", [0, 1, 2])
def test_update_velocities_other_haar_levels(W, oracle, levels):
    """Haar levels 0..2 (blocks of 2, 4, 8): bit-equal to the oracle."""
    _step("update_velocities levels %d" % levels)
    m = channel(W, nx=40, ny=24, maxiter_picard=2)
    m.solver_params.levels = levels
    o = oracle.OracleModel(m)
    spec = _gpu_model(W, m)
    lib = spec._lib
    assert lib.wavi_update_geometry(spec.ctx) == 0
    o.update_geometry()
    rep = W._lib.SolveReport()
    assert lib.wavi_update_velocities(spec.ctx, ctypes.byref(rep)) == 0
    orep = o.update_velocities()
    assert rep.as_dict()["coarse_cg_iters"] == orep["coarse_cg_iters"]
    _assert_fields_equal(W, m, o)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_state_that_meets_the_test_keeps_velocities_and_warm_start(W, oracle, precision):
    """update_velocities! on a state that already meets tol_picard returns at
    its first test, with no V-cycle (update_velocities.jl:20-24): the
    velocities and the coarse warm start (the one field a V-cycle writes)
    come out as before.  The fields the first test recomputes from the
    velocities (bed speed, drag, viscosity) follow the reference; in FP64
    every field equals the oracle's."""
    _step("converged state, %s" % precision)
    m = channel(W, nx=40, ny=24, maxiter_picard=60, precision=precision)
    W.update_geometry(m)
    W.update_velocities(m)
    assert m.last_report["converged"]
    before = {fid: np.array(m._array(c, a), dtype=float, copy=True) for c, a, fid in W.model.FIELD_MAP}
    m.parallel_spec.close()  # a new context, as for a model loaded from a checkpoint
    o = oracle.OracleModel(m) if precision == "fp64" else None
    W.update_velocities(m)
    rep = m.last_report
    assert rep["picard_iters"] == 1 and rep["converged"] and rep["coarse_cg_iters"][0] == 0
    for c, a, fid in W.model.FIELD_MAP:
        if fid in ("U_U", "V_V", "WU_CORRECTION_COARSE", "WV_CORRECTION_COARSE"):
            assert np.array_equal(np.asarray(m._array(c, a), dtype=float), before[fid]), fid
    if o is not None:
        orep = o.update_velocities()
        assert rep == orep
        _assert_fields_equal(W, m, o)


@pytest.mark.parametrize("tile", [(16, 16), (32, 16), (16, 32)])
def test_smoother_tile_shapes_bit_equal(W, oracle, tile):
    """Each tile shape of the smoother kernel (the autotuner's candidates)
    gives the oracle's bits: a 3-Picard solve on 70 x 40, whose tiles are
    partial in both directions for every shape."""
    _step("smoother tile %dx%d" % tile)
    m = channel(W, nx=70, ny=40, maxiter_picard=3)
    o = oracle.OracleModel(m)
    spec = _gpu_model(W, m)
    lib = spec._lib
    assert lib.wavi_set_smoother_tile(spec.ctx, *tile) == 0
    assert lib.wavi_update_geometry(spec.ctx) == 0
    o.update_geometry()
    rep = W._lib.SolveReport()
    assert lib.wavi_update_velocities(spec.ctx, ctypes.byref(rep)) == 0
    orep = o.update_velocities()
    assert rep.as_dict() == orep
    _assert_fields_equal(W, m, o)
    tx, ty = ctypes.c_int(), ctypes.c_int()
    assert lib.wavi_smoother_tile(spec.ctx, ctypes.byref(tx), ctypes.byref(ty)) == 0
    assert (tx.value, ty.value) == tile


@pytest.mark.parametrize("nx,ny", [(80, 10), (48, 12), (40, 24)])
def test_whole_grid_smoother_bit_equal_to_oracle(W, oracle, nx, ny):
    """The smoothing of a small grid in one workgroup (k_smooth_grid, every
    iteration in one launch; wavi_set_smoother_tile(-1, -1)) gives the
    oracle's bits: a 3-Picard FP64 solve on grids of one (48 x 12) and two
    (80 x 10, 40 x 24) window points per thread."""
    _step("whole-grid smoother %dx%d" % (nx, ny))
    m = channel(W, nx=nx, ny=ny, maxiter_picard=3)
    o = oracle.OracleModel(m)
    spec = _gpu_model(W, m)
    lib = spec._lib
    assert lib.wavi_set_smoother_tile(spec.ctx, -1, -1) == 0
    assert lib.wavi_update_geometry(spec.ctx) == 0
    o.update_geometry()
    rep = W._lib.SolveReport()
    assert lib.wavi_update_velocities(spec.ctx, ctypes.byref(rep)) == 0
    orep = o.update_velocities()
    assert rep.as_dict() == orep
    _assert_fields_equal(W, m, o)
    tx, ty = ctypes.c_int(), ctypes.c_int()
    assert lib.wavi_smoother_tile(spec.ctx, ctypes.byref(tx), ctypes.byref(ty)) == 0
    assert (tx.value, ty.value) == (-1, -1)


def test_whole_grid_smoother_ragged_domain_bit_equal_to_oracle(W, oracle):
    """The whole-grid smoother on the irregular domain (moving calving front,
    interior holes, interior Dirichlet points; 40 x 24): a 3-Picard FP64
    solve gives the oracle's bits."""
    _step("whole-grid smoother, ragged 40x24")
    m = ragged(W, nx=40, ny=24, maxiter_picard=3)
    o = oracle.OracleModel(m)
    spec = _gpu_model(W, m)
    lib = spec._lib
    assert lib.wavi_set_smoother_tile(spec.ctx, -1, -1) == 0
    assert lib.wavi_update_geometry(spec.ctx) == 0
    o.update_geometry()
    rep = W._lib.SolveReport()
    assert lib.wavi_update_velocities(spec.ctx, ctypes.byref(rep)) == 0
    orep = o.update_velocities()
    assert rep.as_dict() == orep
    _assert_fields_equal(W, m, o)


def test_whole_grid_smoother_mixed_and_autotuned(W):
    """Mixed precision on 80 x 10: the whole-grid smoother, a tile shape and
    the autotuner's choice (which may be either) give the same bits; a grid
    whose window does not fit one workgroup is refused."""
    runs = []
    for tile in ((-1, -1), (16, 16), (0, 0)):
        _step("mixed smoother %dx%d on 80x10" % tile)
        m = channel(W, nx=80, ny=10, precision="mixed", maxiter_picard=4)
        spec = _gpu_model(W, m)
        lib = spec._lib
        assert lib.wavi_set_smoother_tile(spec.ctx, *tile) == 0
        assert lib.wavi_update_geometry(spec.ctx) == 0
        rep = W._lib.SolveReport()
        assert lib.wavi_update_velocities(spec.ctx, ctypes.byref(rep)) == 0
        W.pull_fields(m)
        tx, ty = ctypes.c_int(), ctypes.c_int()
        assert lib.wavi_smoother_tile(spec.ctx, ctypes.byref(tx), ctypes.byref(ty)) == 0
        if tile == (0, 0):
            assert (tx.value, ty.value) in ((-1, -1), (16, 16), (32, 16), (16, 32))
            _step("autotuned smoother %dx%d" % (tx.value, ty.value))
        else:
            assert (tx.value, ty.value) == tile
        runs.append((m.fields.gu.u.copy(), m.fields.gv.v.copy(), rep.as_dict()))
    for u, v, r in runs[1:]:
        assert np.array_equal(u, runs[0][0]) and np.array_equal(v, runs[0][1]) and r == runs[0][2]
    big = channel(W, nx=70, ny=40, precision="mixed")
    spec = _gpu_model(W, big)
    assert spec._lib.wavi_set_smoother_tile(spec.ctx, -1, -1) == W._lib.WAVI_EARG


def test_smoother_autotuner_and_shapes_in_mixed_precision(W):
    """The mixed solve gives the same bits whatever the smoother tile shape;
    the autotuner picks one of the shapes; an unknown shape is rejected."""
    _step("smoother autotuner, mixed")
    runs = []
    for tile in ((0, 0), (16, 16), (32, 16), (16, 32)):
        m = channel(W, nx=96, ny=40, precision="mixed", maxiter_picard=3)
        spec = _gpu_model(W, m)
        lib = spec._lib
        assert lib.wavi_set_smoother_tile(spec.ctx, *tile) == 0
        assert lib.wavi_update_geometry(spec.ctx) == 0
        rep = W._lib.SolveReport()
        assert lib.wavi_update_velocities(spec.ctx, ctypes.byref(rep)) == 0
        W.pull_fields(m)
        tx, ty = ctypes.c_int(), ctypes.c_int()
        assert lib.wavi_smoother_tile(spec.ctx, ctypes.byref(tx), ctypes.byref(ty)) == 0
        if tile == (0, 0):
            assert (tx.value, ty.value) in ((16, 16), (32, 16), (16, 32))
            _step("autotuned smoother tile %dx%d" % (tx.value, ty.value))
        else:
            assert (tx.value, ty.value) == tile
        runs.append((m.fields.gu.u.copy(), m.fields.gv.v.copy(), rep.as_dict()))
    for u, v, r in runs[1:]:
        assert np.array_equal(u, runs[0][0]) and np.array_equal(v, runs[0][1]) and r == runs[0][2]
    assert lib.wavi_set_smoother_tile(spec.ctx, 8, 8) == W._lib.WAVI_EARG
