"""Host-side API (the mirror of WAVI.jl's constructors) and the C ABI of
libwavi_hip.so, without a GPU.  Error conventions follow the reference's
tests (test/test_grids.jl, test_models.jl, test_timesteppingparams.jl)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_grid_defaults_and_orientations(W):
    g = W.Grid(nx=10, ny=5)
    assert g.h_mask.all() and not g.u_iszero.any() and not g.v_iszero.any()
    assert np.allclose(g.quadrature_weights, np.r_[0.5, np.ones(g.nσ - 2), 0.5] / (g.nσ - 1))
    for s, sl in (("north", np.s_[0, :]), ("south", np.s_[-1, :]), ("west", np.s_[:, 0]), ("east", np.s_[:, -1])):
        g = W.Grid(nx=10, ny=5, u_iszero=[s], v_iszero=[s])
        assert g.u_iszero[sl].all() and g.v_iszero[sl].all()
    assert g.xxu[0, 0] == 0.0 and g.xxh[0, 0] == 4000.0 and g.yyv[0, 0] == -40000.0


def test_grid_errors(W):
    with pytest.raises(W.DimensionMismatch):
        W.Grid(nx=10, ny=10, h_mask=np.ones((5, 5), bool))
    with pytest.raises(W.DimensionMismatch):
        W.Grid(nσ=4, quadrature_weights=np.r_[0.5, np.ones(3), 0.5] / 3)
    with pytest.raises(W.DimensionMismatch):
        W.Grid(nx=10, ny=10, u_iszero=np.ones((10, 10)))
    with pytest.raises(W.DimensionMismatch):
        W.Grid(nx=10, ny=10, v_iszero=np.ones((10, 10)))
    with pytest.raises(W.ArgumentError):
        W.Grid(nx=10, ny=10, h_mask=2.0 * np.ones((10, 10)))
    with pytest.raises(W.ArgumentError):
        W.Grid(nx=1, ny=10)
    with pytest.raises(W.DimensionMismatch):
        W.Grid(nσ=7, σ=np.linspace(0, 1, 6))


def test_timestepping_params(W):
    assert W.TimesteppingParams(dt=0.5, end_time=100.0).n_iter_total == 200
    assert W.TimesteppingParams(dt=0.5, n_iter_total=100).end_time == 50
    for kw in (dict(dt=-0.5, end_time=1.0), dict(end_time=-1.0), dict(n_iter_total=-1),
               dict(end_time=1.0, dt=1.0, n_iter_total=10), dict(dt=0.5), dict(dt=0.5, end_time=np.inf)):
        with pytest.raises(W.ArgumentError):
            W.TimesteppingParams(**kw)


def test_model_construction_and_errors(W):
    grid = W.Grid(nx=12, ny=6)
    bed = -500.0 * np.ones((12, 6))
    m = W.Model(grid=grid, bed_elevation=bed, params=W.Params(weertman_c=1.0, accumulation_rate=1.0, glen_a_ref=1.0))
    assert np.all(m.params.weertman_c == 1.0) and m.params.weertman_c.shape == (12, 6)
    p = m.params
    ic = m.initial_conditions
    assert np.all(ic.initial_thickness == p.default_thickness)
    assert np.all(ic.initial_viscosity == p.default_viscosity) and ic.initial_viscosity.shape == (12, 6, 4)
    assert np.allclose(m.fields.gh.quad_f2, m.fields.gh.h / (3 * m.fields.gh.ηav))
    assert m.fields.wu.mask.all() and m.fields.gu.mask_inner.sum() == m.fields.gu.ni
    ic = W.InitialConditions(initial_thickness=3.14159 * np.ones((12, 6)),
                             initial_viscosity=2.7182818 * np.ones((12, 6, 4)))
    m = W.Model(grid=grid, bed_elevation=bed, initial_conditions=ic)
    assert np.all(m.fields.gh.h == 3.14159) and np.all(m.fields.g3d.η == 2.7182818)
    with pytest.raises(W.ArgumentError):
        W.Model(bed_elevation=bed)
    with pytest.raises(W.ArgumentError):
        W.Model(grid=grid)
    for bad in (dict(initial_thickness=np.ones((11, 6))), dict(initial_u_veloc=np.ones((12, 6))),
                dict(initial_v_veloc=np.ones((12, 6))), dict(initial_viscosity=np.ones((12, 6, 3)))):
        with pytest.raises(W.DimensionMismatch):
            W.Model(grid=grid, bed_elevation=bed, initial_conditions=W.InitialConditions(**bad))
    with pytest.raises(W.DimensionMismatch):
        W.Model(grid=grid, bed_elevation=bed, params=W.Params(weertman_c=np.ones((11, 5))))
    # bed as a function of (x, y) (Model.jl:64-67)
    m = W.Model(grid=grid, bed_elevation=W.mismip_plus_bed)
    assert m.fields.gh.b[0, 0] == pytest.approx(W.mismip_plus_bed(grid.xxh[0, 0], grid.yyh[0, 0]))


def test_masks_follow_setup_fields(W):
    h = np.ones((6, 5), bool)
    h[0, 0] = False
    h[3, 2] = False
    u, v, c = W.get_u_mask(h), W.get_v_mask(h), W.get_c_mask(h)
    assert u.shape == (7, 5) and v.shape == (6, 6) and c.shape == (5, 4)
    assert not u[0, 0] and u[1, 0] and u[3, 2] and u[4, 2]
    assert not c[0, 0] and not c[2, 1] and not c[3, 2] and c[4, 3]
    g = W.Grid(nx=6, ny=5, h_mask=h, u_iszero=["north"])
    m = W.Model(grid=g, bed_elevation=np.zeros((6, 5)))
    assert not m.fields.gu.mask[0, :].any()


def _header_functions():
    text = open(os.path.join(ROOT, "include", "wavi_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(wavi_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_header_symbol(W):
    lib_path = W._lib.LIB_PATH
    if not os.path.exists(lib_path):
        pytest.skip("libwavi_hip.so not built")
    names = _header_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(lib_path)
    for n in names:
        assert hasattr(lib, n), n
    # the Python binding declares a signature for each of them
    assert set(names) == set(W._lib.SIGNATURES), set(names) ^ set(W._lib.SIGNATURES)


def test_field_enum_matches_header(W):
    text = open(os.path.join(ROOT, "include", "wavi_hip.h")).read()
    body = text[text.index("enum wavi_field {"):text.index("WAVI_NFIELDS")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"WAVI_(\w+)", body)
    assert names == W._lib.FIELD_NAMES


def test_create_errors_without_gpu(W):
    """Argument checks run before any device call, so they work here too."""
    lib_path = W._lib.LIB_PATH
    if not os.path.exists(lib_path):
        pytest.skip("libwavi_hip.so not built")
    lib = W._lib.load()
    m = W.workloads.channel_solve(nx=16, ny=8, nσ=4, maxiter_picard=1)
    keep, desc, phys, solver = W.model.grid_desc(m)
    out = ctypes.c_void_p()
    solver.n_iter_viscosity = 3
    assert lib.wavi_ctx_create(ctypes.byref(desc), ctypes.byref(phys), ctypes.byref(solver), 0,
                               ctypes.byref(out)) == W._lib.WAVI_EARG
    assert b"n_iter_viscosity" in lib.wavi_last_error(None)
    solver.n_iter_viscosity = 2
    solver.levels = 4
    assert lib.wavi_ctx_create(ctypes.byref(desc), ctypes.byref(phys), ctypes.byref(solver), 0,
                               ctypes.byref(out)) == W._lib.WAVI_EARG
    solver.levels = 3
    solver.maxiter_coarse = (1 << 30) + 1
    assert lib.wavi_ctx_create(ctypes.byref(desc), ctypes.byref(phys), ctypes.byref(solver), 0,
                               ctypes.byref(out)) == W._lib.WAVI_EARG
    assert b"maxiter_coarse" in lib.wavi_last_error(None)
    solver.maxiter_coarse = 1000
    assert lib.wavi_ctx_create(ctypes.byref(desc), ctypes.byref(phys), ctypes.byref(solver), 7,
                               ctypes.byref(out)) == W._lib.WAVI_EARG
    assert lib.wavi_ctx_create(None, ctypes.byref(phys), ctypes.byref(solver), 0, ctypes.byref(out)) == W._lib.WAVI_EARG
    # 16400 x 16400 padded to whole tiles exceeds 2^28 points per array; the
    # check runs before any array is read
    desc.nx, desc.ny = 16400, 16400
    assert lib.wavi_ctx_create(ctypes.byref(desc), ctypes.byref(phys), ctypes.byref(solver), 0,
                               ctypes.byref(out)) == W._lib.WAVI_EDIM
    assert b"grid too large" in lib.wavi_last_error(None)
    with pytest.raises(W.ArgumentError):
        W.HIPSpec("bf16")


def test_bench_rejects_multi_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stdout


def test_batch_and_gradient_argument_checks_without_gpu(W):
    """The batch and adjoint entry points reject invalid arguments before any
    device call."""
    if not os.path.exists(W._lib.LIB_PATH):
        pytest.skip("libwavi_hip.so not built")
    lib = W._lib.load()
    EARG = W._lib.WAVI_EARG
    null_ctxs = (ctypes.c_void_p * 2)(None, None)
    assert lib.wavi_update_velocities_batch(None, -1, 1, None, None) == EARG
    assert lib.wavi_update_velocities_batch(null_ctxs, 2, 0, None, None) == EARG
    assert lib.wavi_update_velocities_batch(null_ctxs, 2, 4, None, None) == EARG  # NULL members
    assert lib.wavi_update_velocities_batch(None, 0, 4, None, None) == W._lib.WAVI_OK  # empty batch
    assert lib.wavi_update_state_batch(null_ctxs, 2, 4, None, None) == EARG
    assert lib.wavi_misfit_gradient_batch(null_ctxs, 2, 4, 3, 1e-13, 60, None, None, None, None, None, None) == EARG
    J = ctypes.c_double()
    assert lib.wavi_misfit_gradient(None, 3, 1e-13, 60, None, None, ctypes.byref(J), None, None) == EARG
    assert lib.wavi_after_solve(None) == EARG


def test_params_reject_unknown_fields(W):
    """As Julia structs do: a misspelt or misplaced field is an error, not a
    silently ignored new attribute (super_implicitness is a solver
    parameter, src/SolverParams.jl:2-14)."""
    p, sp = W.Params(), W.SolverParams()
    with pytest.raises(AttributeError):
        p.super_implicitness = 0.6
    with pytest.raises(AttributeError):
        sp.tol_coarses = 1e-6
    sp.super_implicitness = 0.6
    p.dt = 0.5
    assert sp.super_implicitness == 0.6 and p.dt == 0.5
