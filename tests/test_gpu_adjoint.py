"""Adjoint basal-drag gradient on the GPU (wavi_misfit_gradient): the
gradient against the oracle's reverse sweep, and against central finite
differences of the map itself in both precisions."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_PICARD = 3
TOL_LIN = 1e-13


def _step(name):
    print("[gpu] %s" % name, flush=True)
    sys.stdout.flush()


def _model(W, logc, precision, nsig=4):
    m = W.workloads.inversion_model(logc, precision, nσ=nsig, n_picard=N_PICARD)
    W.update_geometry(m)
    return m


def _setup(W, nx, ny):
    lt = W.workloads.inversion_logc_true(nx, ny, seed=4)
    m = _model(W, lt, "fp64")
    zu, zv = np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1))
    W.misfit_gradient(m, N_PICARD, zu, zv, TOL_LIN, gradient=False)
    uo, vo = W.workloads.inversion_observations(m.fields.gu.u, m.fields.gv.v, seed=4)
    return np.log(1e4) * np.ones((nx, ny)), uo, vo


def test_gradient_matches_oracle(W, oracle):
    nx, ny = 64, 16
    _step("adjoint vs oracle %dx%d" % (nx, ny))
    l0, uo, vo = _setup(W, nx, ny)
    m = _model(W, l0, "fp64")
    J, g, rep = W.misfit_gradient(m, N_PICARD, uo, vo, TOL_LIN)
    assert rep["unconverged"] == 0
    o = oracle.OracleModel(W.workloads.inversion_model(l0, "fp64", nσ=4, n_picard=N_PICARD))
    o.update_geometry()
    Jo, go, repo = o.misfit_gradient(N_PICARD, TOL_LIN, 60, uo, vo)
    assert repo["unconverged"] == 0
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    assert np.linalg.norm(g - go) <= 1e-8 * np.linalg.norm(go)
    # the context holds the forward's last state: the same velocities as the oracle's
    uf = o.field("U_U")
    assert np.max(np.abs(m.fields.gu.u - uf)) <= 1e-9 * np.max(np.abs(uf))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gradient_matches_finite_differences(W, precision):
    nx, ny = 96, 24
    _step("adjoint FD check %s %dx%d" % (precision, nx, ny))
    l0, uo, vo = _setup(W, nx, ny)
    J0, g, rep = W.misfit_gradient(_model(W, l0, precision), N_PICARD, uo, vo, TOL_LIN)
    assert rep["unconverged"] == 0 and rep["max_fwd_rel"] <= TOL_LIN and rep["max_adj_rel"] <= TOL_LIN
    rng = np.random.default_rng(11)
    eps = 1e-6
    for trial in range(3):
        d = rng.standard_normal((nx, ny))
        Jp = W.misfit_gradient(_model(W, l0 + eps * d, precision), N_PICARD, uo, vo, TOL_LIN, gradient=False)[0]
        Jm = W.misfit_gradient(_model(W, l0 - eps * d, precision), N_PICARD, uo, vo, TOL_LIN, gradient=False)[0]
        fd = (Jp - Jm) / (2 * eps)
        gd = float(np.sum(g * d))
        print("  direction %d: fd %.12e adjoint %.12e rel %.2e" % (trial, fd, gd, abs(fd - gd) / abs(gd)))
        assert abs(fd - gd) <= 1e-5 * abs(gd)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gradient_on_a_ragged_domain(W, oracle, precision):
    """An irregular domain (a calving front that moves from row to row,
    interior holes, grounded and floating parts): the gradient against the
    oracle's reverse sweep."""
    from helpers import ragged
    nx, ny = 40, 24
    _step("adjoint on a ragged domain %s %dx%d" % (precision, nx, ny))

    def make(logc, prec):
        return ragged(W, nx=nx, ny=ny, precision=prec, maxiter_picard=N_PICARD, dirichlet=False,
                      weertman_c=np.asfortranarray(np.exp(logc)))

    l0 = np.log(1e4) * np.ones((nx, ny))
    lt = l0 + 0.3 * np.random.default_rng(8).standard_normal((nx, ny))
    ot = oracle.OracleModel(make(lt, "fp64"))
    ot.update_geometry()
    ot.misfit_gradient(N_PICARD, TOL_LIN, 60, np.zeros((nx + 1) * ny), np.zeros(nx * (ny + 1)), gradient=False)
    uo, vo = W.workloads.inversion_observations(ot.field("U_U").reshape((nx + 1, ny), order="F"),
                                                ot.field("V_V").reshape((nx, ny + 1), order="F"), seed=8)
    o = oracle.OracleModel(make(l0, "fp64"))
    o.update_geometry()
    Jo, go, repo = o.misfit_gradient(N_PICARD, TOL_LIN, 60, uo, vo)
    assert repo["unconverged"] == 0
    m = make(l0, precision)
    W.update_geometry(m)
    J, g, rep = W.misfit_gradient(m, N_PICARD, uo, vo, TOL_LIN)
    err = np.linalg.norm(g - go) / np.linalg.norm(go)
    print("  J rel %.2e, gradient rel %.2e" % (abs(J - Jo) / abs(Jo), err), flush=True)
    assert rep["unconverged"] == 0
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    assert err <= (1e-8 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("levels", [1, 3])
def test_coarse_diagonal_matches_the_oracle_operator(W, oracle, levels):
    """wavi_coarse_diag, the Jacobi scaling of the coarse CG in the
    adjoint's linear solves, against the diagonal of the oracle's coarse
    operator R A P (update_preconditioners.jl:73) read off column by column:
    a ragged domain of 37 x 22 (partial tiles and Haar blocks, holes, a
    moving front), Haar levels 1 and 3."""
    from helpers import ragged
    nx, ny = 37, 22
    _step("coarse diagonal vs oracle, levels %d" % levels)
    m = ragged(W, nx=nx, ny=ny, maxiter_picard=1, dirichlet=False)
    m.solver_params.levels = levels
    W.update_geometry(m)
    spec = m.parallel_spec
    lib = spec._lib
    o = oracle.OracleModel(m)
    assert lib.wavi_inner_update(spec.ctx) == 0
    o.inner_update()
    nu, nv = (nx + 1) * ny, nx * (ny + 1)
    du, dv = np.zeros(nu), np.zeros(nv)
    assert lib.wavi_coarse_diag(spec.ctx, W._lib.dptr(du), W._lib.dptr(dv)) == 0
    ref = np.zeros(nu + nv)
    for k in range(nu + nv):
        e = np.zeros(nu + nv)
        e[k] = 1.0
        au, av = o.coarse_op(e[:nu].reshape((nx + 1, ny), order="F"), e[nu:].reshape((nx, ny + 1), order="F"))
        ref[k] = au.ravel(order="F")[k] if k < nu else av.ravel(order="F")[k - nu]
    got = np.concatenate([du, dv])
    scale = np.abs(ref).max()
    assert scale > 0 and np.count_nonzero(ref) > (nu + nv) // 2
    assert np.array_equal(got == 0.0, ref == 0.0)
    err = np.abs(got - ref)
    assert np.all(err <= 1e-10 * np.abs(ref) + 1e-13 * scale), err.max() / scale


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_jacobi_coarse_cg_gives_the_same_gradient_with_fewer_coarse_iterations(W, precision):
    """The adjoint's linear solves with the Jacobi-scaled coarse CG (the
    default) and with the V-cycle of update_velocities!: the same J and
    gradient to the solve tolerance, in fewer V-cycles."""
    nx, ny = 96, 24
    _step("Jacobi coarse CG in the adjoint solves %s %dx%d" % (precision, nx, ny))
    l0, uo, vo = _setup(W, nx, ny)
    res = {}
    for on in (0, 1):
        m = _model(W, l0, precision)
        assert m.parallel_spec._lib.wavi_set_linear_coarse_jacobi(m.parallel_spec.ctx, on) == 0
        res[on] = W.misfit_gradient(m, N_PICARD, uo, vo, TOL_LIN)
    (J0, g0, r0), (J1, g1, r1) = res[0], res[1]
    v0, v1 = r0["fwd_vcycles"] + r0["adj_vcycles"], r1["fwd_vcycles"] + r1["adj_vcycles"]
    c0, c1 = r0["coarse_cg_iters"], r1["coarse_cg_iters"]
    err = np.linalg.norm(g1 - g0) / np.linalg.norm(g0)
    print("  V-cycles %d plain, %d Jacobi; coarse-CG iterations %d plain, %d Jacobi; J rel %.2e, gradient rel %.2e"
          % (v0, v1, c0, c1, abs(J1 - J0) / abs(J0), err))
    assert r0["unconverged"] == 0 and r1["unconverged"] == 0
    assert abs(J1 - J0) <= 1e-11 * abs(J0)
    assert err <= 1e-9
    assert v1 <= v0 and c1 < c0
    if precision == "fp64":
        assert 2 * c1 < c0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_jacobi_coarse_cg_other_haar_level_and_one_workgroup(W, precision):
    """The Jacobi-scaled coarse CG at Haar level 1 (blocks of 4): the same
    gradient as the plain V-cycle to the solve tolerance; and the coarse CG
    run whole in one workgroup (wavi_set_coarse_cg) gives the launched
    Jacobi CG's gradient bit for bit."""
    nx, ny = 48, 12
    _step("Jacobi coarse CG, levels 1 and one workgroup, %s %dx%d" % (precision, nx, ny))
    l0, uo, vo = _setup(W, nx, ny)
    out = {}
    for key, jac, mode in (("plain", 0, W._lib.WAVI_CG_LAUNCHES), ("jacobi", 1, W._lib.WAVI_CG_LAUNCHES),
                           ("jacobi_wg", 1, W._lib.WAVI_CG_WORKGROUP)):
        m = W.workloads.inversion_model(l0, precision, nσ=4, n_picard=N_PICARD)
        m.solver_params.levels = 1
        W.update_geometry(m)
        spec = m.parallel_spec
        assert spec._lib.wavi_set_linear_coarse_jacobi(spec.ctx, jac) == 0
        assert spec._lib.wavi_set_coarse_cg(spec.ctx, mode) == 0
        out[key] = W.misfit_gradient(m, N_PICARD, uo, vo, TOL_LIN)
    (Jp, gp, rp), (Jj, gj, rj), (Jw, gw, rw) = out["plain"], out["jacobi"], out["jacobi_wg"]
    assert rp["unconverged"] == 0 and rj["unconverged"] == 0
    assert abs(Jj - Jp) <= 1e-11 * abs(Jp)
    assert np.linalg.norm(gj - gp) <= 1e-9 * np.linalg.norm(gp)
    assert rj["coarse_cg_iters"] < rp["coarse_cg_iters"]
    assert Jw == Jj and np.array_equal(gw, gj) and rw == rj


def test_fixed_velocities_rejected(W):
    _step("adjoint rejects fixed velocities")
    nx, ny = 32, 8
    m = W.workloads.channel_solve(nx=nx, ny=ny, nσ=4, maxiter_picard=2)
    m.grid.u_isfixed[3, 3] = True
    W.update_geometry(m)
    with pytest.raises(W.ArgumentError):
        W.misfit_gradient(m, 2, np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1)))


def test_torch_autograd_binding(W):
    """MisfitLogC.apply(logc, problem).backward() gives the adjoint gradient,
    and the problem re-evaluates from its fixed starting state."""
    import importlib
    torch = pytest.importorskip("torch")
    autodiff = importlib.import_module("wavi_jl_amd.autodiff")
    nx, ny = 64, 16
    _step("adjoint through torch.autograd %dx%d" % (nx, ny))
    l0, uo, vo = _setup(W, nx, ny)
    problem = autodiff.MisfitProblem(_model(W, l0, "mixed"), N_PICARD, uo, vo, TOL_LIN)
    logc = torch.tensor(l0, dtype=torch.float64, requires_grad=True)
    J = autodiff.MisfitLogC.apply(logc, problem)
    J.backward()
    Jr, gr, _ = W.misfit_gradient(_model(W, l0, "mixed"), N_PICARD, uo, vo, TOL_LIN)
    assert float(J.detach()) == Jr
    np.testing.assert_array_equal(logc.grad.numpy(), gr)
    J2 = autodiff.MisfitLogC.apply(logc.detach(), problem)
    assert float(J2) == Jr
