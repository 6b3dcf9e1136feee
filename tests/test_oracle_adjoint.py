"""The oracle's adjoint basal-drag gradient pinned by the mathematics of
the map it differentiates: central finite differences along seeded
directions, and the second-order decay of the Taylor remainder
|J(l + h d) - J(l) - h g.d| (a wrong or missing term in the reverse sweep
leaves a first-order remainder).  The GPU gradient is compared with this
oracle in tests/test_gpu_adjoint.py."""
import numpy as np
import pytest

NX, NY, N_PICARD, TOL = 48, 12, 3, 1e-13


@pytest.fixture(scope="module")
def inversion(W, oracle):
    lt = W.workloads.inversion_logc_true(NX, NY, seed=4)

    def make(logc):
        o = oracle.OracleModel(W.workloads.inversion_model(logc, "fp64", nσ=4, n_picard=N_PICARD))
        o.update_geometry()
        return o

    o = make(lt)
    z = (np.zeros((NX + 1) * NY), np.zeros(NX * (NY + 1)))
    o.misfit_gradient(N_PICARD, TOL, 60, *z, gradient=False)
    uo, vo = W.workloads.inversion_observations(o.field("U_U"), o.field("V_V"), seed=4)
    l0 = np.log(1e4) * np.ones((NX, NY))
    J0, g, rep = make(l0).misfit_gradient(N_PICARD, TOL, 60, uo, vo)
    J = lambda logc: make(logc).misfit_gradient(N_PICARD, TOL, 60, uo, vo, gradient=False)[0]
    return l0, J0, g, rep, J


def test_forward_only_matches(inversion):
    l0, J0, g, rep, J = inversion
    assert rep["unconverged"] == 0 and rep["max_fwd_rel"] <= TOL and rep["max_adj_rel"] <= TOL
    assert J(l0) == J0
    assert np.all(np.isfinite(g)) and np.linalg.norm(g) > 0


def test_central_differences(inversion):
    l0, J0, g, rep, J = inversion
    rng = np.random.default_rng(3)
    eps = 1e-6
    for _ in range(2):
        d = rng.standard_normal(l0.shape)
        fd = (J(l0 + eps * d) - J(l0 - eps * d)) / (2 * eps)
        gd = float(np.sum(g * d))
        assert abs(fd - gd) <= 1e-6 * abs(gd), (fd, gd)


def test_taylor_remainder_is_second_order(inversion):
    l0, J0, g, rep, J = inversion
    d = np.random.default_rng(5).standard_normal(l0.shape)
    gd = float(np.sum(g * d))
    r = [abs(J(l0 + h * d) - J0 - h * gd) for h in (1e-2, 5e-3)]
    # halving h divides a second-order remainder by about 4
    assert 3.0 < r[0] / r[1] < 5.0, r


def test_central_differences_on_a_ragged_domain(W, oracle):
    """The reverse sweep on an irregular domain (a calving front that moves
    from row to row, interior holes, grounded and floating parts): central
    differences along seeded directions."""
    from helpers import ragged
    nx, ny = 24, 12

    def make(logc):
        m = ragged(W, nx=nx, ny=ny, maxiter_picard=N_PICARD, dirichlet=False,
                   weertman_c=np.asfortranarray(np.exp(logc)))
        o = oracle.OracleModel(m)
        o.update_geometry()
        return o

    l0 = np.log(1e4) * np.ones((nx, ny))
    z = (np.zeros((nx + 1) * ny), np.zeros(nx * (ny + 1)))
    o = make(l0 + 0.3 * np.random.default_rng(8).standard_normal((nx, ny)))
    o.misfit_gradient(N_PICARD, TOL, 60, *z, gradient=False)
    uo, vo = W.workloads.inversion_observations(o.field("U_U"), o.field("V_V"), seed=8)
    J0, g, rep = make(l0).misfit_gradient(N_PICARD, TOL, 60, uo, vo)
    assert rep["unconverged"] == 0
    J = lambda logc: make(logc).misfit_gradient(N_PICARD, TOL, 60, uo, vo, gradient=False)[0]  # noqa: E731
    rng = np.random.default_rng(9)
    eps = 1e-6
    for _ in range(2):
        d = rng.standard_normal(l0.shape)
        fd = (J(l0 + eps * d) - J(l0 - eps * d)) / (2 * eps)
        gd = float(np.sum(g * d))
        assert abs(fd - gd) <= 1e-6 * abs(gd), (fd, gd)
