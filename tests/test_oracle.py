"""Pins of the CPU oracle (oracle/wavi_oracle.c) before it is trusted as the
checker of the GPU library: the reference's golden vectors, its known-answer
unit tests, closed forms and invariants that a dropped term, a wrong sign or
a transposed operand would break."""
import math

import numpy as np
import pytest

from helpers import channel, iceberg, rel_l2

GOLDEN = {
    "h": "H_H", "u": "U_U", "v": "V_V", "viscosity": "H_ETA_AV",
    "grounded_fraction": "H_GROUNDED_FRACTION", "bed_speed": "H_BED_SPEED",
}


def _fixture(tag):
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    return np.load(os.path.join(here, "golden", "mismip_plus_%s.npz" % tag))


@pytest.fixture(scope="module")
def mismip_run(W, oracle):
    """MISMIP+ 80x10, dt 0.1, 1000 steps, maxiter_picard 1: the run behind
    test/version_update_test_verification/test_version_updates.jl:6-70."""
    m = W.workloads.mismip_plus_fixture()
    m.params.dt = 0.1
    o = oracle.OracleModel(m)
    o.run_steps(1000)
    return {k: o.field(f) for k, f in GOLDEN.items()}


def test_mismip_plus_matches_reference_golden_v1_8_3(mismip_run):
    # The reference compares with isapprox (rtol sqrt(eps)); the restated
    # oracle's rounding differs from Julia's, and 1000 coupled steps amplify
    # it, so the bound here is 1e-6 (observed: at most 3e-8).
    d = _fixture("v1_8_3")
    for k in GOLDEN:
        assert rel_l2(mismip_run[k], d[k]) < 1e-6, k


def test_older_golden_family_is_a_different_code_revision(mismip_run):
    # The 1.5.1/1.6.x fixtures agree with one another to 1e-6 but differ from
    # the 1.8.3 one by up to 0.24 (viscosity): they were made by an older
    # revision of src/.  The oracle follows the current source.
    for tag in ("v1_5_1", "v1_6_1", "v1_6_2"):
        d = _fixture(tag)
        assert rel_l2(mismip_run["viscosity"], d["viscosity"]) > 1e-2


def test_det_pow_within_one_ulp_of_libm(oracle):
    rng = np.random.default_rng(0)
    xs = np.exp(rng.uniform(-40, 40, 20000))
    ys = rng.choice([-2.0 / 3.0, 1.0 / 3.0 - 1.0, -0.5, 0.5, 2.0, -1.7, 0.0], 20000)
    worst = 0.0
    for x, y in zip(xs, ys):
        a, b = oracle.det_pow(x, y), math.pow(x, y)
        worst = max(worst, abs(a - b) / math.ulp(b))
    assert worst <= 1.0
    assert oracle.det_pow(8.0, 1.0 / 3.0) == pytest.approx(2.0, rel=1e-15)
    assert oracle.det_pow(0.0, -1.0) == math.inf
    assert oracle.det_pow(5.0, 0.0) == 1.0


@pytest.mark.parametrize("levels", [0, 1, 2, 3])
def test_wavelet_matrix_inverse_pair_and_block_structure(oracle, levels):
    B = 2 ** (levels + 1)
    for n in list(range(1, 2 * B + 3)) + [37, 81]:
        F = oracle.wavelet_matrix(n, levels, True)
        R = oracle.wavelet_matrix(n, levels, False)
        assert np.array_equal(F @ R, np.eye(n))
        # block diagonal with aligned blocks of B: what the GPU tiles rely on
        for b0 in range(0, n, B):
            bl = min(B, n - b0)
            assert np.array_equal(F[b0:b0 + bl, b0:b0 + bl], oracle.wavelet_matrix(bl, levels, True))
            mask = np.ones((n, n), bool)
            mask[b0:b0 + bl, b0:b0 + bl] = False
            assert not np.any(F[b0:b0 + bl][:, mask[b0:b0 + bl].any(axis=0)])


def test_wavelet_known_answers(oracle):
    # one lifting level on a pair: a = (x0+x1)/2 in the odd slot, d = (x1-x0)/2
    F = oracle.wavelet_matrix(2, 0, True)
    assert np.array_equal(F @ np.array([3.0, 7.0]), np.array([5.0, 2.0]))
    # a constant lands only in the scaling slot of each 16-block (levels=3)
    F = oracle.wavelet_matrix(16, 3, True)
    y = F @ np.full(16, 4.0)
    assert y[0] == 4.0 and np.all(y[1:] == 0.0)
    # inverse entries are +-1
    R = oracle.wavelet_matrix(16, 3, False)
    assert set(np.unique(R)) <= {-1.0, 0.0, 1.0}


def test_haar_2d_adjoint_and_inverse(W, oracle):
    m = channel(W, nx=37, ny=21)
    o = oracle.OracleModel(m)
    rng = np.random.default_rng(1)
    for comp, shape in (("u", (38, 21)), ("v", (37, 22))):
        x = rng.standard_normal(shape)
        y = rng.standard_normal(shape)
        Px = o.haar(comp, "inverse", x)
        Rty = o.haar(comp, "inverse_t", y)
        assert abs(np.sum(Px * y) - np.sum(x * Rty)) < 1e-12 * np.abs(Px * y).sum()
        assert np.allclose(o.haar(comp, "forward", Px), x, rtol=0, atol=1e-13)


def test_pos_fraction_reference_cases(oracle):
    # test/test_utils.jl:4-31
    z = np.array([[-1.0, -1, -1], [-1, 1, -1], [-1, -1, -1]])
    h, u, v = oracle.pos_fraction(z)
    assert np.array_equal(h, [[0, 0, 0], [0, 0.5, 0], [0, 0, 0]])
    assert np.array_equal(u, [[0, 0, 0], [0, 0.25, 0], [0, 0.25, 0], [0, 0, 0]])
    assert np.array_equal(v, [[0, 0, 0, 0], [0, 0.25, 0.25, 0], [0, 0, 0, 0]])
    mask = np.array([[0, 0, 0], [0, 1, 0], [0, 0, 0]], bool)
    h, u, v = oracle.pos_fraction(z, mask)
    assert np.array_equal(h, [[0, 0, 0], [0, 1.0, 0], [0, 0, 0]])
    assert np.array_equal(u, [[0, 0, 0], [0, 0.5, 0], [0, 0.5, 0], [0, 0, 0]])
    assert np.array_equal(v, [[0, 0, 0, 0], [0, 0.5, 0.5, 0], [0, 0, 0, 0]])
    h, u, v = oracle.pos_fraction(np.ones((3, 3)), mask)
    assert np.array_equal(h, [[0, 0, 0], [0, 1.0, 0], [0, 0, 0]])
    assert np.array_equal(u, [[0, 0, 0], [0, 0.5, 0], [0, 0.5, 0], [0, 0, 0]])


def test_pos_fraction_planar_area(oracle):
    # a plane crossing zero: grounded fraction equals the exact area of the
    # bilinear quadrant triangles, which for a linear field in x is the
    # fraction of the cell with x above the zero line
    M, N = 9, 5
    x = (np.arange(M) + 0.5)[:, None] * np.ones((1, N))
    z = x - 4.25  # zero line 3/4 of the way across cell 4 (x in [4, 5))
    h, u, _ = oracle.pos_fraction(z)
    assert np.all(h[:4] == 0.0) and np.all(h[5:] == 1.0)
    assert np.allclose(h[4], 0.75, rtol=0, atol=1e-15)
    # the u face at x = 4 sees half cells [3.5, 4] (below) and [4, 4.5] (half above)
    assert np.allclose(u[4], 0.25, rtol=0, atol=1e-15) and np.all(u[5:M] == 1.0)
    assert np.all(u[M] == 0.5)  # the last face sees only the half cell inside the grid


def test_volume_above_floatation_closed_form(W, oracle):
    # test/test_utils.jl:33-42
    grid = W.Grid(nx=20, ny=25, dx=1000.0, dy=1000.0)
    ic = W.InitialConditions(initial_thickness=500.0 * np.ones((20, 25)))
    model = W.Model(grid=grid, bed_elevation=-200.0 * np.ones((20, 25)), initial_conditions=ic)
    o = oracle.OracleModel(model)
    o.update_state()
    o.store_fields(model)
    vaf = W.volume_above_floatation(model.fields.gh.h, model.fields.gh.b, model.params, model.grid)
    assert vaf == pytest.approx((500.0 - (1028.0 / 918.0) * 200.0) * 25 * 20 * 1000.0 ** 2, rel=1e-12)


def test_quadrature_weights(W):
    # test/test_fields.jl:65-77, test/test_grids.jl:19-20
    for σ in (np.linspace(0, 1, 10), np.linspace(0, 1, 10) ** 2):
        g = W.Grid(nx=10, ny=10, nσ=10, σ=σ)
        assert np.sum(g.quadrature_weights) == pytest.approx(1.0)
        assert np.dot(g.quadrature_weights, g.σ) == pytest.approx(0.5)
    g = W.Grid()
    assert np.allclose(g.quadrature_weights, np.r_[0.5, np.ones(g.nσ - 2), 0.5] / (g.nσ - 1))


def test_glen_b_gives_A_2e17(W):
    p = W.Params()
    B = W.glen_b(265.700709, 0.0, p.glen_a_ref, p.glen_n, p.glen_a_activation_energy,
                 p.glen_temperature_ref, p.gas_const)
    assert B ** (-p.glen_n) == pytest.approx(2.0e-17, rel=1e-6)


def _assemble(o, masks, nx, ny):
    ui = masks["u_inner"].ravel(order="F")
    vi = masks["v_inner"].ravel(order="F")
    idx = [("u", k) for k in np.flatnonzero(ui)] + [("v", k) for k in np.flatnonzero(vi)]
    n = len(idx)
    A = np.zeros((n, n))
    nu, nv = (nx + 1) * ny, nx * (ny + 1)
    for c, (comp, k) in enumerate(idx):
        xu = np.zeros(nu)
        xv = np.zeros(nv)
        (xu if comp == "u" else xv)[k] = 1.0
        yu, yv = o.op_apply(xu.reshape((nx + 1, ny), order="F"), xv.reshape((nx, ny + 1), order="F"))
        A[:, c] = np.r_[yu.ravel(order="F")[ui], yv.ravel(order="F")[vi]]
    return A, idx


def test_operator_symmetric_negative_definite_local(W, oracle):
    m = channel(W, nx=12, ny=6)
    o = oracle.OracleModel(m)
    o.update_geometry()
    # a nontrivial velocity field so that viscosity and drag vary
    rng = np.random.default_rng(2)
    o._view(W._lib.FIELD["U_U"])[:] = 100 + 50 * rng.standard_normal((13 * 6))
    o._view(W._lib.FIELD["V_V"])[:] = 10 * rng.standard_normal((12 * 7))
    o.inner_update()
    masks = o.masks()
    A, idx = _assemble(o, masks, 12, 6)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).max() < 0
    # stencil reach: u-u and v-v couplings only to the 4 neighbours
    for a, (ca, ka) in enumerate(idx):
        for b, (cb, kb) in enumerate(idx):
            if A[a, b] != 0 and ca == cb:
                nxg = 13 if ca == "u" else 12
                ia, ja = ka % nxg, ka // nxg
                ib, jb = kb % nxg, kb // nxg
                assert abs(ia - ib) + abs(ja - jb) <= 1
    du, dv = o.op_diag()
    d = np.r_[du.ravel(order="F")[masks["u_inner"].ravel(order="F")], dv.ravel(order="F")[masks["v_inner"].ravel(order="F")]]
    assert np.array_equal(d, np.diag(A))


def test_operator_invariants_on_a_ragged_domain(W, oracle):
    """The operator on an irregular domain (a calving front that moves from
    row to row, interior holes, interior Dirichlet points, grounded and
    floating parts) keeps the invariants of the rectangular case: symmetric,
    negative definite, probed diagonal equal to diag(A)."""
    from helpers import ragged
    m = ragged(W, nx=20, ny=12)
    o = oracle.OracleModel(m)
    o.update_geometry()
    gf = o.field("H_GROUNDED_FRACTION").reshape((20, 12), order="F")[m.fields.gh.mask]
    assert 0.0 < gf.mean() < 1.0
    rng = np.random.default_rng(4)
    o._view(W._lib.FIELD["U_U"])[:] = 100 + 50 * rng.standard_normal(21 * 12)
    o._view(W._lib.FIELD["V_V"])[:] = 10 * rng.standard_normal(20 * 13)
    o.inner_update()
    masks = o.masks()
    assert masks["u_inner"].sum() < m.fields.gu.mask.sum()  # Dirichlet points are not unknowns
    A, idx = _assemble(o, masks, 20, 12)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).max() < 0
    du, dv = o.op_diag()
    d = np.r_[du.ravel(order="F")[masks["u_inner"].ravel(order="F")], dv.ravel(order="F")[masks["v_inner"].ravel(order="F")]]
    assert np.array_equal(d, np.diag(A))


def test_linear_rheology_solve_matches_direct_solve(W, oracle):
    # glen_n = weertman_m = 1 make viscosity and drag constant, so Picard
    # iterations of one V-cycle each converge to the solution of one linear
    # system, which numpy solves directly from the assembled operator.
    m = W.workloads.channel_solve(nx=16, ny=8, nσ=4, maxiter_picard=60, tol_picard=1e-13,
                                  tol_coarse=1e-12)
    m.params.glen_n = 1.0
    m.params.weertman_m = 1.0
    m.fields.g3d.glen_b[:] = 1e8
    m.fields.g3d.η[:] = 0.5e8
    m.fields.gh.ηav[:] = 0.5e8
    o = oracle.OracleModel(m)
    o.update_geometry()
    o.inner_update()
    masks = o.masks()
    A, idx = _assemble(o, masks, 16, 8)
    bu, bv = o.rhs()
    b = np.r_[bu.ravel(order="F")[masks["u_inner"].ravel(order="F")], bv.ravel(order="F")[masks["v_inner"].ravel(order="F")]]
    x = np.linalg.solve(A, b)
    rep = o.update_velocities()
    u = o.field("U_U").ravel(order="F")[masks["u_inner"].ravel(order="F")]
    v = o.field("V_V").ravel(order="F")[masks["v_inner"].ravel(order="F")]
    assert rep["rel_resid"] < 1e-10
    assert rel_l2(np.r_[u, v], x) < 1e-8


def test_canonical_dot_is_a_sum(W, oracle):
    m = channel(W, nx=70, ny=33)
    o = oracle.OracleModel(m)
    rng = np.random.default_rng(4)
    a, b = rng.standard_normal((71, 33)), rng.standard_normal((70, 34))
    c, d = rng.standard_normal((71, 33)), rng.standard_normal((70, 34))
    exact = math.fsum(np.r_[(a * c).ravel(), (b * d).ravel()])
    assert o.canonical_dot(a, b, c, d) == pytest.approx(exact, rel=1e-13)
    assert o.canonical_dot(a, b) == pytest.approx(math.fsum(np.r_[(a * a).ravel(), (b * b).ravel()]), rel=1e-13)


@pytest.mark.slow
def test_iceberg_analytic(W, oracle):
    # test/verification_tests/iceberg_test.jl:85-93 (freely floating,
    # spinning iceberg: steady thickness and linear spreading + solid body
    # rotation), 10 000 steps of dt 0.1
    model, solid_u, solid_v = iceberg(W)
    model.params.dt = 0.1
    o = oracle.OracleModel(model)
    o.run_steps(10000)
    o.store_fields(model)
    f = model.fields
    h0 = ((36.0 * 0.3 / (2.0e-17)) * (1.0 / (9.81 * 918.0 * (1 - 918.0 / 1028.0))) ** 3) ** (1.0 / 4.0)
    u0 = model.grid.xxu * 0.3 / (2.0 * h0) + solid_u
    v0 = model.grid.yyv * 0.3 / (2.0 * h0) + solid_v
    hm = f.gh.mask
    relerr_h = np.linalg.norm(f.gh.h[hm] - h0) / np.linalg.norm(h0 * np.ones(hm.sum()))
    relerr_u = np.linalg.norm(f.gu.u[f.gu.mask] - u0[f.gu.mask]) / np.linalg.norm(u0[f.gu.mask])
    relerr_v = np.linalg.norm(f.gv.v[f.gv.mask] - v0[f.gv.mask]) / np.linalg.norm(v0[f.gv.mask])
    assert relerr_h < 1e-4
    assert relerr_u < 3e-4
    assert relerr_v < 3e-4


def _butterfly(x, levels, transposed):
    """The lane butterflies of the mixed-precision kernels (haar_lines_P /
    haar_lines_R) on a zero-padded 16-point line: pairs (o, o + h), o a
    multiple of 2h; P runs h = B/2 .. 1, R runs h = 1 .. B/2."""
    v = np.zeros(16)
    v[:len(x)] = x
    B = 2 << levels
    hs = [1 << k for k in range(levels + 1)]
    for h in (hs if transposed else hs[::-1]):
        y = v.copy()
        for pos in range(16):
            if pos % h:
                continue
            q = v[pos ^ h]
            if transposed:
                y[pos] = v[pos] - q if pos & h else v[pos] + q
            else:
                y[pos] = q + v[pos] if pos & h else v[pos] - q
        v = y
    return v[:len(x)]


@pytest.mark.parametrize("levels", [0, 1, 2, 3])
def test_butterfly_haar_matches_wavelet_matrix(oracle, levels):
    """The butterfly form of P = idwt and R = idwt^T used by the FP32 coarse
    CG equals wavelet_matrix's inverse and its transpose on every block
    length, partial blocks at the grid end included (zero padding)."""
    rng = np.random.default_rng(levels)
    for n in range(1, 17):
        Wi = oracle.wavelet_matrix(n, levels, False)
        x = rng.integers(-8, 9, n).astype(float)  # small integers: sums are exact in any order
        assert np.array_equal(_butterfly(x, levels, False), Wi @ x), n
        assert np.array_equal(_butterfly(x, levels, True), Wi.T @ x), n


def test_fixed_thickness_cells_keep_their_thickness(W, oracle):
    # update_thickness! (src/Simulations/run_simulation.jl:24-42) zeroes the
    # update on h_isfixed cells: over 10 steps of the channel they keep their
    # initial 100 m while the free cells evolve
    m = W.workloads.channel_transient()
    m.grid.h_isfixed[0, :] = True
    m.grid.h_isfixed[30:34, 4:6] = True
    m.params.dt = 0.5
    o = oracle.OracleModel(m)
    o.set_dt(0.5)
    o.run_steps(10)
    h = o.field("H_H").reshape((80, 10), order="F")
    fixed = m.grid.h_isfixed
    np.testing.assert_array_equal(h[fixed], 100.0)
    assert np.all(h[~fixed] != 100.0)


def test_implicitness_enters_affinely(W, oracle):
    # theta (super_implicitness) multiplies dt dsdh in the implicit thickness
    # term of the operator (update_velocities.jl:381) and of the right-hand
    # side (:110), and nothing else depends on it: both are affine in theta,
    # so at theta = 0.6 they equal 0.4 x (theta = 0) + 0.6 x (theta = 1)
    nx, ny = 12, 6
    rng = np.random.default_rng(5)
    xu, xv = rng.standard_normal((nx + 1, ny)), rng.standard_normal((nx, ny + 1))
    out = {}
    for th in (0.0, 0.6, 1.0):
        m = channel(W, nx=nx, ny=ny)
        m.solver_params.super_implicitness = th
        m.params.dt = 2.0
        m.params.accumulation_rate = 0.3 * np.ones((nx, ny))
        o = oracle.OracleModel(m)
        o.set_dt(2.0)
        o.update_geometry()
        o.inner_update()
        out[th] = o.op_apply(xu, xv) + o.rhs()
    for k in range(4):  # operator u, v, then right-hand side u, v
        a0, a6, a1 = (out[t][k] for t in (0.0, 0.6, 1.0))
        scale = max(np.abs(a0).max(), np.abs(a1).max())
        assert np.abs(a1 - a0).max() > 1e-6 * scale  # theta does enter
        assert np.abs(a6 - (0.4 * a0 + 0.6 * a1)).max() <= 1e-10 * scale


def _textbook_cg(A, b, x0, tol_coarse, maxiter):
    # Hestenes-Stiefel CG with the stop rule IterativeSolvers documents for
    # cg!: iterate while ||r|| > max(reltol ||r0||, abstol) and fewer than
    # maxiter iterations, reltol = sqrt(eps), abstol = tol_coarse ||b||
    x = x0.copy()
    r = b - A @ x
    tol = max(math.sqrt(np.finfo(float).eps) * np.linalg.norm(r), tol_coarse * np.linalg.norm(b))
    p = np.zeros_like(b)
    rr_old, it = 1.0, 0
    while np.linalg.norm(r) > tol and it < maxiter:
        rr = r @ r
        p = r + (rr / rr_old) * p
        q = A @ p
        a = rr / (p @ q)
        x += a * p
        r -= a * q
        rr_old = rr
        it += 1
    return it, x


def test_coarse_cg_follows_the_documented_cg(W, oracle):
    # The coarse solve (update_preconditioners.jl:73 calls IterativeSolvers'
    # cg!, which is not vendored) against a textbook CG on the assembled
    # coarse operator A_c = R A P: the absolute-tolerance stop, the
    # relative-tolerance stop, the iteration cap and the warm start
    nx, ny = 12, 6
    m = channel(W, nx=nx, ny=ny)
    o = oracle.OracleModel(m)
    o.update_geometry()
    o.inner_update()
    nu, nv = (nx + 1) * ny, nx * (ny + 1)
    n = nu + nv
    split = lambda z: (z[:nu].reshape((nx + 1, ny), order="F"), z[nu:].reshape((nx, ny + 1), order="F"))  # noqa: E731
    join = lambda a, b: np.r_[a.ravel(order="F"), b.ravel(order="F")]  # noqa: E731
    A = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        A[:, k] = join(*o.coarse_op(*split(e)))
    assert np.abs(A - A.T).max() <= 1e-10 * np.abs(A).max()
    rng = np.random.default_rng(8)
    b = A @ rng.standard_normal(n)  # in the range of A_c
    x0 = 0.1 * rng.standard_normal(n)
    for tol_coarse, maxiter in ((1e-3, 1000), (0.0, 1000), (0.0, 5)):
        it, xu, xv = o.coarse_cg(*split(b), *split(x0), tol_coarse, maxiter)
        it_ref, x_ref = _textbook_cg(A, b, x0, tol_coarse, maxiter)
        x = join(xu, xv)
        if maxiter == 5:
            assert it == it_ref == 5
            np.testing.assert_allclose(x, x_ref, rtol=0, atol=1e-10 * np.abs(x_ref).max())
        else:
            # rounding makes the two drift apart after many iterations on
            # this small, ill-conditioned coarse system: within 3%
            assert abs(it - it_ref) <= max(2, 0.03 * it_ref) and it < maxiter, (tol_coarse, it, it_ref)
            assert np.linalg.norm(b - A @ x) <= 1.01 * max(1.5e-8 * np.linalg.norm(b - A @ x0),
                                                           tol_coarse * np.linalg.norm(b))
    # a warm start that already meets the tolerance: no iteration, x unchanged
    _, x_ref = _textbook_cg(A, b, x0, 0.0, 1000)
    it, xu, xv = o.coarse_cg(*split(b), *split(x_ref), 1e-6, 1000)
    assert it == 0
    np.testing.assert_array_equal(join(xu, xv), x_ref)


def test_viscosity_closed_form_on_a_floating_slab(W, oracle):
    # inner_update_viscosity! (update_velocities.jl:264-283) on a floating
    # slab (grounded fraction 0, so no basal drag and tau = 0) with u = a x
    # and v = 0: the strain rate is a in every cell, so on every level the
    # viscosity is the fixed point eta = 1/2 B (a^2 + eps_reg^2)^((1/n-1)/2)
    nx, ny, ns = 10, 6, 4
    a = 2.0e-3
    grid = W.Grid(nx=nx, ny=ny, nσ=ns, dx=1000.0, dy=1000.0)
    params = W.Params(default_thickness=100.0, default_temperature=265.700709)
    m = W.Model(grid=grid, bed_elevation=-2000.0 * np.ones((nx, ny)), params=params)
    o = oracle.OracleModel(m)
    o._view(W._lib.FIELD["U_U"])[:] = (a * grid.xxu).ravel(order="F")
    o._view(W._lib.FIELD["V_V"])[:] = 0.0
    o.update_geometry()
    assert np.all(o.field("H_GROUNDED_FRACTION") == 0.0)
    o.inner_update()
    eta = o.field("S_ETA").reshape((nx, ny, ns), order="F")
    B = o.field("S_GLEN_B").reshape((nx, ny, ns), order="F")
    p = m.params
    expect = 0.5 * B * (a * a + p.glen_reg_strain_rate ** 2) ** ((1.0 / p.glen_n - 1.0) / 2.0)
    assert np.abs(eta / expect - 1.0).max() <= 1e-13
    # and n enters: at n = 3 the viscosity differs from 1/2 B
    assert np.abs(eta / (0.5 * B) - 1.0).min() > 0.5


def test_basal_drag_closed_form_on_a_sliding_slab(W, oracle):
    # update_bed_speed! and update_beta! (update_velocities.jl:227-255) from
    # the cold start (beta = 0): a grounded slab sliding uniformly, u = U,
    # v = 0, has bed speed U in every cell, so
    # beta = C (U^2 + reg^2)^((1/m - 1)/2) with C = weertman_c (grounded fraction 1)
    nx, ny = 10, 6
    U = 250.0
    grid = W.Grid(nx=nx, ny=ny, nσ=4, dx=1000.0, dy=1000.0)
    params = W.Params(default_thickness=100.0, default_temperature=265.700709, weertman_c=2.0e4)
    m = W.Model(grid=grid, bed_elevation=np.zeros((nx, ny)), params=params)
    o = oracle.OracleModel(m)
    o._view(W._lib.FIELD["U_U"])[:] = U
    o._view(W._lib.FIELD["V_V"])[:] = 0.0
    o.update_geometry()
    assert np.all(o.field("H_GROUNDED_FRACTION") == 1.0)
    assert np.all(o.field("H_BETA") == 0.0)
    o.inner_update()
    p = m.params
    expect = 2.0e4 * (U * U + p.weertman_reg_speed ** 2) ** ((1.0 / p.weertman_m - 1.0) / 2.0)
    beta = o.field("H_BETA")
    assert np.abs(beta / expect - 1.0).max() <= 1e-14
