"""Model builders shared by the tests (small grids the oracle runs in
seconds)."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def channel(W, nx=48, ny=12, precision="fp64", seed=3, maxiter_picard=3, tol_coarse=1e-8,
            maxiter_coarse=1000, nσ=4):
    """A channel single-solve setup on a small grid."""
    m = W.workloads.channel_solve(nx=nx, ny=ny, precision=precision, seed=seed, nσ=nσ,
                                  maxiter_picard=maxiter_picard, tol_coarse=tol_coarse)
    m.solver_params.maxiter_coarse = maxiter_coarse
    return m


def iceberg(W, precision="fp64", nx=20, ny=20):
    """The reference's iceberg verification setup
    (test/verification_tests/iceberg_test.jl:50-121): circular shelf with
    non-homogeneous Dirichlet points on the centre lines."""
    x0 = y0 = -50000.0
    dx = dy = 5000.0
    u_isfixed = np.zeros((nx + 1, ny), bool)
    u_isfixed[nx // 2, :] = True
    v_isfixed = np.zeros((nx, ny + 1), bool)
    v_isfixed[:, ny // 2] = True
    grid = W.Grid(nx=nx, ny=ny, nσ=4, x0=x0, y0=y0, dx=dx, dy=dy, u_isfixed=u_isfixed, v_isfixed=v_isfixed)
    h_mask = np.sqrt(grid.xxh ** 2 + grid.yyh ** 2) < 45000
    grid = W.Grid(nx=nx, ny=ny, nσ=4, x0=x0, y0=y0, dx=dx, dy=dy, u_isfixed=u_isfixed,
                  v_isfixed=v_isfixed, h_mask=h_mask)
    thick = np.zeros((nx, ny))
    thick[h_mask] = 200.0
    u_drift, v_drift = 5.0, 3.0
    rate = 2 * np.pi / 10000.0
    solid_u = u_drift * np.ones((nx + 1, ny)) - rate * grid.yyu
    solid_v = v_drift * np.ones((nx, ny + 1)) + rate * grid.xxv
    u_bc = np.zeros((nx + 1, ny))
    v_bc = np.zeros((nx, ny + 1))
    u_bc[u_isfixed] = solid_u[u_isfixed]
    v_bc[v_isfixed] = solid_v[v_isfixed]
    ic = W.InitialConditions(initial_thickness=thick, initial_u_veloc=u_bc, initial_v_veloc=v_bc)
    params = W.Params(accumulation_rate=0.3, default_temperature=265.700709)
    sp = W.SolverParams(maxiter_picard=1)
    model = W.Model(grid=grid, bed_elevation=-500.0 * np.ones((nx, ny)), params=params,
                    solver_params=sp, initial_conditions=ic, parallel_spec=W.HIPSpec(precision))
    return model, solid_u, solid_v


def ragged(W, nx=40, ny=24, precision="fp64", seed=7, maxiter_picard=3, nσ=4, dirichlet=True, weertman_c=1e4):
    """An irregular domain: a calving front whose position varies from row to
    row, interior holes (ice-free cells inside the ice, as around nunataks),
    a retrograde bed that grounds the upstream half and floats the rest, and
    interior Dirichlet points on u with nonzero values (the inflow of a
    tributary).  The walls follow the channel (u zero on the north edge, v zero
    on the west and east edges).  dirichlet=False leaves out the interior
    Dirichlet points (the adjoint gradient does not take them)."""
    rng = np.random.default_rng(seed)
    dx = dy = 4000.0
    h_mask = np.ones((nx, ny), bool)
    # The front moves in steps of two rows: a floating finger one cell wide
    # has no c-grid corner (no shear coupling) and no drag, so a rigid
    # translation of it along y is a null vector of the operator, and the
    # problem would be singular.  For the same reason the holes keep apart.
    front = np.repeat(nx - 2 - rng.integers(0, max(nx // 5, 2), (ny + 1) // 2), 2)[:ny]
    for j in range(ny):
        h_mask[front[j]:, j] = False
    holes = []
    while len(holes) < 3:
        i = int(rng.integers(nx // 5, nx // 2))
        j = int(rng.integers(2, ny - 4))
        if all(abs(i - a) > 3 or abs(j - b) > 3 for a, b in holes):
            holes.append((i, j))
            h_mask[i:i + 2, j:j + 2] = False
    hi, hj = nx // 3, ny // 2
    if all(abs(hi - a) > 3 or abs(hj - b) > 3 for a, b in holes):
        h_mask[hi, hj] = False
    u_isfixed = np.zeros((nx + 1, ny), bool)
    u_isfixed[nx // 4, 1:4] = dirichlet
    grid = W.Grid(nx=nx, ny=ny, nσ=nσ, x0=0.0, y0=-ny * dy / 2, dx=dx, dy=dy, h_mask=h_mask,
                  u_iszero=["north"], v_iszero=["west", "east"], u_isfixed=u_isfixed)
    x = grid.xxh
    bed = 150.0 - 900.0 * x / (nx * dx) + rng.normal(0.0, 10.0, x.shape)
    thick = np.where(h_mask, np.maximum(1200.0 * (1.0 - 0.8 * x / (nx * dx)), 50.0), 0.0)
    thick = np.asfortranarray(thick + np.where(h_mask, rng.normal(0.0, 5.0, x.shape), 0.0))
    u0 = np.zeros((nx + 1, ny))
    u0[u_isfixed] = 80.0
    ic = W.InitialConditions(initial_thickness=thick, initial_u_veloc=u0)
    params = W.Params(default_temperature=265.700709, accumulation_rate=0.3, weertman_c=weertman_c)
    sp = W.SolverParams(maxiter_picard=maxiter_picard, tol_coarse=1e-8)
    return W.Model(grid=grid, bed_elevation=np.asfortranarray(bed), params=params, solver_params=sp,
                   initial_conditions=ic, parallel_spec=W.HIPSpec(precision))
