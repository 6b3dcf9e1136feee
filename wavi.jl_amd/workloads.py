"""Seeded synthetic inputs for the velocity-solve workloads.

Every workload is a WAVI model setup:
  mismip_plus_fixture  MISMIP+ 80x10, nσ=4, dt 0.1, maxiter_picard=1: the setup of
                       the reference's regression fixtures
                       (test/version_update_test_verification/test_version_updates.jl:6-70)
  channel_transient    80x10 at 8 km, nσ=4, H0=100 m, acc 0.3 m/a, dt 0.5 a,
                       maxiter_picard=1, tol_coarse=1e-6, 200 steps
  channel_solve        a single velocity solve on an nx x ny grid (1280x160 at
                       500 m or 2560x320 at 250 m), nσ=8, 20 Picard iterations
                       to tol 1e-8, tol_coarse 1e-8, 3 Haar levels, dt 1.0
  inversion_*          the basal-drag inversion: true log C = log 1e4 + a
                       seeded Gaussian random field (std 0.5, correlation
                       length 10 km), prior log C = log 1e4, observations =
                       the FP64 velocities of the true field times
                       (1 + 0.01 N(0, 1)), 10 Picard iterations
  ensemble_members     64 channel solves of seeded sizes and parameters
                       (grid from {160x20, 320x40, 640x80, 1280x160},
                       nσ from {4, 8}, A log-uniform in [1e-17, 4e-17],
                       log C = log 1e4 + a seeded random field of std 0.3)

Channel bed: the MISMIP+ along-flow sixth-order polynomial (from -150 m down
to the -720 m floor, with its retrograde section), plus side-wall ridges of
Gaussian shape 1000*exp(-((y -+ 40 km)/4 km)^2) at y = +-40 km, plus seeded
roughness: iid N(0, 10^2) per cell on the 8 km grid; on the finer grids a
Gaussian random field (std 50 m, correlation exp(-r^2/l^2) with l = 5 km, made
by filtering seeded white noise with exp(-k^2 l^2/8) in Fourier space and
rescaling to the target standard deviation).
Thickness of the single solves: H(x) = 1800 (1 - x/Lx)^0.5 + N(0, 20^2),
clipped at 10 m.  Physics: Glen n=3 with A = 2e-17 Pa^-3 a^-1 (temperature
265.700709 K with the default reference rate factor), Weertman m=3,
C = 1e4, rho_i 918, rho_w 1028; u fixed to zero on "north" (x index 1) and v
on "east"/"west".
"""
import numpy as np

from .core import Grid, InitialConditions, Params, SolverParams, mismip_plus_bed
from .model import HIPSpec, Model

A_TEMPERATURE = 265.700709  # temperature at which glen_b gives A = 2e-17


def bed_polynomial(x):
    xbar = 300000.0
    b0, b2, b4, b6 = -150.0, -728.8, 343.91, -50.75
    return b0 + b2 * (x / xbar) ** 2 + b4 * (x / xbar) ** 4 + b6 * (x / xbar) ** 6


def side_ridges(y, height=1000.0, width=4000.0, yc=40000.0):
    return height * (np.exp(-((y - yc) / width) ** 2) + np.exp(-((y + yc) / width) ** 2))


def gaussian_random_field(nx, ny, dx, dy, std, corr_len, seed):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((nx, ny))
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=dy)
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * np.exp(-k2 * corr_len ** 2 / 8.0)))
    f = f - f.mean()
    return np.asfortranarray(f * (std / f.std()))


def channel_bed(grid, roughness="white", seed=1):
    b = np.maximum(bed_polynomial(grid.xxh) + side_ridges(grid.yyh), -720.0)
    if roughness == "white":
        b = b + np.random.default_rng(seed).normal(0.0, 10.0, b.shape)
    elif roughness == "grf":
        b = b + gaussian_random_field(grid.nx, grid.ny, grid.dx, grid.dy, 50.0, 5000.0, seed)
    return np.asfortranarray(b)


def mismip_plus_fixture(precision="fp64"):
    """Setup of the reference's MISMIP+ regression fixtures (dt is set by the
    Simulation to 0.1; 1000 steps)."""
    grid = Grid(nx=80, ny=10, nσ=4, x0=0.0, y0=-40000.0, dx=8000.0, dy=8000.0,
                u_iszero=["north"], v_iszero=["west", "east"])
    solver_params = SolverParams(maxiter_picard=1)
    params = Params(default_thickness=100.0, accumulation_rate=0.3,
                    default_temperature=A_TEMPERATURE)
    return Model(grid=grid, bed_elevation=mismip_plus_bed, params=params,
                 solver_params=solver_params, parallel_spec=HIPSpec(precision))


def channel_transient(precision="fp64", seed=1, nx=80, ny=10):
    """80x10 channel transient run (200 steps of dt 0.5 set by Simulation)."""
    grid = Grid(nx=nx, ny=ny, nσ=4, x0=0.0, y0=-40000.0, dx=8000.0 * 80 / nx,
                dy=8000.0 * 10 / ny, u_iszero=["north"], v_iszero=["west", "east"])
    solver_params = SolverParams(maxiter_picard=1, tol_coarse=1e-6)
    params = Params(default_thickness=100.0, accumulation_rate=0.3,
                    default_temperature=A_TEMPERATURE, weertman_c=1e4, weertman_m=3.0)
    bed = channel_bed(grid, "white", seed)
    return Model(grid=grid, bed_elevation=bed, params=params, solver_params=solver_params,
                 parallel_spec=HIPSpec(precision))


def channel_solve(nx=1280, ny=160, precision="fp64", seed=2, nσ=8, maxiter_picard=20,
                  tol_picard=1e-8, tol_coarse=1e-8, weertman_c=1e4, glen_a_ref=None):
    """Single velocity solve on the channel (1280x160 at 500 m by default).
    The domain is 640 km x 80 km for every nx, ny."""
    Lx, Ly = 640000.0, 80000.0
    grid = Grid(nx=nx, ny=ny, nσ=nσ, x0=0.0, y0=-40000.0, dx=Lx / nx, dy=Ly / ny,
                u_iszero=["north"], v_iszero=["west", "east"])
    solver_params = SolverParams(maxiter_picard=maxiter_picard, tol_picard=tol_picard,
                                 tol_coarse=tol_coarse, levels=3)
    kw = {}
    if glen_a_ref is not None:
        kw["glen_a_ref"] = glen_a_ref
    params = Params(default_temperature=A_TEMPERATURE, weertman_c=weertman_c, weertman_m=3.0, **kw)
    bed = channel_bed(grid, "grf", seed)
    rng = np.random.default_rng(seed + 1000)
    x = grid.xxh
    H = 1800.0 * np.sqrt(np.maximum(1.0 - x / Lx, 0.0)) + rng.normal(0.0, 20.0, x.shape)
    H = np.asfortranarray(np.maximum(H, 10.0))
    ic = InitialConditions(initial_thickness=H)
    return Model(grid=grid, bed_elevation=bed, params=params, solver_params=solver_params,
                 initial_conditions=ic, parallel_spec=HIPSpec(precision))


ENSEMBLE_SIZES = ((160, 20), (320, 40), (640, 80), (1280, 160))


def ensemble_specs(n=64, seed=5):
    """Seeded member list of the ensemble: dicts with nx, ny, nσ, A (Pa^-3 a^-1)
    and the seed of the member's log C field and roughness."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n):
        nx, ny = ENSEMBLE_SIZES[rng.integers(len(ENSEMBLE_SIZES))]
        nσ = (4, 8)[rng.integers(2)]
        A = float(np.exp(rng.uniform(np.log(1e-17), np.log(4e-17))))
        specs.append({"nx": int(nx), "ny": int(ny), "nσ": int(nσ), "A": A,
                      "seed": int(seed * 1000 + k)})
    return specs


def ensemble_member(spec, precision="mixed", maxiter_picard=20):
    """One ensemble member: channel_solve at the member's size and nσ, with
    Glen A scaled to spec["A"] (A is proportional to glen_a_ref at a fixed
    temperature) and C = 1e4 exp(f), f a Gaussian random field of std 0.3 and
    correlation length 10 km."""
    nx, ny = spec["nx"], spec["ny"]
    Lx, Ly = 640000.0, 80000.0
    glen_a_ref = Params().glen_a_ref * spec["A"] / 2e-17
    f = gaussian_random_field(nx, ny, Lx / nx, Ly / ny, 0.3, 10000.0, spec["seed"])
    return channel_solve(nx=nx, ny=ny, precision=precision, seed=spec["seed"], nσ=spec["nσ"],
                         maxiter_picard=maxiter_picard, weertman_c=np.asfortranarray(1e4 * np.exp(f)),
                         glen_a_ref=glen_a_ref)


def ensemble_members(n=64, seed=5, precision="mixed", maxiter_picard=20):
    """The n members of the seeded ensemble, as Models."""
    return [ensemble_member(sp, precision, maxiter_picard) for sp in ensemble_specs(n, seed)]


def inversion_logc_true(nx=1280, ny=160, seed=4):
    """True log C of the inversion workload: log 1e4 + GRF(std 0.5, 10 km)."""
    Lx, Ly = 640000.0, 80000.0
    return np.log(1e4) + gaussian_random_field(nx, ny, Lx / nx, Ly / ny, 0.5, 10000.0, seed)


def inversion_model(logc, precision="fp64", nσ=8, n_picard=10, seed=2):
    """Channel model of the inversion workload with C = exp(logc)."""
    nx, ny = logc.shape
    return channel_solve(nx=nx, ny=ny, precision=precision, seed=seed, nσ=nσ, maxiter_picard=n_picard,
                         weertman_c=np.asfortranarray(np.exp(logc)))


def inversion_observations(u, v, seed=4):
    """Synthetic observations: velocities times (1 + 0.01 N(0, 1))."""
    rng = np.random.default_rng(seed + 7)
    return (np.asfortranarray(u * (1.0 + 0.01 * rng.standard_normal(u.shape))),
            np.asfortranarray(v * (1.0 + 0.01 * rng.standard_normal(v.shape))))
