"""Host-side parameter and grid objects, mirroring WAVI.jl's constructors.

Arrays use WAVI.jl's shapes: (nx, ny) for the h grid, (nx+1, ny) for u,
(nx, ny+1) for v, (nx, ny, nsigma) for the sigma grid, stored in Fortran
order so that x is the fastest index as in Julia.
"""
import math
from dataclasses import dataclass, field

import numpy as np


class ArgumentError(ValueError):
    """Raised where WAVI.jl throws ArgumentError."""


class DimensionMismatch(ValueError):
    """Raised where WAVI.jl throws DimensionMismatch."""


class BoundsError(IndexError):
    """Raised where WAVI.jl throws BoundsError."""


def _bool2d(a, shape, name):
    try:
        a = np.asarray(a)
        if a.dtype != np.bool_:
            if not np.all((a == 0) | (a == 1)):
                raise ValueError
            a = a.astype(np.bool_)
    except (TypeError, ValueError):
        raise ArgumentError("%s must be Boolean (or equivalent)" % name)
    if a.shape != shape:
        raise DimensionMismatch("%s size must be %s" % (name, shape))
    return np.asfortranarray(a)


def orientations2bc(orientations, M, N):
    """M x N mask true on the named edges (src/Grid.jl:211-222): "north" is
    x index 1, "south" x index end, "west" y index 1, "east" y index end."""
    A = np.zeros((M, N), dtype=np.bool_, order="F")
    for o in orientations:
        o = o.lower()
        if o == "north":
            A[0, :] = True
        if o == "south":
            A[-1, :] = True
        if o == "east":
            A[:, -1] = True
        if o == "west":
            A[:, 0] = True
    return A


class Grid:
    """Grid(; nx, ny, dx, dy, nσ, x0, y0, h_mask, ...) (src/Grid.jl:66-202)."""

    def __init__(self, nx=80, ny=10, dx=8000.0, dy=8000.0, nσ=4, x0=0.0, y0=-40000.0,
                 h_mask=None, h_isfixed=None, u_iszero=None, v_iszero=None,
                 u_isfixed=None, v_isfixed=None, quadrature_weights=None, σ=None,
                 basin_ID=None):
        for name, v in (("nx", nx), ("ny", ny), ("nσ", nσ)):
            if not (isinstance(v, (int, np.integer)) and v > 1):
                raise ArgumentError("number of grid cells (%s) must a positive integer larger than one" % name)
        if isinstance(u_iszero, (list, tuple)) and all(isinstance(s, str) for s in u_iszero):
            u_iszero = orientations2bc(u_iszero, nx + 1, ny)
        if isinstance(v_iszero, (list, tuple)) and all(isinstance(s, str) for s in v_iszero):
            v_iszero = orientations2bc(v_iszero, nx, ny + 1)
        if isinstance(h_isfixed, (list, tuple)) and all(isinstance(s, str) for s in h_isfixed):
            h_isfixed = orientations2bc(h_isfixed, nx, ny)
        if isinstance(u_isfixed, (list, tuple)) and all(isinstance(s, str) for s in u_isfixed):
            u_isfixed = orientations2bc(u_isfixed, nx + 1, ny)
        if isinstance(v_isfixed, (list, tuple)) and all(isinstance(s, str) for s in v_isfixed):
            v_isfixed = orientations2bc(v_isfixed, nx, ny + 1)
        self.nx, self.ny, self.nσ = int(nx), int(ny), int(nσ)
        self.dx, self.dy = float(dx), float(dy)
        self.x0, self.y0 = float(x0), float(y0)
        self.h_mask = _bool2d(np.ones((nx, ny), bool) if h_mask is None else h_mask, (nx, ny), "h_mask")
        self.h_isfixed = _bool2d(np.zeros((nx, ny), bool) if h_isfixed is None else h_isfixed, (nx, ny), "h_isfixed")
        self.u_iszero = _bool2d(np.zeros((nx + 1, ny), bool) if u_iszero is None else u_iszero, (nx + 1, ny), "u_iszero")
        self.v_iszero = _bool2d(np.zeros((nx, ny + 1), bool) if v_iszero is None else v_iszero, (nx, ny + 1), "v_iszero")
        self.u_isfixed = _bool2d(np.zeros((nx + 1, ny), bool) if u_isfixed is None else u_isfixed, (nx + 1, ny), "u_isfixed")
        self.v_isfixed = _bool2d(np.zeros((nx, ny + 1), bool) if v_isfixed is None else v_isfixed, (nx, ny + 1), "v_isfixed")
        self.basin_ID = np.ones((nx, ny), order="F") if basin_ID is None else np.asfortranarray(basin_ID, dtype=float)
        if self.basin_ID.shape != (nx, ny):
            raise DimensionMismatch("Basin_ID size must be (nx x ny)")
        i = np.arange(1, nx + 2, dtype=float)[:, None]
        j = np.arange(1, ny + 2, dtype=float)[None, :]
        self.xxh = np.asfortranarray(np.broadcast_to(x0 + (i[:nx] - 0.5) * dx, (nx, ny)))
        self.yyh = np.asfortranarray(np.broadcast_to(y0 + (j[:, :ny] - 0.5) * dy, (nx, ny)))
        self.xxu = np.asfortranarray(np.broadcast_to(x0 + (i - 1.0) * dx, (nx + 1, ny)))
        self.yyu = np.asfortranarray(np.broadcast_to(y0 + (j[:, :ny] - 0.5) * dy, (nx + 1, ny)))
        self.xxv = np.asfortranarray(np.broadcast_to(x0 + (i[:nx] - 0.5) * dx, (nx, ny + 1)))
        self.yyv = np.asfortranarray(np.broadcast_to(y0 + (j - 1.0) * dy, (nx, ny + 1)))
        self.xxc = np.asfortranarray(np.broadcast_to(x0 + i[:nx - 1] * dx, (nx - 1, ny - 1)))
        self.yyc = np.asfortranarray(np.broadcast_to(y0 + j[:, :ny - 1] * dy, (nx - 1, ny - 1)))
        if σ is None:
            σ = np.linspace(0.0, 1.0, nσ)
        σ = np.asarray(σ, dtype=float)
        if σ.shape != (nσ,):
            raise DimensionMismatch("number of sigma levels (= %d) must match number of σ grid points (=%d)" % (σ.size, nσ))
        if not (σ[0] == 0.0 and σ[-1] == 1.0):
            raise AssertionError("σ must start at 0 and end at 1")
        self.σ = σ
        self.ζ = 1.0 - σ
        if quadrature_weights is None:
            quadrature_weights = 0.5 * np.concatenate(([σ[1] - σ[0]], σ[2:] - σ[:-2], [σ[-1] - σ[-2]]))
        quadrature_weights = np.asarray(quadrature_weights, dtype=float)
        if quadrature_weights.shape != (nσ,):
            raise DimensionMismatch("quadrature weights must have size (nσ,)")
        self.quadrature_weights = quadrature_weights


class _Sealed:
    """Fields fixed at construction, as for a Julia struct: assigning a name
    that is not a field raises AttributeError instead of adding it."""
    _sealed = False

    def __setattr__(self, name, value):
        if self._sealed and not hasattr(self, name):
            raise AttributeError("%s has no field %s" % (type(self).__name__, name))
        object.__setattr__(self, name, value)


@dataclass
class SolverParams(_Sealed):
    """SolverParams (src/SolverParams.jl:2-14)."""
    n_iter_viscosity: int = 2
    tol_picard: float = 1e-5
    maxiter_picard: int = 30
    tol_coarse: float = 1e-5
    maxiter_coarse: int = 1000
    levels: int = 3
    wavelet_threshold: float = 10.0
    nsmooth: int = 5
    smoother_omega: float = 1.0
    stencil_margin: int = 3
    super_implicitness: float = 1.0

    def __post_init__(self):
        if self.n_iter_viscosity != 2:
            raise AssertionError("n_iter_viscosity == 2")
        self._sealed = True


class Params(_Sealed):
    """Params(; ...) (src/Params.jl:280-331). dt defaults to 1.0 and is set
    by Simulation."""

    def __init__(self, g=9.81, density_ice=918.0, density_ocean=1028.0, gas_const=8.314,
                 sec_per_year=3600 * 24 * 365.25, default_thickness=100.0,
                 default_viscosity=1.0e7, default_temperature=263.15, default_damage=0.0,
                 accumulation_rate=0.0, glen_a_activation_energy=5.8631e+04,
                 glen_a_ref=4.9e-16 * 3600 * 24 * 365.25 * 1.0e-9, glen_temperature_ref=263.15,
                 glen_n=3.0, glen_reg_strain_rate=1.0e-5, weertman_c=1.0e4, weertman_m=3.0,
                 weertman_reg_speed=1.0e-5, sea_level_wrt_geoid=0.0, minimum_thickness=50.0,
                 evolveShelves=True, smallHAF=1.0):
        self.dt = 1.0
        self.g = g
        self.density_ice = density_ice
        self.density_ocean = density_ocean
        self.gas_const = gas_const
        self.sec_per_year = sec_per_year
        self.default_thickness = default_thickness
        self.default_viscosity = default_viscosity
        self.default_temperature = default_temperature
        self.default_damage = default_damage
        self.accumulation_rate = accumulation_rate
        self.glen_a_activation_energy = glen_a_activation_energy
        self.glen_a_ref = glen_a_ref
        self.glen_temperature_ref = glen_temperature_ref
        self.glen_n = glen_n
        self.glen_reg_strain_rate = glen_reg_strain_rate
        self.weertman_c = weertman_c
        self.weertman_m = weertman_m
        self.weertman_reg_speed = weertman_reg_speed
        self.sea_level_wrt_geoid = sea_level_wrt_geoid
        self.minimum_thickness = minimum_thickness
        self.evolveShelves = evolveShelves
        self.smallHAF = smallHAF
        self._sealed = True


class InitialConditions:
    """InitialConditions (src/InitialConditions.jl:27-35); unset entries are
    NaN and replaced by defaults when the Model is built."""

    def __init__(self, initial_thickness=None, initial_grounded_fraction=None,
                 initial_u_veloc=None, initial_v_veloc=None, initial_viscosity=None,
                 initial_temperature=None, initial_damage=None):
        nan2 = np.full((1, 1), np.nan)
        nan3 = np.full((1, 1, 1), np.nan)
        self.initial_thickness = nan2 if initial_thickness is None else np.asarray(initial_thickness, float)
        self.initial_grounded_fraction = nan2 if initial_grounded_fraction is None else np.asarray(initial_grounded_fraction, float)
        self.initial_u_veloc = nan2 if initial_u_veloc is None else np.asarray(initial_u_veloc, float)
        self.initial_v_veloc = nan2 if initial_v_veloc is None else np.asarray(initial_v_veloc, float)
        self.initial_viscosity = nan3 if initial_viscosity is None else np.asarray(initial_viscosity, float)
        self.initial_temperature = nan3 if initial_temperature is None else np.asarray(initial_temperature, float)
        self.initial_damage = nan3 if initial_damage is None else np.asarray(initial_damage, float)


@dataclass
class UniformMeltRate:
    """UniformMeltRate(; m=0.0) (src/MeltRate/MeltRate.jl:37-53)."""
    m: float = 0.0


def compute_iterations_and_end_time(end_time, n_iter_total, dt):
    """src/TimesteppingParams.jl:81-98."""
    if not (isinstance(dt, (int, float)) and dt > 0):
        raise ArgumentError("timestep dt must be a positive number")
    if n_iter_total is not None and end_time is not None:
        if not math.isclose(end_time, n_iter_total * dt, rel_tol=1.5e-8):
            raise ArgumentError("end_time must equal n_iter_total * dt")
    elif n_iter_total is None and end_time is not None:
        n_iter_total = math.inf if end_time == math.inf else int(round(end_time / dt))
    elif n_iter_total is not None and end_time is None:
        end_time = math.inf if n_iter_total == math.inf else dt * n_iter_total
    else:
        raise ArgumentError("You must pass at least one of end_time or n_iter_total")
    if not (isinstance(n_iter_total, int) and n_iter_total > 0):
        raise ArgumentError("n_iter_total must be a positive integer")
    if not end_time > 0:
        raise ArgumentError("end_time must be a positive number")
    return end_time, n_iter_total


class TimesteppingParams:
    """TimesteppingParams(; niter0, dt, end_time, n_iter_total, chkpt_freq,
    pchkpt_freq, chkpt_path, step_thickness) (src/TimesteppingParams.jl:43-74).
    Checkpoints are written every n_iter_chkpt steps (alternating files A and
    B) and every n_iter_pchkpt steps (permanent) by run_simulation."""

    def __init__(self, niter0=0, dt=1.0, end_time=None, n_iter_total=None, chkpt_freq=math.inf,
                 pchkpt_freq=math.inf, chkpt_path="./", step_thickness=True):
        import os
        import warnings
        self.niter0 = niter0
        self.dt = float(dt)
        self.end_time, self.n_iter_total = compute_iterations_and_end_time(end_time, n_iter_total, dt)
        self.t0 = niter0 * dt if niter0 > 0 else 0.0
        self.chkpt_freq = chkpt_freq
        self.pchkpt_freq = pchkpt_freq
        self.n_iter_chkpt = math.inf if chkpt_freq == math.inf else int(round(chkpt_freq / dt))
        self.n_iter_pchkpt = math.inf if pchkpt_freq == math.inf else int(round(pchkpt_freq / dt))
        if not chkpt_path.endswith("/"):
            chkpt_path = chkpt_path + "/"
        if not os.path.isdir(chkpt_path):
            warnings.warn("Did not find output path %s. Any outputs will go to the working directory %s"
                          % (chkpt_path, os.getcwd()))
            chkpt_path = "./"
        self.chkpt_path = chkpt_path
        self.step_thickness = step_thickness


class OutputParams:
    """OutputParams(; outputs, output_freq, output_format, prefix, output_path,
    dump_vel, zip_format, output_start) (src/OutputParams/OutputParams.jl:2-64).
    Of the outputs, this package writes the velocity dump (dump_vel) only;
    field outputs (a non-empty outputs) are rejected."""

    def __init__(self, outputs=(), output_freq=math.inf, output_format="jld2", prefix="outfile",
                 output_path="./", dump_vel=False, zip_format="none", output_start=False):
        if len(outputs) > 0:
            raise ArgumentError("field outputs are not supported; only dump_vel")
        self.outputs = outputs
        self.output_freq = output_freq
        self.output_format = output_format
        self.prefix = prefix
        self.output_path = output_path
        self.dump_vel = bool(dump_vel)
        self.zip_format = zip_format
        self.output_start = output_start
        self.n_iter_out = math.inf


def glen_b(temperature, damage, glen_a_ref, glen_n, glen_a_activation_energy,
           glen_temperature_ref, gas_const):
    """Stiffness B of Glen's law (src/utilities.jl:462-466)."""
    glen_a0 = glen_a_ref * np.exp(+glen_a_activation_energy / (glen_temperature_ref * gas_const))
    return (1 - damage) * (glen_a0 * np.exp(-glen_a_activation_energy / (temperature * gas_const))) ** (-1.0 / glen_n)


def mismip_plus_bed(x, y):
    """MISMIP+ bed (src/utilities.jl:401-409)."""
    xbar = 300000.0
    b0, b2, b4, b6 = -150.0, -728.8, 343.91, -50.75
    wc, fc, dc = 24000.0, 4000.0, 500.0
    bx = b0 + b2 * (x / xbar) ** 2 + b4 * (x / xbar) ** 4 + b6 * (x / xbar) ** 6
    by = dc * ((1 + np.exp(-2 * (y - wc) / fc)) ** (-1) + (1 + np.exp(2 * (y + wc) / fc)) ** (-1))
    return np.maximum(bx + by, -720.0)


def get_u_mask(h_mask):
    """src/utilities.jl:475-482."""
    nx, ny = h_mask.shape
    u = np.zeros((nx + 1, ny), bool, order="F")
    u[:-1, :] |= h_mask
    u[1:, :] |= h_mask
    return u


def get_v_mask(h_mask):
    """src/utilities.jl:489-496."""
    nx, ny = h_mask.shape
    v = np.zeros((nx, ny + 1), bool, order="F")
    v[:, :-1] |= h_mask
    v[:, 1:] |= h_mask
    return v


def get_c_mask(h_mask):
    """src/utilities.jl:503-507."""
    return np.asfortranarray(h_mask[:-1, :-1] & h_mask[:-1, 1:] & h_mask[1:, :-1] & h_mask[1:, 1:])


def clip(trial_mask):
    """src/utilities.jl:515-528."""
    good = get_c_mask(trial_mask)
    nx, ny = trial_mask.shape
    m = np.zeros((nx, ny), bool, order="F")
    m[:-1, :-1] |= good
    m[:-1, 1:] |= good
    m[1:, :-1] |= good
    m[1:, 1:] |= good
    return m


def height_above_floatation(h, b, params):
    """src/utilities.jl:445."""
    return h - (params.density_ocean / params.density_ice) * (params.sea_level_wrt_geoid - b)


def get_glx(model):
    """Grounding-line x positions (src/utilities.jl:417-426).  Wherever the
    sign of the height above floatation drops from +1 to -1 between the h
    cells (i, j) and (i+1, j), the zero of the straight line through the two
    cell centres.  The crossings come in the reference's column-major order:
    x fastest, then y."""
    haf = np.asarray(model.fields.gh.haf, dtype=float)
    xx = np.asarray(model.grid.xxh, dtype=float)
    s = np.sign(haf)
    cross = (s[1:, :] - s[:-1, :]) == -2
    jj, ii = np.nonzero(cross.T)  # row-major over (j, i): i varies fastest
    x1, x2 = xx[ii, jj], xx[ii + 1, jj]
    h1, h2 = haf[ii, jj], haf[ii + 1, jj]
    return x1 + (x2 - x1) * (0.0 - h1) / (h2 - h1)


def volume_above_floatation(h, b, params, grid):
    """src/utilities.jl:453."""
    haf = height_above_floatation(h, b, params)
    return np.sum(haf[haf > 0]) * grid.dx * grid.dy
