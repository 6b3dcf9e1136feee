"""Restricted additive Schwarz (RAS) velocity solve: SharedMemorySpec
(src/SharedMemorySpec.jl).

The full-domain model keeps one device context (SharedMemorySpec is a
HIPSpec) for its inner update, right-hand side, residual, geometry and the
steps after the solve.  Every subdomain is a Model of its own with its own
device context.  update_preconditioner! builds the subdomain models from
the full model's fields once per solve, as the reference does: on the host
the first time, then on the device into the persistent subdomain contexts
(wavi_schwarz_refresh).
Each Picard iteration then runs on the device (wavi_schwarz_precondition):
restriction, update_state! of every subdomain (host threads, one stream
per subdomain, as the reference spawns one task per subdomain), damping and
the partition-of-unity prolongation, with no field crossing to the host.
The host functions below (schwarz_restrict_velocities,
schwarz_prolong_velocities, schwarz_partition_of_unity) restate the
reference's for host Models.  Subdomain contexts persist across solves while
the subdomain grid is unchanged.
"""
import copy
import ctypes
import math

import numpy as np

from . import _lib
from .core import ArgumentError, BoundsError, Grid, InitialConditions
from .model import HIPSpec, Model, _check, _report_solve, pull_fields, push_fields, update_preconditioner


class SharedMemorySpec(HIPSpec):
    """SharedMemorySpec(; ngridsx, ngridsy, overlap, niterations, damping)
    (SharedMemorySpec.jl:6-13).  precision applies to the full domain's and
    the subdomains' device solves; nthreads host threads drive the
    subdomains."""

    def __init__(self, ngridsx=1, ngridsy=1, overlap=1, niterations=0, damping=0.0, precision="fp64",
                 nthreads=8):
        super().__init__(precision)
        self.ngridsx, self.ngridsy = int(ngridsx), int(ngridsy)
        self.overlap, self.niterations = int(overlap), int(niterations)
        self.damping = float(damping)
        self.nthreads = int(nthreads)
        self.schwarz_model_array = [[None] * self.ngridsy for _ in range(self.ngridsx)]

    def close(self):
        for row in getattr(self, "schwarz_model_array", []):
            for mg in row:
                if mg is not None:
                    mg.parallel_spec.close()
        super().close()


def _bounds(model, igrid, jgrid, ngridsx, ngridsy, overlap):
    """Index ranges (0-based, stop exclusive) of subdomain (igrid, jgrid)
    (1-based, as in the reference) with its overlap (SharedMemorySpec.jl:183-199)."""
    nx, ny = model.grid.nx, model.grid.ny
    if nx % ngridsx != 0:
        raise ArgumentError("Model domain is not an integer number of subdomains in x-direction")
    if ny % ngridsy != 0:
        raise ArgumentError("Model domain is not an integer number of subdomains in y-direction")
    if not 1 <= igrid <= ngridsx:
        raise BoundsError("igrid %d out of 1..%d" % (igrid, ngridsx))
    if not 1 <= jgrid <= ngridsy:
        raise BoundsError("jgrid %d out of 1..%d" % (jgrid, ngridsy))
    nxd, nyd = nx // ngridsx, ny // ngridsy
    i0 = max((igrid - 1) * nxd + 1 - overlap, 1)
    i1 = min(igrid * nxd + overlap, nx)
    j0 = max((jgrid - 1) * nyd + 1 - overlap, 1)
    j1 = min(jgrid * nyd + overlap, ny)
    return i0 - 1, i1, j0 - 1, j1


def schwarz_model(model, igrid=1, jgrid=1, ngridsx=1, ngridsy=1, overlap=1, parallel_spec=None):
    """schwarzModel (SharedMemorySpec.jl:181-287): the subdomain model, its
    halo velocities fixed, its fields taken from the full model."""
    g = model.grid
    f = model.fields
    i0, i1, j0, j1 = _bounds(model, igrid, jgrid, ngridsx, ngridsy, overlap)
    nxg, nyg = i1 - i0, j1 - j0
    u_isfixed = g.u_isfixed[i0:i1 + 1, j0:j1].copy()
    v_isfixed = g.v_isfixed[i0:i1, j0:j1 + 1].copy()
    # halo points become fixed-velocity points
    if igrid != 1:
        u_isfixed[0, :] = True
        v_isfixed[0, :] = True
    if igrid != ngridsx:
        u_isfixed[nxg, :] = True
        v_isfixed[nxg - 1, :] = True
    if jgrid != 1:
        v_isfixed[:, 0] = True
        u_isfixed[:, 0] = True
    if jgrid != ngridsy:
        v_isfixed[:, nyg] = True
        u_isfixed[:, nyg - 1] = True
    grid_g = Grid(nx=nxg, ny=nyg, dx=g.dx, dy=g.dy, nσ=g.nσ, x0=g.x0 + i0 * g.dx, y0=g.y0 + j0 * g.dy,
                  h_mask=g.h_mask[i0:i1, j0:j1], h_isfixed=g.h_isfixed[i0:i1, j0:j1],
                  u_iszero=g.u_iszero[i0:i1 + 1, j0:j1], v_iszero=g.v_iszero[i0:i1, j0:j1 + 1],
                  u_isfixed=u_isfixed, v_isfixed=v_isfixed, quadrature_weights=g.quadrature_weights, σ=g.σ)
    p = model.params
    params_g = copy.copy(p)
    params_g.weertman_c = np.asfortranarray(p.weertman_c[i0:i1, j0:j1])
    params_g.accumulation_rate = np.asfortranarray(p.accumulation_rate[i0:i1, j0:j1])
    params_g.glen_a_ref = np.asfortranarray(np.broadcast_to(p.glen_a_ref, (g.nx, g.ny))[i0:i1, j0:j1])
    ic = InitialConditions(
        initial_thickness=f.gh.h[i0:i1, j0:j1],
        initial_grounded_fraction=f.gh.grounded_fraction[i0:i1, j0:j1],
        initial_u_veloc=f.gu.u[i0:i1 + 1, j0:j1],
        initial_v_veloc=f.gv.v[i0:i1, j0:j1 + 1],
        initial_viscosity=f.g3d.η[i0:i1, j0:j1, :],
        initial_temperature=f.g3d.θ[i0:i1, j0:j1, :],
        initial_damage=f.g3d.Φ[i0:i1, j0:j1, :])
    return Model(grid=grid_g, bed_elevation=f.gh.b[i0:i1, j0:j1], params=params_g,
                 solver_params=model.solver_params, initial_conditions=ic, melt_rate=model.melt_rate,
                 parallel_spec=parallel_spec if parallel_spec is not None else HIPSpec(model.parallel_spec.precision))


def schwarz_restrict_velocities(model_g, model, igrid=1, jgrid=1, ngridsx=1, ngridsy=1, overlap=1):
    """schwarzRestrictVelocities! (SharedMemorySpec.jl:306-331)."""
    i0, i1, j0, j1 = _bounds(model, igrid, jgrid, ngridsx, ngridsy, overlap)
    model_g.fields.gu.u[:, :] = model.fields.gu.u[i0:i1 + 1, j0:j1]
    model_g.fields.gv.v[:, :] = model.fields.gv.v[i0:i1, j0:j1 + 1]
    return model_g


def schwarz_partition_of_unity(m, n, leavei1, leaveim, leavej1, leavejn, overlapi, overlapj):
    """schwarzPartitionOfUnity (SharedMemorySpec.jl:395-407): an m x n array
    ramping from 1 inside to 0 at the edges not marked to be left at 1."""
    if not m > ((not leavei1) and overlapi) + ((not leaveim) and overlapi):
        raise ArgumentError("subdomain too small for the overlap in i")
    if not n > ((not leavej1) and overlapj) + ((not leavejn) and overlapj):
        raise ArgumentError("subdomain too small for the overlap in j")
    i = np.arange(1, m + 1, dtype=float)[:, None]
    j = np.arange(1, n + 1, dtype=float)[None, :]
    inf = math.inf
    with np.errstate(divide="ignore", invalid="ignore"):
        pi = np.minimum(1.0, np.minimum(inf if leavei1 else (i - 1) / overlapi, inf if leaveim else (m - i) / overlapi))
        pj = np.minimum(1.0, np.minimum(inf if leavej1 else (j - 1) / overlapj, inf if leavejn else (n - j) / overlapj))
    return np.asfortranarray(pi * pj * np.ones((m, n)))


def schwarz_prolong_velocities(model, model_g, igrid=1, jgrid=1, ngridsx=1, ngridsy=1, overlap=1, damping=0.0):
    """schwarzProlongVelocities! (SharedMemorySpec.jl:351-382)."""
    i0, i1, j0, j1 = _bounds(model, igrid, jgrid, ngridsx, ngridsy, overlap)
    nxg, nyg = i1 - i0, j1 - j0
    upou = schwarz_partition_of_unity(nxg + 1, nyg, igrid == 1, igrid == ngridsx, jgrid == 1, jgrid == ngridsy,
                                      2 * overlap, 2 * overlap - 1)
    vpou = schwarz_partition_of_unity(nxg, nyg + 1, igrid == 1, igrid == ngridsx, jgrid == 1, jgrid == ngridsy,
                                      2 * overlap - 1, 2 * overlap)
    model.fields.gu.u[i0:i1 + 1, j0:j1] += (1.0 - damping) * upou * model_g.fields.gu.u
    model.fields.gv.v[i0:i1, j0:j1 + 1] += (1.0 - damping) * vpou * model_g.fields.gv.v
    return model


def update_preconditioner_schwarz(model):
    """update_preconditioner!(model, ::SharedMemorySpec) (SharedMemorySpec.jl:51-68):
    fresh subdomain models from the full model's fields.  The first time,
    each subdomain's Model and device context are built on the host
    (schwarz_model); after that the contexts persist (the subdomain grids
    follow the full model's grid, which does not change) and their fields are
    rebuilt on the device from the full model's (wavi_schwarz_refresh): the
    full model's context must hold its current fields."""
    spec = model.parallel_spec
    update_preconditioner(model)  # the full domain's context
    lib = spec._lib
    ready = getattr(spec, "_subdomains_built", False)
    if ready:
        members = [spec.schwarz_model_array[ig][jg] for ig in range(spec.ngridsx) for jg in range(spec.ngridsy)]
        for mg in members:
            mg.params.dt = model.params.dt
            update_preconditioner(mg)  # keeps the subdomain's dt in step
        subs = (ctypes.c_void_p * len(members))(*[mg.parallel_spec.ctx for mg in members])
        _check(lib, spec.ctx, lib.wavi_schwarz_refresh(spec.ctx, subs, spec.ngridsx, spec.ngridsy, spec.overlap))
        return model
    for ig in range(spec.ngridsx):
        for jg in range(spec.ngridsy):
            old = spec.schwarz_model_array[ig][jg]
            keep = None
            if old is not None:
                keep = old.parallel_spec
            mg = schwarz_model(model, ig + 1, jg + 1, spec.ngridsx, spec.ngridsy, spec.overlap,
                               parallel_spec=HIPSpec(spec.precision) if keep is None else keep)
            if old is not None and not _same_grid(old.grid, mg.grid):
                keep.close()
                mg.parallel_spec = HIPSpec(spec.precision)
            spec.schwarz_model_array[ig][jg] = mg
            update_preconditioner(mg)
            push_fields(mg)
    spec._subdomains_built = True
    return model


def _same_grid(a, b):
    return (a.nx == b.nx and a.ny == b.ny and a.nσ == b.nσ and a.dx == b.dx and a.dy == b.dy and
            all(np.array_equal(getattr(a, k), getattr(b, k))
                for k in ("h_mask", "h_isfixed", "u_iszero", "v_iszero", "u_isfixed", "v_isfixed",
                          "quadrature_weights", "σ")))


def precondition_schwarz(model):
    """precondition!(model, ::SharedMemorySpec) (SharedMemorySpec.jl:76-153)
    on the device (wavi_schwarz_precondition): the residual of the start
    guess, then niterations of restriction, subdomain update_state!, damping
    and prolongation, with every field on the device.  The model's device
    context must hold its current fields.  Returns (converged, rel_resid)."""
    spec = model.parallel_spec
    lib = spec._lib
    members = [spec.schwarz_model_array[ig][jg] for ig in range(spec.ngridsx) for jg in range(spec.ngridsy)]
    n = len(members)
    subs = (ctypes.c_void_p * n)(*[mg.parallel_spec.ctx for mg in members])
    conv = ctypes.c_int(0)
    rel = ctypes.c_double(0.0)
    _check(lib, spec.ctx, lib.wavi_schwarz_precondition(spec.ctx, subs, spec.ngridsx, spec.ngridsy, spec.overlap,
                                                        spec.niterations, spec.damping, spec.nthreads,
                                                        ctypes.byref(conv), ctypes.byref(rel)))
    return bool(conv.value), rel.value


def _velocities_on_device(model, verbose=False):
    """update_velocities! with the RAS preconditioner on fields that are
    already on the full model's device context (and stay there): the
    subdomain models from those fields (the first time from the host fields,
    which must then be current), then the Picard loop of inner_update! and
    precondition!(model, ::SharedMemorySpec)."""
    spec = model.parallel_spec
    if not getattr(spec, "_subdomains_built", False):
        pull_fields(model)
    update_preconditioner_schwarz(model)
    converged, it, rel = False, 0, math.inf
    while not converged and it < model.solver_params.maxiter_picard:
        it += 1
        _check(spec._lib, spec.ctx, spec._lib.wavi_inner_update(spec.ctx))
        converged, rel = precondition_schwarz(model)
    rep = _lib.SolveReport()
    rep.picard_iters, rep.converged, rep.rel_resid = it, int(converged), rel
    _report_solve(model, rep, verbose)
    return model


def update_velocities_schwarz(model, verbose=False):
    """update_velocities! (update_velocities.jl:11-29) with the RAS
    preconditioner: update_preconditioner! builds the subdomain models from
    the model's fields, then the Picard loop of inner_update! and
    precondition!(model, ::SharedMemorySpec) runs on the device, the fields
    going to the device once and coming back once.  The subdomain Models'
    host fields are those update_preconditioner! gave them the first time;
    their device contexts hold the subdomain solves."""
    update_preconditioner(model)
    push_fields(model)
    _velocities_on_device(model, verbose)
    pull_fields(model)
    return model


def state_steps_on_device(model, nsteps, dt, step_thickness=True, verbose=False):
    """nsteps of timestep! (run_simulation.jl:6-17) of a model whose
    parallel_spec is a SharedMemorySpec, on fields already on its device
    context: update_state! (geometry, the RAS velocity solve, the steps after
    the solve) and, with step_thickness, update_thickness!, each step on the
    device."""
    spec = model.parallel_spec
    lib = spec._lib
    for _ in range(nsteps):
        _check(lib, spec.ctx, lib.wavi_update_geometry(spec.ctx))
        _velocities_on_device(model, verbose)
        _check(lib, spec.ctx, lib.wavi_after_solve(spec.ctx))
        if step_thickness:
            _check(lib, spec.ctx, lib.wavi_update_thickness(spec.ctx, dt))
    return model


def update_state_schwarz(model, verbose=False):
    """update_state! (update_state.jl:7-21) of a model whose parallel_spec is
    a SharedMemorySpec: geometry and the steps after the solve on the full
    domain's context, the velocity solve by RAS, with the fields on the
    device throughout."""
    update_preconditioner(model)
    push_fields(model)
    state_steps_on_device(model, 1, model.params.dt, step_thickness=False, verbose=verbose)
    pull_fields(model)
    return model
