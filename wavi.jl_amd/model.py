"""Model, Fields and the velocity-solve entry points, mirroring WAVI.jl.

`Model(; grid, bed_elevation, params, solver_params, initial_conditions,
melt_rate, parallel_spec)` follows src/Models/Model.jl:34-93 and
setup_fields (src/Fields/Fields.jl:26-115).  The only parallel spec is
`HIPSpec`, the MI355X drop-in for `BasicParallelSpec`: update_velocities,
inner_update, update_state and update_wavelets run in libwavi_hip.so on the
GPU.  Host arrays in `model.fields` are the reference's fields; they are
copied to the device before a call and back after it.
"""
import ctypes

import numpy as np

from . import _lib
from .core import (ArgumentError, DimensionMismatch, Grid, InitialConditions, Params,
                   SolverParams, UniformMeltRate, get_c_mask, get_u_mask, get_v_mask, glen_b)


class _Namespace:
    def __repr__(self):
        return "<%s %s>" % (type(self).__name__, ", ".join(sorted(vars(self))))


class HGrid(_Namespace):
    pass


class UGrid(_Namespace):
    pass


class VGrid(_Namespace):
    pass


class CGrid(_Namespace):
    pass


class SigmaGrid(_Namespace):
    pass


class Wavelets(_Namespace):
    pass


class Fields(_Namespace):
    """Fields{gh, gu, gv, gc, g3d, wu, wv} (src/Fields/Fields.jl:10-18)."""


# (container, attribute, library field id); masks travel as 0.0/1.0
FIELD_MAP = [
    ("gh", "b", "H_B"), ("gh", "h", "H_H"), ("gh", "s", "H_S"), ("gh", "dhdt", "H_DHDT"),
    ("gh", "accumulation", "H_ACCUMULATION"), ("gh", "basal_melt", "H_BASAL_MELT"),
    ("gh", "haf", "H_HAF"), ("gh", "grounded_fraction", "H_GROUNDED_FRACTION"),
    ("gh", "dsdh", "H_DSDH"), ("gh", "shelf_strain_rate", "H_SHELF_STRAIN_RATE"),
    ("gh", "av_speed", "H_AV_SPEED"), ("gh", "u", "H_U"), ("gh", "v", "H_V"),
    ("gh", "us", "H_US"), ("gh", "vs", "H_VS"), ("gh", "ub", "H_UB"), ("gh", "vb", "H_VB"),
    ("gh", "bed_speed", "H_BED_SPEED"), ("gh", "weertman_c", "H_WEERTMAN_C"),
    ("gh", "β", "H_BETA"), ("gh", "βeff", "H_BETA_EFF"), ("gh", "τbed", "H_TAU_BED"),
    ("gh", "ηav", "H_ETA_AV"), ("gh", "quad_f1", "H_QUAD_F1"), ("gh", "quad_f2", "H_QUAD_F2"),
    ("params", "weertman_c", "P_WEERTMAN_C"), ("params", "accumulation_rate", "P_ACCUMULATION_RATE"),
    ("gu", "s", "U_S"), ("gu", "h", "U_H"), ("gu", "grounded_fraction", "U_GROUNDED_FRACTION"),
    ("gu", "βeff", "U_BETA_EFF"), ("gu", "u", "U_U"), ("gu", "residual", "U_RESIDUAL"),
    ("gv", "s", "V_S"), ("gv", "h", "V_H"), ("gv", "grounded_fraction", "V_GROUNDED_FRACTION"),
    ("gv", "βeff", "V_BETA_EFF"), ("gv", "v", "V_V"), ("gv", "residual", "V_RESIDUAL"),
    ("g3d", "η", "S_ETA"), ("g3d", "glen_b", "S_GLEN_B"),
    ("wu", "wavelets", "WU_WAVELETS"), ("wu", "mask", "WU_MASK"),
    ("wu", "correction_coarse", "WU_CORRECTION_COARSE"),
    ("wv", "wavelets", "WV_WAVELETS"), ("wv", "mask", "WV_MASK"),
    ("wv", "correction_coarse", "WV_CORRECTION_COARSE"),
]


def _f64(shape):
    return np.zeros(shape, dtype=np.float64, order="F")


def setup_fields(grid, initial_conditions, solver_params, params, bed_array):
    """setup_fields (src/Fields/Fields.jl:26-115)."""
    nx, ny, ns = grid.nx, grid.ny, grid.nσ
    h_mask = grid.h_mask
    u_mask = get_u_mask(h_mask)
    v_mask = get_v_mask(h_mask)
    c_mask = get_c_mask(h_mask)
    u_mask[grid.u_iszero] = False
    v_mask[grid.v_iszero] = False
    f = Fields()
    gh = HGrid()
    gh.nxh, gh.nyh, gh.mask, gh.h_isfixed = nx, ny, h_mask, grid.h_isfixed
    gh.n = int(h_mask.sum())
    gh.b = np.asfortranarray(bed_array, dtype=float)
    gh.h = np.asfortranarray(initial_conditions.initial_thickness, dtype=float).copy(order="F")
    gh.grounded_fraction = np.asfortranarray(initial_conditions.initial_grounded_fraction, dtype=float).copy(order="F")
    gh.ηav = np.asfortranarray(initial_conditions.initial_viscosity[:, :, 0], dtype=float).copy(order="F")
    for name in ("s", "dhdt", "accumulation", "basal_melt", "haf", "shelf_strain_rate", "av_speed",
                 "u", "v", "us", "vs", "ub", "vb", "bed_speed", "weertman_c", "β", "βeff", "τbed",
                 "quad_f1", "quad_f2"):
        setattr(gh, name, _f64((nx, ny)))
    gh.dsdh = np.ones((nx, ny), order="F")
    gh.quad_f2[h_mask] = gh.h[h_mask] / (3 * gh.ηav[h_mask])  # HGrid.jl:229
    gu = UGrid()
    gu.nxu, gu.nyu, gu.mask, gu.u_isfixed = nx + 1, ny, u_mask, grid.u_isfixed
    gu.mask_inner = u_mask & ~grid.u_isfixed
    gu.n, gu.ni = int(u_mask.sum()), int(gu.mask_inner.sum())
    gu.levels = solver_params.levels
    gu.u = np.asfortranarray(initial_conditions.initial_u_veloc, dtype=float).copy(order="F")
    for name in ("s", "h", "βeff", "residual"):
        setattr(gu, name, _f64((nx + 1, ny)))
    gu.grounded_fraction = np.ones((nx + 1, ny), order="F")
    gv = VGrid()
    gv.nxv, gv.nyv, gv.mask, gv.v_isfixed = nx, ny + 1, v_mask, grid.v_isfixed
    gv.mask_inner = v_mask & ~grid.v_isfixed
    gv.n, gv.ni = int(v_mask.sum()), int(gv.mask_inner.sum())
    gv.levels = solver_params.levels
    gv.v = np.asfortranarray(initial_conditions.initial_v_veloc, dtype=float).copy(order="F")
    for name in ("s", "h", "βeff", "residual"):
        setattr(gv, name, _f64((nx, ny + 1)))
    gv.grounded_fraction = np.ones((nx, ny + 1), order="F")
    gc = CGrid()
    gc.nxc, gc.nyc, gc.mask = nx - 1, ny - 1, c_mask
    gc.n = int(c_mask.sum())
    g3d = SigmaGrid()
    g3d.nxs, g3d.nys, g3d.nσs = nx, ny, ns
    g3d.σ, g3d.ζ, g3d.quadrature_weights = grid.σ, grid.ζ, grid.quadrature_weights
    g3d.η = np.asfortranarray(initial_conditions.initial_viscosity, dtype=float).copy(order="F")
    g3d.θ = np.asfortranarray(initial_conditions.initial_temperature, dtype=float)
    g3d.Φ = np.asfortranarray(initial_conditions.initial_damage, dtype=float)
    glen_a_ref = np.broadcast_to(np.asarray(params.glen_a_ref, dtype=float), (nx, ny))
    g3d.glen_b = np.asfortranarray(glen_b(g3d.θ, g3d.Φ, glen_a_ref[:, :, None], params.glen_n,
                                          params.glen_a_activation_energy,
                                          params.glen_temperature_ref, params.gas_const))
    wu = Wavelets()
    wu.nxuw, wu.nyuw, wu.levels = nx + 1, ny, solver_params.levels
    wu.mask = np.ones((nx + 1, ny), bool, order="F")
    wu.wavelets = _f64((nx + 1, ny))
    wu.correction_coarse = _f64((nx + 1, ny))
    wv = Wavelets()
    wv.nxvw, wv.nyvw, wv.levels = nx, ny + 1, solver_params.levels
    wv.mask = np.ones((nx, ny + 1), bool, order="F")
    wv.wavelets = _f64((nx, ny + 1))
    wv.correction_coarse = _f64((nx, ny + 1))
    f.gh, f.gu, f.gv, f.gc, f.g3d, f.wu, f.wv = gh, gu, gv, gc, g3d, wu, wv
    return f


def check_initial_conditions(ic, params, grid):
    """src/Models/model_utilities.jl:7-60."""
    nx, ny, ns = grid.nx, grid.ny, grid.nσ
    if np.all(np.isnan(ic.initial_thickness)):
        ic.initial_thickness = params.default_thickness * np.ones((nx, ny))
    if np.all(np.isnan(ic.initial_grounded_fraction)):
        ic.initial_grounded_fraction = np.ones((nx, ny))
    if np.all(np.isnan(ic.initial_u_veloc)):
        ic.initial_u_veloc = np.zeros((nx + 1, ny))
    if np.all(np.isnan(ic.initial_v_veloc)):
        ic.initial_v_veloc = np.zeros((nx, ny + 1))
    if np.all(np.isnan(ic.initial_viscosity)):
        ic.initial_viscosity = params.default_viscosity * np.ones((nx, ny, ns))
    if np.all(np.isnan(ic.initial_temperature)):
        ic.initial_temperature = params.default_temperature * np.ones((nx, ny, ns))
    if np.all(np.isnan(ic.initial_damage)):
        ic.initial_damage = params.default_damage * np.ones((nx, ny, ns))
    checks = [("Initial thickness", ic.initial_thickness, (nx, ny)),
              ("Initial grounded fraction", ic.initial_grounded_fraction, (nx, ny)),
              ("Initial u_velocity", ic.initial_u_veloc, (nx + 1, ny)),
              ("Initial v_velocity", ic.initial_v_veloc, (nx, ny + 1)),
              ("Initial temperature", ic.initial_temperature, (nx, ny, ns)),
              ("Initial viscosity", ic.initial_viscosity, (nx, ny, ns)),
              ("Initial damage", ic.initial_damage, (nx, ny, ns))]
    for name, a, shape in checks:
        if a.shape != shape:
            raise DimensionMismatch("%s field is not compatible with grid size: %s, must be %s" % (name, a.shape, shape))
    return ic


class HIPSpec:
    """The MI355X parallel spec: the counterpart of a new
    `AbstractParallelSpec` subtype dispatched on by update_velocities!.
    precision: "fp64" (the reference scheme) or "mixed" (FP32 V-cycle inside
    FP64 residual refinement)."""

    def __init__(self, precision="fp64"):
        self.ctx = None
        self._lib = None
        if precision not in _lib.PRECISION:
            raise ArgumentError("precision must be one of %s" % list(_lib.PRECISION))
        self.precision = precision
        self.ctx = None
        self._lib = None
        self._dt = None

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "ctx", None) is not None and getattr(self, "_lib", None) is not None:
            self._lib.wavi_ctx_destroy(self.ctx)
            self.ctx = None


BasicParallelSpec = HIPSpec


class Model:
    """Model(; grid, bed_elevation, params, solver_params, initial_conditions,
    melt_rate, parallel_spec) (src/Models/Model.jl:34-93)."""

    def __init__(self, grid=None, bed_elevation=None, params=None, solver_params=None,
                 initial_conditions=None, melt_rate=None, parallel_spec=None):
        if grid is None:
            raise ArgumentError("You must specify an input grid")
        if bed_elevation is None:
            raise ArgumentError("You must input a bed elevation")
        params = Params() if params is None else params
        solver_params = SolverParams() if solver_params is None else solver_params
        initial_conditions = InitialConditions() if initial_conditions is None else initial_conditions
        melt_rate = UniformMeltRate() if melt_rate is None else melt_rate
        parallel_spec = HIPSpec() if parallel_spec is None else parallel_spec
        if not isinstance(melt_rate, UniformMeltRate):
            raise ArgumentError("only UniformMeltRate is supported by the device state update")
        if callable(bed_elevation):
            bed = np.asfortranarray(bed_elevation(grid.xxh, grid.yyh), dtype=float)
        else:
            try:
                bed = np.asfortranarray(bed_elevation, dtype=float)
            except (TypeError, ValueError):
                raise ArgumentError("bed elevation must be of type function or array")
        initial_conditions = check_initial_conditions(initial_conditions, params, grid)
        for name in ("weertman_c", "accumulation_rate", "glen_a_ref"):
            v = getattr(params, name)
            if np.isscalar(v):
                setattr(params, name, v * np.ones((grid.nx, grid.ny), order="F"))
            else:
                v = np.asfortranarray(v, dtype=float)
                if v.shape != (grid.nx, grid.ny):
                    raise DimensionMismatch("Size of input %s must match grid size (i.e. %d x %d)" % (name, grid.nx, grid.ny))
                setattr(params, name, v)
        self.grid = grid
        self.params = params
        self.solver_params = solver_params
        self.initial_conditions = initial_conditions
        self.melt_rate = melt_rate
        self.parallel_spec = parallel_spec
        self.fields = setup_fields(grid, initial_conditions, solver_params, params, bed)
        self.last_report = None

    # -- field plumbing -------------------------------------------------
    def _array(self, container, attr):
        obj = self.params if container == "params" else getattr(self.fields, container)
        return getattr(obj, attr)

    def _set_array(self, container, attr, value):
        obj = self.params if container == "params" else getattr(self.fields, container)
        old = getattr(obj, attr)
        if old.dtype == np.bool_:
            value = value > 0.5
        setattr(obj, attr, np.asfortranarray(value.reshape(old.shape, order="F"), dtype=old.dtype))


def _check(lib, ctx, rc):
    if rc == 0:
        return
    msg = lib.wavi_last_error(ctx)
    msg = msg.decode() if msg else "error %d" % rc
    if rc == _lib.WAVI_EDIM:
        raise DimensionMismatch(msg)
    if rc == _lib.WAVI_EARG:
        raise ArgumentError(msg)
    raise RuntimeError("libwavi_hip: %s (code %d)" % (msg, rc))


def grid_desc(model):
    """Build the C descriptor structs; returns (keepalive, desc, phys, solver)."""
    g, p, s = model.grid, model.params, model.solver_params
    keep = []

    def u8(a):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.uint8).ravel(order="F"))
        keep.append(a)
        return _lib.u8ptr(a)

    def f8(a):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))
        keep.append(a)
        return _lib.dptr(a)

    desc = _lib.GridDesc(g.nx, g.ny, g.nσ, g.dx, g.dy, u8(g.h_mask), u8(g.h_isfixed),
                         u8(g.u_iszero), u8(g.v_iszero), u8(g.u_isfixed), u8(g.v_isfixed),
                         f8(g.σ), f8(g.quadrature_weights))
    phys = _lib.PhysParams(p.dt, p.g, p.density_ice, p.density_ocean, p.glen_n,
                           p.glen_reg_strain_rate, p.weertman_m, p.weertman_reg_speed,
                           p.sea_level_wrt_geoid, p.minimum_thickness, model.melt_rate.m,
                           1 if p.evolveShelves else 0)
    solver = _lib.SolverParamsC(s.n_iter_viscosity, s.tol_picard, s.maxiter_picard, s.tol_coarse,
                                s.maxiter_coarse, s.levels, s.wavelet_threshold, s.nsmooth,
                                s.smoother_omega, s.stencil_margin, s.super_implicitness)
    return keep, desc, phys, solver


def update_preconditioner(model):
    """update_preconditioner!(model, ::HIPSpec) (dispatch hook,
    src/SharedMemorySpec.jl:31): create the device context on first use and
    keep its dt in step with params.dt."""
    spec = model.parallel_spec
    if not isinstance(spec, HIPSpec):
        raise ArgumentError("parallel_spec must be HIPSpec")
    lib = _lib.load()
    if spec.ctx is None:
        if not model.params.evolveShelves:
            raise ArgumentError("evolveShelves=false is not supported")
        keep, desc, phys, solver = grid_desc(model)
        out = ctypes.c_void_p()
        rc = lib.wavi_ctx_create(ctypes.byref(desc), ctypes.byref(phys), ctypes.byref(solver),
                                 _lib.PRECISION[spec.precision], ctypes.byref(out))
        if rc != 0:
            msg = lib.wavi_last_error(None)
            msg = msg.decode() if msg else "error %d" % rc
            if rc == _lib.WAVI_EDIM:
                raise DimensionMismatch(msg)
            if rc == _lib.WAVI_EARG:
                raise ArgumentError(msg)
            raise RuntimeError("wavi_ctx_create: %s (code %d)" % (msg, rc))
        spec.ctx = out.value
        spec._lib = lib
        spec._dt = model.params.dt
    elif spec._dt != model.params.dt:
        _check(lib, spec.ctx, lib.wavi_set_dt(spec.ctx, model.params.dt))
        spec._dt = model.params.dt
    return model


def push_fields(model):
    """Copy every field of model.fields to the device."""
    spec = model.parallel_spec
    lib = spec._lib
    for cont, attr, fid in FIELD_MAP:
        a = model._array(cont, attr)
        buf = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))
        _check(lib, spec.ctx, lib.wavi_set_field(spec.ctx, _lib.FIELD[fid], _lib.dptr(buf), buf.size))


def pull_fields(model):
    """Copy every field back from the device into model.fields."""
    spec = model.parallel_spec
    lib = spec._lib
    for cont, attr, fid in FIELD_MAP:
        a = model._array(cont, attr)
        buf = np.empty(a.size, dtype=np.float64)
        _check(lib, spec.ctx, lib.wavi_get_field(spec.ctx, _lib.FIELD[fid], _lib.dptr(buf), buf.size))
        model._set_array(cont, attr, buf)


def _report_solve(model, rep, verbose):
    model.last_report = rep.as_dict()
    if verbose:
        print("Solved momentum equation with residual %.3g at iteration %d"
              % (rep.rel_resid, rep.picard_iters))


def _schwarz(model):
    return getattr(model.parallel_spec, "schwarz_model_array", None) is not None


def update_velocities(model, verbose=False):
    """update_velocities!(model) (src/Models/update_velocities.jl:11-29) on
    the GPU: Picard iteration of inner_update! and one V-cycle per
    iteration (with a SharedMemorySpec, the restricted additive Schwarz
    preconditioner instead of the V-cycle)."""
    if _schwarz(model):
        from .schwarz import update_velocities_schwarz
        return update_velocities_schwarz(model, verbose)
    update_preconditioner(model)
    spec = model.parallel_spec
    push_fields(model)
    rep = _lib.SolveReport()
    _check(spec._lib, spec.ctx, spec._lib.wavi_update_velocities(spec.ctx, ctypes.byref(rep)))
    pull_fields(model)
    _report_solve(model, rep, verbose)
    return model


def update_velocities_batch(models, nthreads=8, verbose=False, lockstep=False):
    """update_velocities! on every model of an ensemble in one native call:
    host threads drive the members concurrently, each on its own stream, so
    small members share the GPU.  With lockstep=True (mixed-precision models
    with one Haar level count) the calling thread drives all members through
    each Picard iteration together instead, and their coarse CGs share one
    launch per CG iteration (wavi_update_velocities_lockstep).  Either way
    equivalent, member by member, to calling update_velocities on each model
    in turn."""
    models = list(models)
    if len({id(m) for m in models}) != len(models):
        raise ArgumentError("update_velocities_batch: a model appears twice in the batch")
    for m in models:
        update_preconditioner(m)
        push_fields(m)
    if not models:
        return models
    lib = models[0].parallel_spec._lib
    ctxs = (ctypes.c_void_p * len(models))(*[m.parallel_spec.ctx for m in models])
    reps = (_lib.SolveReport * len(models))()
    codes = (ctypes.c_int * len(models))()
    if lockstep:
        rc = lib.wavi_update_velocities_lockstep(ctxs, len(models), reps, codes)
    else:
        rc = lib.wavi_update_velocities_batch(ctxs, len(models), int(nthreads), reps, codes)
    if rc != _lib.WAVI_OK:
        for m, code in zip(models, codes):
            _check(lib, m.parallel_spec.ctx, code)
        _check(lib, None, rc)
    for m, rep in zip(models, reps):
        pull_fields(m)
        _report_solve(m, rep, verbose)
    return models


def update_geometry(model):
    """The update_state! steps that precede the velocity solve (surface,
    geometry on u/v, height above floatation, grounded fractions,
    weertman_c = C * grounded fraction, dsdh; update_state.jl:8-15)."""
    update_preconditioner(model)
    spec = model.parallel_spec
    push_fields(model)
    _check(spec._lib, spec.ctx, spec._lib.wavi_update_geometry(spec.ctx))
    pull_fields(model)
    return model


def misfit_gradient(model, n_picard, u_obs, v_obs, tol_linear=1e-13, maxit_linear=60, gradient=True):
    """Adjoint basal-drag gradient (no reference counterpart; see
    wavi_misfit_gradient in include/wavi_hip.h): J = 1/2 ||u - u_obs||^2 over
    inner u and v points after n_picard Picard iterations with converged
    linear solves, and dJ/d log C on the h grid.  The model's fields must
    already hold the state after update_geometry (as update_velocities
    expects).  Returns (J, grad (nx, ny) or None, report dict); the model is
    left in the forward's last Picard state."""
    update_preconditioner(model)
    spec = model.parallel_spec
    lib = spec._lib
    push_fields(model)
    nx, ny = model.grid.nx, model.grid.ny
    uo = np.ascontiguousarray(np.asarray(u_obs, dtype=np.float64).ravel(order="F"))
    vo = np.ascontiguousarray(np.asarray(v_obs, dtype=np.float64).ravel(order="F"))
    if uo.size != (nx + 1) * ny or vo.size != nx * (ny + 1):
        raise DimensionMismatch("u_obs must have (nx+1)*ny values and v_obs nx*(ny+1)")
    J = ctypes.c_double(0.0)
    g = np.zeros(nx * ny) if gradient else None
    rep = _lib.AdjointReport()
    _check(lib, spec.ctx, lib.wavi_misfit_gradient(spec.ctx, int(n_picard), float(tol_linear), int(maxit_linear),
                                                   _lib.dptr(uo), _lib.dptr(vo), ctypes.byref(J),
                                                   _lib.dptr(g) if gradient else None, ctypes.byref(rep)))
    pull_fields(model)
    return J.value, (g.reshape((nx, ny), order="F") if gradient else None), rep.as_dict()


def misfit_gradient_batch(models, n_picard, observations, tol_linear=1e-13, maxit_linear=60, nthreads=8,
                          gradient=True):
    """misfit_gradient on every model of an ensemble in one native call
    (host threads, one stream per model).  observations: one (u_obs, v_obs)
    pair per model.  Returns a list of (J, grad or None, report dict)."""
    models = list(models)
    n = len(models)
    if n == 0:
        return []
    if len({id(m) for m in models}) != n:
        raise ArgumentError("misfit_gradient_batch: a model appears twice in the batch")
    obs = []
    for m, (uo, vo) in zip(models, observations):
        update_preconditioner(m)
        push_fields(m)
        nx, ny = m.grid.nx, m.grid.ny
        a = np.ascontiguousarray(np.asarray(uo, dtype=np.float64).ravel(order="F"))
        b = np.ascontiguousarray(np.asarray(vo, dtype=np.float64).ravel(order="F"))
        if a.size != (nx + 1) * ny or b.size != nx * (ny + 1):
            raise DimensionMismatch("u_obs must have (nx+1)*ny values and v_obs nx*(ny+1)")
        obs.append((a, b))
    if len(obs) != n:
        raise ArgumentError("one (u_obs, v_obs) pair per model is required")
    lib = models[0].parallel_spec._lib
    ctxs = (ctypes.c_void_p * n)(*[m.parallel_spec.ctx for m in models])
    up = (_lib._dp * n)(*[_lib.dptr(a) for a, _ in obs])
    vp = (_lib._dp * n)(*[_lib.dptr(b) for _, b in obs])
    J = np.zeros(n)
    grads = [np.zeros(m.grid.nx * m.grid.ny) for m in models] if gradient else None
    gp = (_lib._dp * n)(*[_lib.dptr(g) for g in grads]) if gradient else None
    reps = (_lib.AdjointReport * n)()
    codes = (ctypes.c_int * n)()
    rc = lib.wavi_misfit_gradient_batch(ctxs, n, int(nthreads), int(n_picard), float(tol_linear), int(maxit_linear),
                                        up, vp, _lib.dptr(J), gp, reps, codes)
    if rc != _lib.WAVI_OK:
        for m, code in zip(models, codes):
            _check(lib, m.parallel_spec.ctx, code)
        _check(lib, None, rc)
    out = []
    for k, m in enumerate(models):
        pull_fields(m)
        g = grads[k].reshape((m.grid.nx, m.grid.ny), order="F") if gradient else None
        out.append((float(J[k]), g, reps[k].as_dict()))
    return out


def inner_update(model):
    """inner_update!(model) (update_velocities.jl:32-45)."""
    update_preconditioner(model)
    spec = model.parallel_spec
    push_fields(model)
    _check(spec._lib, spec.ctx, spec._lib.wavi_inner_update(spec.ctx))
    pull_fields(model)
    return model


def update_state(model, clock=None, verbose=False):
    """update_state!(model[, clock]) (src/Models/update_state.jl:7-42)."""
    if _schwarz(model):
        from .schwarz import update_state_schwarz
        return update_state_schwarz(model, verbose)
    update_preconditioner(model)
    spec = model.parallel_spec
    push_fields(model)
    rep = _lib.SolveReport()
    _check(spec._lib, spec.ctx, spec._lib.wavi_update_state(spec.ctx, ctypes.byref(rep)))
    pull_fields(model)
    _report_solve(model, rep, verbose)
    return model


def update_wavelets(model):
    """update_wavelets!(model) (src/Wavelets/update_wavelets.jl:6-32)."""
    update_preconditioner(model)
    spec = model.parallel_spec
    push_fields(model)
    _check(spec._lib, spec.ctx, spec._lib.wavi_update_wavelets(spec.ctx))
    pull_fields(model)
    return model


# ---- single pieces of the solve on host arrays (test hooks) ----------

def _fullvec(model, a, which):
    shape = (model.grid.nx + 1, model.grid.ny) if which == "u" else (model.grid.nx, model.grid.ny + 1)
    a = np.asarray(a, dtype=np.float64)
    if a.shape != shape:
        raise DimensionMismatch("expected shape %s, got %s" % (shape, a.shape))
    return np.ascontiguousarray(a.ravel(order="F"))


def _shape_back(model, buf, which):
    shape = (model.grid.nx + 1, model.grid.ny) if which == "u" else (model.grid.nx, model.grid.ny + 1)
    return buf.reshape(shape, order="F")


def op_apply(model, xu, xv, sampled=True):
    """op_fun! (src/utilities.jl:61-145) with the model's current coefficients."""
    spec = model.parallel_spec
    bu, bv = _fullvec(model, xu, "u"), _fullvec(model, xv, "v")
    yu, yv = np.empty_like(bu), np.empty_like(bv)
    _check(spec._lib, spec.ctx, spec._lib.wavi_op_apply(spec.ctx, _lib.dptr(bu), _lib.dptr(bv),
                                                        _lib.dptr(yu), _lib.dptr(yv), 1 if sampled else 0))
    return _shape_back(model, yu, "u"), _shape_back(model, yv, "v")


def get_rhs(model):
    """get_rhs (update_velocities.jl:106-182) in full layout."""
    spec = model.parallel_spec
    bu = np.empty((model.grid.nx + 1) * model.grid.ny)
    bv = np.empty(model.grid.nx * (model.grid.ny + 1))
    _check(spec._lib, spec.ctx, spec._lib.wavi_rhs(spec.ctx, _lib.dptr(bu), _lib.dptr(bv)))
    return _shape_back(model, bu, "u"), _shape_back(model, bv, "v")


def get_op_diag(model):
    """get_op_diag (src/Wavelets/update_preconditioners.jl:92-109)."""
    spec = model.parallel_spec
    du = np.empty((model.grid.nx + 1) * model.grid.ny)
    dv = np.empty(model.grid.nx * (model.grid.ny + 1))
    _check(spec._lib, spec.ctx, spec._lib.wavi_op_diag(spec.ctx, _lib.dptr(du), _lib.dptr(dv)))
    return _shape_back(model, du, "u"), _shape_back(model, dv, "v")


def vcycle(model, ru, rv):
    """One V-cycle precondition!(x, p, b) (update_preconditioners.jl:54-83)."""
    spec = model.parallel_spec
    bu, bv = _fullvec(model, ru, "u"), _fullvec(model, rv, "v")
    eu, ev = np.empty_like(bu), np.empty_like(bv)
    it = ctypes.c_int32(0)
    _check(spec._lib, spec.ctx, spec._lib.wavi_vcycle(spec.ctx, _lib.dptr(bu), _lib.dptr(bv),
                                                      _lib.dptr(eu), _lib.dptr(ev), ctypes.byref(it)))
    return _shape_back(model, eu, "u"), _shape_back(model, ev, "v"), it.value


def haar(model, comp, direction, x):
    """2-D Haar transform: direction "forward" (dwt), "inverse" (idwt) or
    "inverse_t" (idwt^T), comp "u" or "v"."""
    spec = model.parallel_spec
    d = {"forward": 0, "inverse": 1, "inverse_t": 2}[direction]
    b = _fullvec(model, x, comp)
    out = np.empty_like(b)
    _check(spec._lib, spec.ctx, spec._lib.wavi_haar(spec.ctx, 0 if comp == "u" else 1, d,
                                                    _lib.dptr(b), _lib.dptr(out)))
    return _shape_back(model, out, comp)
