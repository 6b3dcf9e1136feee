"""
    WAVIHIP

MI355X velocity solve for WAVI.jl models.  `HIPSpec` is a new
`AbstractParallelSpec` (WAVI.jl `src/WAVI.jl:17`); models built with
`parallel_spec = HIPSpec()` dispatch `update_velocities!` and
`update_state!` to `libwavi_hip.so` through `ccall`.  The fields stay in the
model's `Fields` struct.  `update_velocities!` moves only the fields the
Picard loop reads to the device and only those it writes back;
`run_simulation!` keeps the whole state resident on the device between the
steps at which the reference writes something (checkpoints, outputs, the
velocity dump), where it brings the fields back and writes the reference's
own JLD2 checkpoints with `@save`.

This file is written against the C ABI in `include/wavi_hip.h`.  It has not
been run: the build image has no Julia.
"""
module WAVIHIP

using WAVI
using ChainRulesCore
using JLD2
using SparseArrays
using LinearAlgebra
import WAVI: update_velocities!, update_state!, AbstractParallelSpec, AbstractModel

const LIB = get(ENV, "WAVI_HIP_LIB", joinpath(@__DIR__, "..", "lib", "libwavi_hip.so"))

# --- C structs (layout of include/wavi_hip.h) ------------------------------
struct GridDesc
    nx::Int32
    ny::Int32
    nsigma::Int32
    dx::Float64
    dy::Float64
    h_mask::Ptr{UInt8}
    h_isfixed::Ptr{UInt8}
    u_iszero::Ptr{UInt8}
    v_iszero::Ptr{UInt8}
    u_isfixed::Ptr{UInt8}
    v_isfixed::Ptr{UInt8}
    sigma::Ptr{Float64}
    quadrature_weights::Ptr{Float64}
end

struct PhysParams
    dt::Float64
    g::Float64
    density_ice::Float64
    density_ocean::Float64
    glen_n::Float64
    glen_reg_strain_rate::Float64
    weertman_m::Float64
    weertman_reg_speed::Float64
    sea_level_wrt_geoid::Float64
    minimum_thickness::Float64
    uniform_melt_rate::Float64
    evolve_shelves::Int32
end

struct SolverParamsC
    n_iter_viscosity::Int32
    tol_picard::Float64
    maxiter_picard::Int32
    tol_coarse::Float64
    maxiter_coarse::Int32
    levels::Int32
    wavelet_threshold::Float64
    nsmooth::Int32
    smoother_omega::Float64
    stencil_margin::Int32
    super_implicitness::Float64
end

struct SolveReport
    picard_iters::Int32
    converged::Int32
    rel_resid::Float64
    coarse_cg_total::Int32
    refinements::Int32
    coarse_cg_iters::NTuple{64,Int32}
    rel_resid_hist::NTuple{64,Float64}
end

struct AdjointReport
    n_picard::Int32
    fwd_vcycles::Int32
    adj_vcycles::Int32
    unconverged::Int32
    max_fwd_rel::Float64
    max_adj_rel::Float64
    coarse_cg_iters::Int64
end

const PRECISION = Dict(:fp64 => Int32(0), :mixed => Int32(1))

"""
    HIPSpec(; precision = :fp64)

Parallel spec that runs the velocity solve on the GPU.  `precision = :mixed`
runs the V-cycle in FP32 inside FP64 residual refinement.
"""
mutable struct HIPSpec <: AbstractParallelSpec
    precision::Int32
    ctx::Ptr{Cvoid}
    dt::Float64
    function HIPSpec(; precision::Symbol = :fp64)
        haskey(PRECISION, precision) || throw(ArgumentError("precision must be :fp64 or :mixed"))
        spec = new(PRECISION[precision], C_NULL, NaN)
        finalizer(spec) do s
            s.ctx == C_NULL || ccall((:wavi_ctx_destroy, LIB), Cint, (Ptr{Cvoid},), s.ctx)
            s.ctx = C_NULL
        end
        return spec
    end
end

function check(ctx, rc)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:wavi_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx))
    rc == -2 && throw(DimensionMismatch(msg))
    rc == -1 && throw(ArgumentError(msg))
    error("libwavi_hip: $msg (code $rc)")
end

# (field id, accessor) in the order of enum wavi_field
fieldlist(m) = begin
    f = m.fields; gh, gu, gv, g3d, wu, wv = f.gh, f.gu, f.gv, f.g3d, f.wu, f.wv
    Any[gh.b, gh.h, gh.s, gh.dhdt, gh.accumulation, gh.basal_melt, gh.haf, gh.grounded_fraction,
        gh.dsdh, gh.shelf_strain_rate, gh.av_speed, gh.u, gh.v, gh.us, gh.vs, gh.ub, gh.vb,
        gh.bed_speed, gh.weertman_c, gh.β, gh.βeff, gh.τbed, gh.ηav, gh.quad_f1, gh.quad_f2,
        m.params.weertman_c, m.params.accumulation_rate,
        gu.s, gu.h, gu.grounded_fraction, gu.βeff, gu.u, gu.residual,
        gv.s, gv.h, gv.grounded_fraction, gv.βeff, gv.v, gv.residual,
        g3d.η, g3d.glen_b]
end

"""
    update_preconditioner!(model)

Dispatch hook (`src/SharedMemorySpec.jl:31`): creates the device context on
first use and keeps its dt in step with `model.params.dt`.
"""
function WAVI.update_preconditioner!(model::AbstractModel{T,N,M,<:HIPSpec}) where {T,N,M}
    spec = model.parallel_spec
    if spec.ctx == C_NULL
        g, p, s = model.grid, model.params, model.solver_params
        σ = collect(Float64, g.σ); w = collect(Float64, g.quadrature_weights)
        masks = map(x -> collect(UInt8, x), (g.h_mask, g.h_isfixed, g.u_iszero, g.v_iszero, g.u_isfixed, g.v_isfixed))
        GC.@preserve σ w masks begin
            desc = GridDesc(g.nx, g.ny, g.nσ, g.dx, g.dy, map(pointer, masks)..., pointer(σ), pointer(w))
            melt = model.melt_rate isa UniformMeltRate ? model.melt_rate.m :
                   throw(ArgumentError("HIPSpec supports UniformMeltRate"))
            phys = PhysParams(p.dt, p.g, p.density_ice, p.density_ocean, p.glen_n, p.glen_reg_strain_rate,
                              p.weertman_m, p.weertman_reg_speed, p.sea_level_wrt_geoid, p.minimum_thickness,
                              melt, p.evolveShelves ? 1 : 0)
            solver = SolverParamsC(s.n_iter_viscosity, s.tol_picard, s.maxiter_picard, s.tol_coarse,
                                   s.maxiter_coarse, s.levels, s.wavelet_threshold, s.nsmooth,
                                   s.smoother_omega, s.stencil_margin, s.super_implicitness)
            out = Ref{Ptr{Cvoid}}(C_NULL)
            rc = ccall((:wavi_ctx_create, LIB), Cint,
                       (Ref{GridDesc}, Ref{PhysParams}, Ref{SolverParamsC}, Cint, Ref{Ptr{Cvoid}}),
                       desc, phys, solver, spec.precision, out)
            check(C_NULL, rc)
            spec.ctx = out[]
            spec.dt = p.dt
        end
    elseif spec.dt != model.params.dt
        check(spec.ctx, ccall((:wavi_set_dt, LIB), Cint, (Ptr{Cvoid}, Cdouble), spec.ctx, model.params.dt))
        spec.dt = model.params.dt
    end
    return model
end

# Field ids (0-based, enum wavi_field) that the Picard loop of
# wavi_update_velocities reads (inner update: velocities, thickness, C * gf,
# grounded fractions, dsdh, glen_b, viscosity, drag, f2; right-hand side:
# surface, accumulation, melt, thickness and surface on u and v) and writes
# (the fields inner_update! sets, velocities, residuals).  Everything else
# (bed, haf, dh/dt, velocities on h, parameters, wavelet coefficients) is
# neither read nor written by the solve.
const SOLVE_IN = (1, 2, 4, 5, 7, 8, 18, 19, 24, 27, 28, 29, 31, 33, 34, 35, 37, 39, 40)
const SOLVE_OUT = (9, 10, 17, 19, 20, 21, 22, 23, 24, 30, 31, 32, 36, 37, 38, 39)
const ALL_FIELDS = 0:40

put_field(ctx, id, a) = check(ctx, ccall((:wavi_set_field, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Csize_t),
                                        ctx, id, a, length(a)))
take_field!(ctx, id, a) = check(ctx, ccall((:wavi_get_field, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Csize_t),
                                         ctx, id, a, length(a)))

# The coarse warm start: compacted on the wavelet mask on the host (the
# reference's correction_coarse[], update_preconditioners.jl:195-216), full
# grid with zeros off the mask on the device.
function push_wavelets!(ctx, model; coefficients::Bool = false)
    wu, wv = model.fields.wu, model.fields.wv
    for (w, idw, idm, idc) in ((wu, 41, 42, 43), (wv, 44, 45, 46))
        coefficients && put_field(ctx, idw, w.wavelets)
        put_field(ctx, idm, Float64.(w.mask))
        full = zeros(Float64, size(w.mask))
        c = w.correction_coarse[]
        length(c) == count(w.mask) && (full[w.mask] .= c)
        put_field(ctx, idc, full)
    end
end

# The wavelet masks, warm start and (coefficients = true) coefficients from
# the device, with the sampling operators rebuilt from the mask as
# update_wavelets! builds them (src/Wavelets/update_wavelets.jl:16-29).
function pull_wavelets!(ctx, model; coefficients::Bool = false)
    wu, wv = model.fields.wu, model.fields.wv
    for (w, idw, idm, idc) in ((wu, 41, 42, 43), (wv, 44, 45, 46))
        coefficients && take_field!(ctx, idw, w.wavelets)
        m = similar(w.wavelets)
        take_field!(ctx, idm, m)
        mask = m .> 0.5
        if mask != w.mask
            w.mask .= mask
            n = count(mask)
            w.n[] = n
            N = length(mask)
            w.crop[] = Diagonal(float(mask[:]))
            w.samp[] = sparse(1:n, (1:N)[mask[:]], ones(n), n, N)
            w.spread[] = sparse(w.samp[]')
        end
        full = similar(w.wavelets)
        take_field!(ctx, idc, full)
        w.correction_coarse[] = full[w.mask]
    end
end

function push_fields!(model, ids = ALL_FIELDS)
    ctx = model.parallel_spec.ctx
    fl = fieldlist(model)
    for id in ids
        put_field(ctx, id, fl[id + 1])
    end
    push_wavelets!(ctx, model; coefficients = ids === ALL_FIELDS)
    return model
end

function pull_fields!(model, ids = ALL_FIELDS)
    ctx = model.parallel_spec.ctx
    fl = fieldlist(model)
    for id in ids
        take_field!(ctx, id, fl[id + 1])
    end
    pull_wavelets!(ctx, model; coefficients = ids === ALL_FIELDS)
    return model
end

"""
    update_velocities!(model::AbstractModel{T,N,M,<:HIPSpec})

`update_velocities!` (`src/Models/update_velocities.jl:11-29`) on the GPU.
"""
function WAVI.update_velocities!(model::AbstractModel{T,N,M,<:HIPSpec}) where {T,N,M}
    WAVI.update_preconditioner!(model)
    push_fields!(model, SOLVE_IN)
    rep = Ref{SolveReport}()
    check(model.parallel_spec.ctx, ccall((:wavi_update_velocities, LIB), Cint, (Ptr{Cvoid}, Ref{SolveReport}),
                                         model.parallel_spec.ctx, rep))
    pull_fields!(model, SOLVE_OUT)
    println("Solved momentum equation on the GPU with residual ", round(rep[].rel_resid, sigdigits=3),
            " at iteration ", rep[].picard_iters)
    return model
end

"""
    run_steps!(simulation, nsteps)

`nsteps` of `timestep!` (`src/Simulations/run_simulation.jl:6-17`) with the
state resident on the device: every field goes to the device once and comes
back once.  With `step_thickness = false` the steps are `update_state!` only.
"""
function run_steps!(simulation, nsteps::Integer)
    model = simulation.model
    nsteps > 0 || return simulation
    WAVI.update_preconditioner!(model)
    push_fields!(model)
    rep = Ref{SolveReport}()
    ctx = model.parallel_spec.ctx
    if simulation.timestepping_params.step_thickness
        check(ctx, ccall((:wavi_run_steps, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{SolveReport}), ctx, nsteps, rep))
    else
        for _ in 1:nsteps
            check(ctx, ccall((:wavi_update_state, LIB), Cint, (Ptr{Cvoid}, Ref{SolveReport}), ctx, rep))
        end
    end
    pull_fields!(model)
    simulation.clock.n_iter += nsteps
    simulation.clock.time += nsteps * simulation.timestepping_params.dt
    return simulation
end

"""
    run_simulation!(simulation) for a model with a `HIPSpec`

`run_simulation!` (`src/Simulations/run_simulation.jl:61-100`) with the time
loop on the device: the steps run by `run_steps!` in stretches that end at
the steps where the reference writes something (temporary and permanent
checkpoints, outputs, the velocity dump at the last step).  There the fields
come back to the host and the reference's own writers run: `@save` of the
whole simulation into `ChkptA/B.jld2` and `PChkpt_<10 digits>.jld2` under
`chkpt_path` (so the reference's `pickup!`, `Simulation.jl:68-86`, reads
them), `write_output` and `write_vel`; then `zip_output`.
"""
function WAVI.run_simulation!(simulation::WAVI.Simulation{<:AbstractModel{T,N,M,<:HIPSpec}}) where {T,N,M}
    tp = simulation.timestepping_params
    op = simulation.output_params
    events = (tp.n_iter_chkpt, tp.n_iter_pchkpt, op.n_iter_out)
    chkpt_tag = "A"
    i = simulation.clock.n_iter
    while i < tp.n_iter_total
        nxt = tp.n_iter_total
        for e in events
            isfinite(e) && (nxt = min(nxt, (i ÷ Int(e) + 1) * Int(e)))
        end
        run_steps!(simulation, nxt - i)
        i = nxt
        if isfinite(tp.n_iter_chkpt) && mod(i, tp.n_iter_chkpt) == 0
            fname = string(tp.chkpt_path, "Chkpt", chkpt_tag, ".jld2")
            @save fname simulation
            chkpt_tag = (chkpt_tag == "A") ? "B" : "A"
            println("making temporary checkpoint at timestep number $(simulation.clock.n_iter)")
        end
        if isfinite(tp.n_iter_pchkpt) && mod(i, tp.n_iter_pchkpt) == 0
            n_iter_string = lpad(simulation.clock.n_iter, 10, "0")
            fname = string(tp.chkpt_path, "PChkpt_", n_iter_string, ".jld2")
            @save fname simulation
            println("making permanent checkpoint at timestep number $(simulation.clock.n_iter)")
        end
        if isfinite(op.n_iter_out) && mod(i, op.n_iter_out) == 0
            WAVI.write_output(simulation)
        end
        if i == tp.n_iter_total && op.dump_vel
            WAVI.write_vel(simulation)
        end
    end
    WAVI.zip_output(simulation)
    return simulation
end

"""
    update_velocities_batch!(models; nthreads = 8, lockstep = false)

`update_velocities!` on every model of an ensemble in one native call (host
threads, one stream per model; the small mixed-precision models as one
lockstep group).  `lockstep = true` puts every model (all mixed precision)
in one lockstep group: their coarse CGs share one launch per CG iteration.
No counterpart in WAVI.jl.
"""
function update_velocities_batch!(models::AbstractVector; nthreads::Integer = 8, lockstep::Bool = false)
    length(unique(objectid.(models))) == length(models) ||
        throw(ArgumentError("update_velocities_batch!: a model appears twice in the batch"))
    for m in models
        WAVI.update_preconditioner!(m)
        push_fields!(m)
    end
    ctxs = Ptr{Cvoid}[m.parallel_spec.ctx for m in models]
    reps = Vector{SolveReport}(undef, length(models))
    codes = zeros(Cint, length(models))
    rc = if lockstep
        ccall((:wavi_update_velocities_lockstep, LIB), Cint,
              (Ptr{Ptr{Cvoid}}, Cint, Ptr{SolveReport}, Ptr{Cint}),
              ctxs, length(models), reps, codes)
    else
        ccall((:wavi_update_velocities_batch, LIB), Cint,
              (Ptr{Ptr{Cvoid}}, Cint, Cint, Ptr{SolveReport}, Ptr{Cint}),
              ctxs, length(models), nthreads, reps, codes)
    end
    if rc != 0
        for (m, c) in zip(models, codes)
            check(m.parallel_spec.ctx, c)
        end
        check(C_NULL, rc)
    end
    for m in models
        pull_fields!(m)
    end
    return models
end

"""
    misfit_gradient(model, n_picard, u_obs, v_obs; tol_linear = 1e-13, maxit_linear = 60)

J = 1/2 ||u - u_obs||^2 after `n_picard` Picard iterations with converged
linear solves, and dJ/d log C (an nx x ny array) by the adjoint sweep.  The
model must hold the state after `update_state!`'s geometry steps.  No
counterpart in WAVI.jl.
"""
function misfit_gradient(model, n_picard::Integer, u_obs::Matrix{Float64}, v_obs::Matrix{Float64};
                         tol_linear::Float64 = 1e-13, maxit_linear::Integer = 60)
    WAVI.update_preconditioner!(model)
    push_fields!(model)
    J = Ref{Cdouble}(0.0)
    grad = zeros(Float64, model.grid.nx, model.grid.ny)
    rep = Ref{AdjointReport}()
    check(model.parallel_spec.ctx,
          ccall((:wavi_misfit_gradient, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cdouble, Cint, Ptr{Float64}, Ptr{Float64}, Ref{Cdouble}, Ptr{Float64},
                 Ref{AdjointReport}),
                model.parallel_spec.ctx, n_picard, tol_linear, maxit_linear, u_obs, v_obs, J, grad, rep))
    pull_fields!(model)
    return J[], grad, rep[]
end

"""
    MisfitProblem(model, n_picard, u_obs, v_obs)

The misfit J(log C) of `misfit_gradient` from a fixed starting state (the
model's fields when the problem is made): `misfit(problem, logC)` is a pure
function of `logC`, and its `rrule` returns the adjoint gradient computed in
the same library call.
"""
struct MisfitProblem{Mo}
    model::Mo
    n_picard::Int
    u_obs::Matrix{Float64}
    v_obs::Matrix{Float64}
    start::Vector{Any}   # copies of fieldlist(model) when the problem was made
end

function MisfitProblem(model, n_picard::Integer, u_obs::Matrix{Float64}, v_obs::Matrix{Float64})
    return MisfitProblem(model, Int(n_picard), u_obs, v_obs, Any[copy(a) for a in fieldlist(model)])
end

"""
    update_geometry!(model)

The steps of `update_state!` before the solve (`src/Models/update_state.jl:8-15`),
including `weertman_c = C * grounded fraction`, on the device.
"""
function update_geometry!(model)
    WAVI.update_preconditioner!(model)
    push_fields!(model)
    check(model.parallel_spec.ctx, ccall((:wavi_update_geometry, LIB), Cint, (Ptr{Cvoid},), model.parallel_spec.ctx))
    pull_fields!(model)
    return model
end

function evaluate(problem::MisfitProblem, logC::AbstractMatrix{Float64})
    model = problem.model
    for (a, a0) in zip(fieldlist(model), problem.start)
        a .= a0
    end
    model.params.weertman_c .= exp.(logC)
    update_geometry!(model)
    return misfit_gradient(model, problem.n_picard, problem.u_obs, problem.v_obs)
end

misfit(problem::MisfitProblem, logC::AbstractMatrix{Float64}) = evaluate(problem, logC)[1]

function ChainRulesCore.rrule(::typeof(misfit), problem::MisfitProblem, logC::AbstractMatrix{Float64})
    J, grad, _ = evaluate(problem, logC)
    misfit_pullback(Jbar) = (NoTangent(), NoTangent(), Jbar .* grad)
    return J, misfit_pullback
end

export HIPSpec, run_steps!, update_velocities_batch!, update_geometry!, misfit_gradient, MisfitProblem, misfit,
       push_fields!, pull_fields!

end # module
