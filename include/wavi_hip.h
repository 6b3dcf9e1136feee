/*
 * wavi_hip.h - C ABI of libwavi_hip.so, the MI355X (gfx950) velocity solve of
 * WAVI.jl.
 *
 * The library replaces, for one model, the reference's dispatch hooks on
 * `parallel_spec` (WAVI.jl src/WAVI.jl:17-18, src/Models/Model.jl:1-9):
 *   update_velocities!(model)            src/Models/update_velocities.jl:11-29
 *   precondition!(model, spec)           src/Models/update_velocities.jl:48-83
 *   update_preconditioner!(model, spec)  src/SharedMemorySpec.jl:31
 * A host binding (Julia `ccall`, Python `ctypes`) keeps the reference's
 * Model / Simulation objects and moves field data through this interface.
 *
 * Conventions
 *  - Every array is Float64 (or uint8 for masks), column-major with the
 *    x index fastest: element (i, j) of an nx-by-ny field is at i + nx*j.
 *    Sigma-grid fields are (nx, ny, nsigma) with the sigma index slowest.
 *  - Host pointers are borrowed for the duration of the call only.  Device
 *    memory is owned by the context.
 *  - Return codes: 0 = success; negative = invalid argument (WAVI_EARG, the
 *    reference's ArgumentError) or size mismatch (WAVI_EDIM, the reference's
 *    DimensionMismatch) or invalid state (WAVI_ESTATE); positive = a HIP
 *    runtime error (WAVI_EHIP).  wavi_last_error() gives the message.
 *  - Non-convergence of Picard or coarse CG is reported in wavi_solve_report,
 *    not as an error, as in the reference (update_velocities.jl:20).
 *  - A context is not re-entrant: use one context per host thread / task.
 */
#ifndef WAVI_HIP_H
#define WAVI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WAVI_OK 0
#define WAVI_EARG (-1)
#define WAVI_EDIM (-2)
#define WAVI_ESTATE (-3)
#define WAVI_EHIP 1

/* Precision of the solve (argument of wavi_ctx_create). */
#define WAVI_PRECISION_FP64 0   /* plain FP64: the reference scheme, bit-reproducible */
#define WAVI_PRECISION_MIXED 1  /* FP32 V-cycle (smoother, Haar transfers and a one-launch-per-
                                  * iteration coarse CG with FP64 cross-workgroup sums) inside
                                  * FP64 residual refinement: every Picard iteration forms the
                                  * FP64 residual b - A(u)u and applies the FP32 correction.
                                  * The FP32 coarse CG's relative stop is 2^-19 ||r0|| (FP32's
                                  * attainable accuracy) in place of sqrt(eps), so its
                                  * coarse_cg_iters can be lower than FP64's.  Once the
                                  * Picard residual is below 1e-7 the coarse CG starts
                                  * from zero instead of the carried warm start
                                  * (*_CORRECTION_COARSE), whose FP32 null-space content
                                  * would otherwise hold the residual near 2e-9 */

/* Grid: the reference's Grid(;...) keyword arguments (src/Grid.jl:66-82).
 * Masks are 0/1 bytes.  u_iszero/v_iszero remove points from the u/v masks
 * (homogeneous Dirichlet); u_isfixed/v_isfixed mark non-homogeneous Dirichlet
 * points whose values are taken from the u/v fields (src/Fields/Fields.jl:34-35,
 * src/Fields/UGrid.jl:76). */
typedef struct wavi_grid_desc {
    int32_t nx, ny, nsigma;
    double dx, dy;
    const uint8_t *h_mask;      /* nx*ny */
    const uint8_t *h_isfixed;   /* nx*ny (may be NULL: all false) */
    const uint8_t *u_iszero;    /* (nx+1)*ny (may be NULL) */
    const uint8_t *v_iszero;    /* nx*(ny+1) (may be NULL) */
    const uint8_t *u_isfixed;   /* (nx+1)*ny (may be NULL) */
    const uint8_t *v_isfixed;   /* nx*(ny+1) (may be NULL) */
    const double *sigma;        /* nsigma, sigma[0]=0, sigma[nsigma-1]=1 */
    const double *quadrature_weights; /* nsigma */
} wavi_grid_desc;

/* Physical parameters used by the solve (src/Params.jl:1-25). */
typedef struct wavi_phys_params {
    double dt;                    /* timestep injected by Simulation (Simulation.jl:56-59) */
    double g;
    double density_ice;
    double density_ocean;
    double glen_n;
    double glen_reg_strain_rate;
    double weertman_m;
    double weertman_reg_speed;
    double sea_level_wrt_geoid;
    double minimum_thickness;
    double uniform_melt_rate;     /* UniformMeltRate(m) (src/MeltRate/MeltRate.jl:37-53) */
    int32_t evolve_shelves;       /* only 1 is supported (run_simulation.jl:30) */
} wavi_phys_params;

/* Solver parameters (src/SolverParams.jl:2-14). */
typedef struct wavi_solver_params {
    int32_t n_iter_viscosity;     /* must be 2 (SolverParams.jl:3) */
    double tol_picard;
    int32_t maxiter_picard;
    double tol_coarse;
    int32_t maxiter_coarse;       /* 0..2^30 */
    int32_t levels;               /* Haar levels: blocks of 2^(levels+1), 0..3 */
    double wavelet_threshold;
    int32_t nsmooth;
    double smoother_omega;
    int32_t stencil_margin;       /* colours used to probe the operator diagonal */
    double super_implicitness;
} wavi_solver_params;

#define WAVI_REPORT_MAX_PICARD 64

/* Outcome of one update_velocities!. */
typedef struct wavi_solve_report {
    int32_t picard_iters;         /* i_picard at loop exit */
    int32_t converged;            /* rel_resid < tol_picard */
    double rel_resid;             /* last ||b - A x|| / ||b|| */
    int32_t coarse_cg_total;      /* sum of coarse CG iterations */
    int32_t refinements;          /* mixed precision: FP64 residual refinements */
    int32_t coarse_cg_iters[WAVI_REPORT_MAX_PICARD]; /* per V-cycle (first 64) */
    double rel_resid_hist[WAVI_REPORT_MAX_PICARD];   /* per Picard iteration */
} wavi_solve_report;

/* Field identifiers: the arrays of the reference's Fields structs
 * (src/Fields/HGrid.jl:1-37, UGrid.jl:1-29, VGrid.jl, SigmaGrid.jl:98-109,
 * src/Wavelets/UWavelets.jl:2-15) plus the two per-cell parameter arrays the
 * state update reads (Params weertman_c and accumulation_rate, Model.jl:63-74). */
enum wavi_field {
    /* h-grid, nx*ny */
    WAVI_H_B = 0, WAVI_H_H, WAVI_H_S, WAVI_H_DHDT, WAVI_H_ACCUMULATION,
    WAVI_H_BASAL_MELT, WAVI_H_HAF, WAVI_H_GROUNDED_FRACTION, WAVI_H_DSDH,
    WAVI_H_SHELF_STRAIN_RATE, WAVI_H_AV_SPEED, WAVI_H_U, WAVI_H_V, WAVI_H_US,
    WAVI_H_VS, WAVI_H_UB, WAVI_H_VB, WAVI_H_BED_SPEED, WAVI_H_WEERTMAN_C,
    WAVI_H_BETA, WAVI_H_BETA_EFF, WAVI_H_TAU_BED, WAVI_H_ETA_AV,
    WAVI_H_QUAD_F1, WAVI_H_QUAD_F2,
    WAVI_P_WEERTMAN_C, WAVI_P_ACCUMULATION_RATE,
    /* u-grid, (nx+1)*ny */
    WAVI_U_S, WAVI_U_H, WAVI_U_GROUNDED_FRACTION, WAVI_U_BETA_EFF, WAVI_U_U,
    WAVI_U_RESIDUAL,
    /* v-grid, nx*(ny+1) */
    WAVI_V_S, WAVI_V_H, WAVI_V_GROUNDED_FRACTION, WAVI_V_BETA_EFF, WAVI_V_V,
    WAVI_V_RESIDUAL,
    /* sigma grid, nx*ny*nsigma */
    WAVI_S_ETA, WAVI_S_GLEN_B,
    /* wavelet grids (same shapes as u and v grids); masks as 0.0/1.0 and the
     * coarse warm start in full (uncompacted) layout, zero off the mask */
    WAVI_WU_WAVELETS, WAVI_WU_MASK, WAVI_WU_CORRECTION_COARSE,
    WAVI_WV_WAVELETS, WAVI_WV_MASK, WAVI_WV_CORRECTION_COARSE,
    WAVI_NFIELDS
};

typedef struct wavi_ctx wavi_ctx;

/* Context life cycle.  Creating the context builds the masks and operators of
 * setup_fields (src/Fields/Fields.jl:26-115) on the device and initialises
 * the fields as the HGrid/UGrid/VGrid/UWavelets constructors do
 * (all-true wavelet masks, quad_f2 = h/(3*eta_av) once h and eta_av are set;
 * use wavi_init_quad_f2). Corresponds to Model(;...) + update_preconditioner!. */
int wavi_ctx_create(const wavi_grid_desc *grid, const wavi_phys_params *phys,
                    const wavi_solver_params *solver, int precision_mode,
                    wavi_ctx **out);
int wavi_ctx_destroy(wavi_ctx *ctx);
const char *wavi_last_error(const wavi_ctx *ctx);
int wavi_set_dt(wavi_ctx *ctx, double dt);

/* Field transfer (host <-> device).  n is the element count of the field. */
int wavi_field_size(const wavi_ctx *ctx, int field, size_t *n);
int wavi_set_field(wavi_ctx *ctx, int field, const double *host, size_t n);
int wavi_get_field(wavi_ctx *ctx, int field, double *host, size_t n);
/* quad_f2 = h/(3 eta_av) on the h mask (src/Fields/HGrid.jl:229). */
int wavi_init_quad_f2(wavi_ctx *ctx);

/* Save / restore the state a velocity solve mutates (u, v, eta, beta,
 * quad_f2, warm start, wavelet masks, ...), device to device.  Used to time
 * repeated solves from one reset state. */
int wavi_state_save(wavi_ctx *ctx);
int wavi_state_restore(wavi_ctx *ctx);

/* update_velocities!(model) (src/Models/update_velocities.jl:11-29):
 * Picard loop of inner_update! + precondition! on device-resident fields. */
int wavi_update_velocities(wavi_ctx *ctx, wavi_solve_report *report);
/* update_velocities! for each of n contexts (an ensemble of independent
 * models, of any grid sizes and precisions), run by nthreads host threads;
 * every context has its own stream, so the GPU overlaps the members'
 * kernels.  Small mixed-precision members (at most 256 tiles of 16 x 16,
 * one device and Haar level count, their coarse-CG arenas within 128 MB)
 * run as one lockstep group (wavi_update_velocities_lockstep) on one of the
 * threads, concurrently with the other members on the rest.  reports and codes (each NULL or n entries) receive each
 * member's report and return code (wavi_last_error of that member's context
 * holds its message).  Returns 0, or the first nonzero member code in member
 * order; WAVI_EARG before any member runs when a context is NULL or appears
 * twice.  No reference counterpart: the reference solves one model per call. */
int wavi_update_velocities_batch(wavi_ctx **ctxs, int n, int nthreads,
                                 wavi_solve_report *reports, int *codes);
/* update_velocities! for n mixed-precision contexts in lockstep, driven by
 * the calling thread.  Each Picard iteration runs every member's inner
 * update and residual on the member's own stream; the members whose residual
 * asks for a V-cycle then run it up to the coarse right-hand side, and their
 * coarse CGs run together: one launch per CG iteration over the tiles of all
 * of them (k_cgm_iter_batch), until every member's CG has stopped; the rest
 * of each V-cycle follows on the members' streams.  Every member is
 * bit-identical to its own wavi_update_velocities.  All contexts must be
 * mixed precision, on one device, with one Haar level count (else
 * WAVI_EARG, as for a NULL or repeated context).  reports and codes as for
 * wavi_update_velocities_batch.  No reference counterpart: the reference
 * solves one model per call. */
int wavi_update_velocities_lockstep(wavi_ctx **ctxs, int n, wavi_solve_report *reports, int *codes);
/* wavi_update_velocities_lockstep without any device allocation: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, checks and results as
 * wavi_update_velocities_lockstep, bit for bit.  The device tables of the
 * shared coarse-CG launches (per member about 100 bytes plus one int per
 * 8 tiles) live in a static buffer of WAVI_NOALLOC_SCRATCH_BYTES on each
 * GPU, held until the process ends and set up by the first wavi_ctx_create
 * on that GPU; the events are the contexts' own.  At most
 * WAVI_NOALLOC_MAX_MEMBERS members, and the tables must fit the buffer:
 * WAVI_EARG otherwise, before any work.  Every call on one GPU uses that
 * GPU's one buffer, which keeps the last call's tables: calls that overlap
 * on one GPU race, so a caller serialises them. */
#define WAVI_NOALLOC_MAX_MEMBERS 64
#define WAVI_NOALLOC_SCRATCH_BYTES (256 * 1024)
int wavi_update_velocities_lockstep_noalloc(wavi_ctx **ctxs, int n, wavi_solve_report *reports, int *codes);
/* update_state! (update_state.jl:7-21) on each of n contexts, like
 * wavi_update_velocities_batch: the subdomain solves of the restricted
 * additive Schwarz preconditioner (src/SharedMemorySpec.jl:96-133 runs
 * update_state! on every subdomain model, one task each). */
int wavi_update_state_batch(wavi_ctx **ctxs, int n, int nthreads, wavi_solve_report *reports, int *codes);

/* precondition!(model, ::SharedMemorySpec) (src/SharedMemorySpec.jl:76-153)
 * on the device: the residual of the full model's start guess (residual
 * fields set, relative residual into *rel_resid, *converged = rel_resid <
 * tol_picard), and unless converged, niterations rounds of restricted
 * additive Schwarz: every subdomain's u and v from the full model's
 * (schwarzRestrictVelocities!, :306-331), update_state! of every
 * subdomain (nthreads host threads, one stream each), the full model's
 * velocities times damping, and each subdomain's velocities weighted by the
 * partition of unity and (1 - damping) added in subdomain order
 * (schwarzProlongVelocities!, :351-382).  subs holds the ngridsx * ngridsy
 * subdomain contexts, subdomain (igrid, jgrid) at (igrid - 1) * ngridsy +
 * jgrid - 1, each built on the grid and fields schwarzModel gives it
 * (:181-287) and on the full model's device.  No field crosses to the host. */
int wavi_schwarz_precondition(wavi_ctx *full, wavi_ctx **subs, int ngridsx, int ngridsy, int overlap,
                              int niterations, double damping, int nthreads, int *converged, double *rel_resid);

/* update_preconditioner!(model, ::SharedMemorySpec) (SharedMemorySpec.jl:51-68)
 * on existing subdomain contexts: every subdomain's fields as schwarzModel
 * (:181-287) and setup_fields (src/Fields/Fields.jl:26-115) build them from
 * the full model's current fields (bed, thickness, grounded fraction, u, v,
 * viscosity, glen_b, Weertman coefficient and accumulation sliced; eta_av
 * the bottom viscosity level, quad_f2 = h / (3 eta_av) on the mask, dsdh and
 * the u/v grounded fractions 1, the wavelet masks set, every other field 0),
 * on the device.  The subdomain contexts must have been created on
 * schwarzModel's grids; subs as for wavi_schwarz_precondition. */
int wavi_schwarz_refresh(wavi_ctx *full, wavi_ctx **subs, int ngridsx, int ngridsy, int overlap);

/* inner_update!(model) (update_velocities.jl:32-45). */
int wavi_inner_update(wavi_ctx *ctx);

/* update_state!(model[, clock]) (src/Models/update_state.jl:7-42) with
 * UniformMeltRate, including update_wavelets!. */
int wavi_update_state(wavi_ctx *ctx, wavi_solve_report *report);
/* The update_state! steps that precede the velocity solve: surface, geometry
 * on u/v, height above floatation, grounded fractions, accumulation, melt,
 * weertman_c, dsdh (update_state.jl:8-15). */
int wavi_update_geometry(wavi_ctx *ctx);
/* The update_state! steps after the solve: velocities on the h grid, dh/dt
 * (update_state.jl:16-20, 164-189) and update_wavelets! (:21). */
int wavi_after_solve(wavi_ctx *ctx);
/* update_thickness!(simulation) (src/Simulations/run_simulation.jl:24-42). */
int wavi_update_thickness(wavi_ctx *ctx, double dt);
/* nsteps of timestep!: update_state! then update_thickness! (run_simulation.jl:6-17). */
int wavi_run_steps(wavi_ctx *ctx, int nsteps, wavi_solve_report *last_report);
/* update_wavelets!(model) (src/Wavelets/update_wavelets.jl:6-32). */
int wavi_update_wavelets(wavi_ctx *ctx);

/* ---- adjoint basal-drag gradient (no reference counterpart: WAVI.jl has
 * no adjoint; this extends its solve for inversions) ----
 * The map.  From the context's current state (velocities, viscosity,
 * drag, quadrature f2, as update_velocities! would start from them):
 * n_picard Picard iterations, each inner_update! (update_velocities.jl:32-45)
 * followed by the linear system A u = b of that iteration solved to
 * ||b - A u|| <= tol_linear ||b|| by flexible CG in FP64 preconditioned
 * with the V-cycle of update_velocities! in the context's precision (at
 * most maxit_linear V-cycles per solve); then
 * J = 1/2 sum over inner u and v points of (u - u_obs)^2.
 * grad_logc (NULL: the forward and J only) receives dJ/d log C on the h
 * grid (nx*ny, column-major), C the per-cell Weertman coefficient
 * P_WEERTMAN_C, through H_WEERTMAN_C = C * grounded fraction: call
 * wavi_update_geometry after setting C.  The gradient is exact for the map
 * (reverse sweep over the checkpointed Picard states, with the adjoint
 * systems A_k lambda = ubar solved like the forward ones: A_k is symmetric).
 * u_obs: (nx+1)*ny, v_obs: nx*(ny+1).  The context is left in the state of
 * the forward's last Picard iteration.  Fixed velocities (u_isfixed,
 * v_isfixed) are rejected with WAVI_EARG.  The first call allocates a device
 * workspace of (n_picard + 1) (N_u + N_v + nsigma N_h + 2 N_h) doubles plus a
 * few vectors, kept with the context (grown when a later call needs more)
 * until wavi_ctx_destroy. */
typedef struct wavi_adjoint_report {
    int32_t n_picard;
    int32_t fwd_vcycles;   /* V-cycles of all forward linear solves */
    int32_t adj_vcycles;   /* V-cycles of all adjoint linear solves */
    int32_t unconverged;   /* linear solves stopped by maxit_linear */
    double max_fwd_rel;    /* largest final ||b - A u|| / ||b|| of a forward solve */
    double max_adj_rel;    /* the same for the adjoint solves */
    int64_t coarse_cg_iters; /* coarse-CG iterations of all those V-cycles */
} wavi_adjoint_report;
int wavi_misfit_gradient(wavi_ctx *ctx, int n_picard, double tol_linear, int maxit_linear,
                         const double *u_obs, const double *v_obs, double *J, double *grad_logc,
                         wavi_adjoint_report *report);
/* wavi_misfit_gradient for each of n contexts (an ensemble), driven by
 * nthreads host threads with one stream per context, like
 * wavi_update_velocities_batch.  u_obs[k], v_obs[k]: member k's
 * observations; J (n entries) receives the misfits; grad_logc (NULL, or n
 * pointers of which any may be NULL) the gradients; reports and codes (each
 * NULL or n entries) the members' reports and return codes.  Returns 0 or
 * the first nonzero member code.  The small mixed-precision members (as in
 * wavi_update_velocities_batch) run on up to 32 threads of their own and
 * meet at every coarse CG: when all of them that are running wait there,
 * their CGs run as one batch (one launch per CG iteration over all their
 * tiles); every member stays bit-identical to its own call.  No reference
 * counterpart. */
int wavi_misfit_gradient_batch(wavi_ctx **ctxs, int n, int nthreads, int n_picard, double tol_linear,
                               int maxit_linear, const double *const *u_obs, const double *const *v_obs,
                               double *J, double *const *grad_logc, wavi_adjoint_report *reports, int *codes);
/* The linear solves of wavi_misfit_gradient (flexible CG, whose
 * preconditioner may be any fixed or varying SPD map) run the V-cycle with
 * its coarse CG preconditioned by the diagonal of the coarse operator
 * (Jacobi), computed for each linear system; on = 0 restores the
 * update_velocities! V-cycle there.  The gradient is the same to the solve
 * tolerance either way.  Default on.  No reference counterpart. */
int wavi_set_linear_coarse_jacobi(wavi_ctx *ctx, int on);

/* ---- test hooks: single pieces of the solve on host arrays ---- */
/* op_fun! (src/utilities.jl:61-145) on full-grid inputs xu ((nx+1)*ny) and
 * xv (nx*(ny+1)); the output holds the operator at inner points and zero
 * elsewhere.  sampled=1 masks the inputs to inner points first (the
 * vecSampled=true path); sampled=0 uses them as given (vecSampled=false). */
int wavi_op_apply(wavi_ctx *ctx, const double *xu, const double *xv,
                  double *yu, double *yv, int sampled);
/* get_rhs (update_velocities.jl:106-182), full layout, zero off inner points. */
int wavi_rhs(wavi_ctx *ctx, double *bu, double *bv);
/* get_op_diag (src/Wavelets/update_preconditioners.jl:92-109). */
int wavi_op_diag(wavi_ctx *ctx, double *du, double *dv);
/* One V-cycle precondition!(x, p, b) (update_preconditioners.jl:54-83) with
 * x starting at zero; uses and updates the coarse warm start. */
int wavi_vcycle(wavi_ctx *ctx, const double *ru, const double *rv,
                double *eu, double *ev, int32_t *coarse_iters);
/* Diagonal of the coarse-CG operator A_c = R A P (with the masks the CG
 * applies; update_preconditioners.jl:73) for the current coefficients
 * (call wavi_inner_update first), in the full-grid layout, zero off the
 * wavelet masks. */
int wavi_coarse_diag(wavi_ctx *ctx, double *du, double *dv);
/* 2-D Haar transform of one component (comp 0 = u grid, 1 = v grid).
 * dir 0: forward dwt (UGrid.jl:89); 1: inverse idwt (UWavelets.jl:48);
 * 2: transposed inverse idwt^T (UWavelets.jl:49). */
int wavi_haar(wavi_ctx *ctx, int comp, int dir, const double *in, double *out);

/* Tile shape of the smoother kernel (one launch per smoothing iteration
 * runs the four colour sweeps of gauss_seidel_smoother!,
 * src/Wavelets/update_preconditioners.jl:133-150, on tiles of tx x ty
 * points).  The shapes are 16 x 16, 32 x 16 and 16 x 32; every shape gives
 * the same bits.  A grid whose window (the grid and a border of one point)
 * holds at most 2048 points and fits 64 KiB of LDS (80 x 10 in either
 * precision) can also be smoothed whole in one workgroup, every smoothing
 * iteration of a V-cycle in one launch: the shape -1 x -1.  By default the
 * library's autotuner picks the variant on the context's first V-cycle: it
 * times one smoothing phase of each (the three shapes walked in bands of
 * tile columns or of tile rows, one band per XCD, and the whole grid where it
 * fits) on the context's own grid, keeps the fastest, and remembers the
 * choice for later contexts of the same grid size and precision in the
 * process.  wavi_set_smoother_tile fixes a shape, walked in column bands, or
 * -1 x -1 (WAVI_EARG if the grid does not fit); 0, 0 tunes again.
 * wavi_smoother_tile reports the shape in use for the context's V-cycle
 * precision (0, 0 before the first V-cycle).
 * No reference counterpart. */
int wavi_set_smoother_tile(wavi_ctx *ctx, int tx, int ty);
int wavi_smoother_tile(const wavi_ctx *ctx, int *tx, int *ty);

/* How the coarse CG of the V-cycle runs (update_preconditioners.jl:73).
 * WAVI_CG_LAUNCHES: one launch per CG iteration over all tiles (three in
 * FP64, one in mixed precision), queued in chunks.  WAVI_CG_WORKGROUP: the
 * whole CG in one launch of one workgroup, which runs each iteration on the
 * tiles in turn, with no launch and no host round trip between iterations.
 * Both give the same bits.  WAVI_CG_AUTO (the default) is the library's
 * choice, at present WAVI_CG_LAUNCHES for every grid: measured on MI355X,
 * the workgroup is 3.4 to 3.6 times slower per iteration even on a grid of
 * six 16 x 16 tiles.  wavi_coarse_cg reports the mode in effect for the
 * context.  Lockstep batches (wavi_update_velocities_lockstep*) and the
 * small members of wavi_misfit_gradient_batch share one launch per CG
 * iteration whatever the mode.  No reference counterpart. */
#define WAVI_CG_AUTO 0
#define WAVI_CG_LAUNCHES 1
#define WAVI_CG_WORKGROUP 2
int wavi_set_coarse_cg(wavi_ctx *ctx, int mode);
int wavi_coarse_cg(const wavi_ctx *ctx, int *mode);

/* Kernel timing with HIP events on the context's stream, for measurement.
 * Sampled single launches (events around every 16th launch, so each time
 * includes the cost of the two event records): WAVI_KT_CG_APPLY (FP64
 * coarse CG: A then R on one tile, dot partials), WAVI_KT_CG_DIR (direction
 * update and prolongation), WAVI_KT_CG_UPDATE (axpys and norm partials),
 * WAVI_KT_COLOUR (one smoother iteration: its four colour sweeps in one
 * launch).  Bursts (events around each
 * chunk of 32 back-to-back iterations; only chunks in which every launch
 * still iterates are counted, and the count is of launches):
 * WAVI_KT_CG_ITER_MIXED (mixed-precision coarse CG, one whole iteration per
 * launch of k_cgm_iter) and WAVI_KT_CG_ITER_FP64 (FP64 coarse CG, one
 * iteration = k_cg_dir + k_cg_apply + k_cg_update).  FP64 chunks alternate
 * between burst timing and the sampled single launches.
 * enable=1 starts (and clears) the accumulation; wavi_profile_read returns
 * the total milliseconds and the number of timed launches (iterations).
 * No reference counterpart. */
#define WAVI_KT_CG_APPLY 0
#define WAVI_KT_CG_DIR 1
#define WAVI_KT_CG_UPDATE 2
#define WAVI_KT_COLOUR 3
#define WAVI_KT_CG_ITER_MIXED 4
#define WAVI_KT_CG_ITER_FP64 5
#define WAVI_KT_COUNT 6
int wavi_profile_enable(wavi_ctx *ctx, int enable);
int wavi_profile_read(wavi_ctx *ctx, int kernel, double *total_ms, int64_t *launches);

/* Device arithmetic probe on host arrays of n elements: fn 0 out = x^y by the
 * library's deterministic power function, 1 out = sqrt(x), 2 out = x / y.
 * No reference counterpart: it checks that the device rounds as the CPU does. */
int wavi_math_probe(int fn, const double *x, const double *y, double *out, size_t n);

#ifdef __cplusplus
}
#endif
#endif /* WAVI_HIP_H */
