"""Benchmark: velocity solves per second of WAVI's update_velocities! on one
MI355X at 1280 x 160 x 8 (the channel single-solve workload: 20 Picard
iterations to 1e-8, coarse CG to 1e-8 with at most 1000 iterations, 3 Haar
levels), with the inputs resident in HBM.

A step is one update_velocities! from a fixed reset state (velocities,
viscosity, drag, quadratures, warm start and wavelet masks restored on the
device before every solve, outside the timed interval).  The headline value
is the mixed-precision solve (FP32 V-cycle with the single-launch coarse CG,
FP64 residual refinement); the FP64 solve of the same state is timed beside
it and gives the speedup and the velocity error of the mixed solve.  The
timed solves run without instrumentation; the dominant kernel's launch time
comes from one more solve of each precision with HIP events on, after the
timed loop.  The default line also carries the basal-drag gradient (adjoint
ms, FP64 and mixed, checked by central differences), the 80 x 10 channel
transient (200 steps) on the GPU beside the CPU oracle, the 64-member
ensemble (forward solves and forward + adjoint gradients, members/s), and
the CPU baseline.  When the mixed velocities are more than 1e-6 from FP64 (relative
L2 or max |du| / max |u|) the line says "valid": false and the exit status
is 1.  Prints one JSON line.  --precision fp64 or mixed times one of them
alone (for profiling runs).
"""
import argparse
import ctypes
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
ERROR_BAR = 1e-6


def metric(nx, ny):
    return "velocity solves/s at %dx%dx8, mixed vs FP64 GPU @1e-6 rel err; adjoint ms" % (nx, ny)


def _package():
    if "wavi_jl_amd" in sys.modules:
        return sys.modules["wavi_jl_amd"]
    pkg_dir = os.path.join(ROOT, "wavi.jl_amd")
    spec = importlib.util.spec_from_file_location(
        "wavi_jl_amd", os.path.join(pkg_dir, "__init__.py"), submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["wavi_jl_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sizes(nx, ny):
    return (nx + 1) * ny, nx * (ny + 1), nx * ny, (nx - 1) * (ny - 1)


def cg_apply_bytes(nx, ny, w=8):
    """Algorithmic HBM bytes of one k_cg_apply launch (FP64 coarse CG, A then
    R on a tile): t_u, t_v in; c_u, c_v out; u_u, u_v in for dot(u, c);
    beta_eff on u, v; thickness on u, v; h*eta_av on h; dimp on h; h*eta_av
    on c (w bytes each); masks (one byte each): inner u, v; wavelet u, v;
    u, v masks; h and c masks."""
    nu, nv, nh, nc = sizes(nx, ny)
    return w * (5 * (nu + nv) + 2 * nh + nc) + 3 * (nu + nv) + nh + nc


def cgm_iter_bytes(nx, ny, w=4):
    """Algorithmic HBM bytes of one k_cgm_iter launch (one whole iteration of
    the mixed-precision coarse CG), averaged over the two alternating
    launch parities: r, w, p, s in and out (8 (N_u + N_v) values) and x in
    and out every other launch (1 (N_u + N_v) on average); beta_eff and
    thickness on u and v; h*eta_av and dimp on h; h*eta_av on c; the ring
    images P r, P s, P w (a 60-point ring written and a 68-point halo read
    per 16 x 16 tile and kind: 3 * 128 / 256 values per point); mask bits
    (one byte per u and v point)."""
    nu, nv, nh, nc = sizes(nx, ny)
    return int(w * (9 * (nu + nv) + 2 * (nu + nv) + 1.5 * (nu + nv) + 2 * nh + nc) + (nu + nv))


def cg_iteration_fp64_bytes(nx, ny, w=8):
    """Algorithmic HBM bytes of one FP64 coarse-CG iteration (k_cg_dir +
    k_cg_apply + k_cg_update).  k_cg_dir: r, u in, u, t out, wavelet and
    inner masks; k_cg_apply as above; k_cg_update: x, u, r, c in, x, r out."""
    nu, nv, nh, nc = sizes(nx, ny)
    return 4 * w * (nu + nv) + 2 * (nu + nv) + cg_apply_bytes(nx, ny, w) + 6 * w * (nu + nv)


DOMINANT = {  # precision -> (what is timed, timing id, profile name, bytes per launch)
    "fp64": ("FP64 CG iteration (k_cg_dir + k_cg_apply + k_cg_update)", "cg_iter_fp64", "cg_iteration",
             lambda nx, ny: cg_iteration_fp64_bytes(nx, ny, 8)),
    "mixed": ("k_cgm_iter", "cg_iter_mixed", "k_cgm_iter", lambda nx, ny: cgm_iter_bytes(nx, ny, 4)),
}


def measured_traffic(nx, ny, precision):
    """HBM bytes per launch (per iteration for FP64) of the dominant kernel
    from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary for this
    workload, scaled by the copy-kernel calibration recorded there; None when
    no summary matches."""
    name = DOMINANT[precision][2]
    path = os.path.join(ROOT, "profiles", "%s_%dx%d_%s_traffic.md" % (name, nx, ny, precision))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch %s:" % name):
            return int(line.split(":")[1])
    return None


def calibration_copy(mib):
    """One device-to-device copy of known size (a plain copy kernel), the
    reference point that scales the FETCH_SIZE / WRITE_SIZE counters."""
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), n) == 0 and hip.hipMalloc(ctypes.byref(b), n) == 0
    assert hip.hipMemset(a, 1, n) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, n, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def setup(W, nx, ny, precision, picard=None, maxiter_coarse=None):
    kw = {} if picard is None else {"maxiter_picard": picard}
    m = W.workloads.channel_solve(nx=nx, ny=ny, precision=precision, **kw)
    if maxiter_coarse is not None:
        m.solver_params.maxiter_coarse = maxiter_coarse
    W.update_preconditioner(m)
    W.push_fields(m)
    spec = m.parallel_spec
    lib = spec._lib
    assert lib.wavi_update_geometry(spec.ctx) == 0
    assert lib.wavi_state_save(spec.ctx) == 0
    return m, spec, lib


def solve(W, spec, lib):
    rep = W._lib.SolveReport()
    rc = lib.wavi_update_velocities(spec.ctx, ctypes.byref(rep))
    if rc != 0:
        raise RuntimeError("wavi_update_velocities failed: %s" % lib.wavi_last_error(spec.ctx))
    return rep.as_dict()


def velocities(W, m, spec):
    """u and v of the model's device state, concatenated."""
    lib = spec._lib
    nx, ny = m.grid.nx, m.grid.ny
    u = np.empty((nx + 1) * ny)
    v = np.empty(nx * (ny + 1))
    assert lib.wavi_get_field(spec.ctx, W._lib.FIELD["U_U"], W._lib.dptr(u), u.size) == 0
    assert lib.wavi_get_field(spec.ctx, W._lib.FIELD["V_V"], W._lib.dptr(v), v.size) == 0
    return np.concatenate([u, v])


def timed_solves(W, nx, ny, precision, steps, warmup, picard=None, maxiter_coarse=None):
    """Solve from the reset state warmup + steps times and time the last
    steps with nothing but the solve inside the timed interval; then one
    more solve with the library's HIP-event timing of the dominant kernel
    on.  Returns (times, report, velocities, roofline of the dominant
    kernel)."""
    log("[bench] setup %dx%dx8 precision=%s" % (nx, ny, precision))
    m, spec, lib = setup(W, nx, ny, precision, picard, maxiter_coarse)
    for k in range(warmup):
        assert lib.wavi_state_restore(spec.ctx) == 0
        t0 = time.perf_counter()
        rep = solve(W, spec, lib)
        log("[bench] %s warmup %d: %.3f s, picard %d, cg %d, rel %.3e" % (
            precision, k, time.perf_counter() - t0, rep["picard_iters"], rep["coarse_cg_total"], rep["rel_resid"]))
    times = []
    for k in range(steps):
        assert lib.wavi_state_restore(spec.ctx) == 0
        t0 = time.perf_counter()
        rep = solve(W, spec, lib)
        times.append(time.perf_counter() - t0)
        log("[bench] %s step %d: %.3f s, picard %d, cg %d, rel %.3e" % (
            precision, k, times[-1], rep["picard_iters"], rep["coarse_cg_total"], rep["rel_resid"]))
    vel = velocities(W, m, spec)
    # kernel timing: a separate solve of the same state with the event timing on
    kernel, kid, _, nbytes = DOMINANT[precision]
    assert lib.wavi_state_restore(spec.ctx) == 0
    lib.wavi_profile_enable(spec.ctx, 1)
    t0 = time.perf_counter()
    solve(W, spec, lib)
    t_prof = time.perf_counter() - t0
    ms = ctypes.c_double(0.0)
    n = ctypes.c_int64(0)
    lib.wavi_profile_read(spec.ctx, W._lib.KT[kid], ctypes.byref(ms), ctypes.byref(n))
    lib.wavi_profile_enable(spec.ctx, 0)
    log("[bench] %s solve with kernel timing: %.3f s" % (precision, t_prof))
    # no complete chunk of 32 iterations (a capped profiling run): no timing
    avg_s = ms.value / 1e3 / n.value if n.value > 0 else None
    B = nbytes(nx, ny)
    achieved = B / avg_s / 1e9 if avg_s else None
    roof = {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None,
            "traffic": measured_traffic(nx, ny, precision),
            "bytes_per_launch": B, "avg_launch_us": avg_s * 1e6 if avg_s else None, "launches": n.value,
            "timing": "HIP events on the solve's stream around chunks of 32 back-to-back CG iterations "
                      "(chunks that ran past the CG stop excluded), in one solve after the timed ones"}
    m.parallel_spec.close()
    return times, rep, vel, roof


def ensemble_run(W, n, nthreads, steps, warmup, precision="mixed", adjoint=False, n_picard=10, max_nx=None,
                 compact=False):
    """The seeded ensemble (channel solves of mixed sizes, nσ, Glen A and
    basal-drag fields) from the reset state, one batched call per step:
    forward solves (wavi_update_velocities_batch, 20 Picard iterations: the
    small members in one lockstep group, the large ones one per thread; also
    timed with every member in one lockstep group), or
    with adjoint=True forward + adjoint gradients (wavi_misfit_gradient_batch,
    n_picard Picard iterations with converged solves; each member's
    observations are its own forward velocities times (1 + 0.01 N(0, 1))).
    Also timed member after member with one thread, for the concurrency
    gain, unless compact.  Returns the JSON line."""
    if max_nx:
        # the same seeded draw, keeping the first n members at most max_nx wide
        specs = [sp for sp in W.workloads.ensemble_specs(16 * n) if sp["nx"] <= max_nx][:n]
    else:
        specs = W.workloads.ensemble_specs(n)
    log("[bench] ensemble: %d members, setup" % n)
    members = []
    for sp in specs:
        m = W.workloads.ensemble_member(sp, precision, n_picard if adjoint else 20)
        W.update_preconditioner(m)
        W.push_fields(m)
        spec = m.parallel_spec
        assert spec._lib.wavi_update_geometry(spec.ctx) == 0
        assert spec._lib.wavi_state_save(spec.ctx) == 0
        members.append(m)
    lib = members[0].parallel_spec._lib
    ctxs = (ctypes.c_void_p * n)(*[m.parallel_spec.ctx for m in members])
    codes = (ctypes.c_int * n)()
    if adjoint:
        obs = []
        for sp, m in zip(specs, members):
            nx, ny = sp["nx"], sp["ny"]
            z = np.zeros((nx + 1) * ny + nx * (ny + 1))
            J = ctypes.c_double(0.0)
            ctx = m.parallel_spec.ctx
            assert lib.wavi_misfit_gradient(ctx, n_picard, 1e-13, 200, W._lib.dptr(z), W._lib.dptr(z),
                                            ctypes.byref(J), None, None) == 0
            u, v = np.empty((nx + 1) * ny), np.empty(nx * (ny + 1))
            assert lib.wavi_get_field(ctx, W._lib.FIELD["U_U"], W._lib.dptr(u), u.size) == 0
            assert lib.wavi_get_field(ctx, W._lib.FIELD["V_V"], W._lib.dptr(v), v.size) == 0
            uo, vo = W.workloads.inversion_observations(u, v, seed=sp["seed"])
            obs.append((np.ascontiguousarray(uo), np.ascontiguousarray(vo)))
        up = (W._lib._dp * n)(*[W._lib.dptr(a) for a, _ in obs])
        vp = (W._lib._dp * n)(*[W._lib.dptr(b) for _, b in obs])
        Js = np.zeros(n)
        grads = [np.zeros(sp["nx"] * sp["ny"]) for sp in specs]
        gp = (W._lib._dp * n)(*[W._lib.dptr(g) for g in grads])
        reps = (W._lib.AdjointReport * n)()
    else:
        reps = (W._lib.SolveReport * n)()

    def run(threads, lockstep=False):
        for m in members:
            assert lib.wavi_state_restore(m.parallel_spec.ctx) == 0
        t0 = time.perf_counter()
        if lockstep:
            rc = lib.wavi_update_velocities_lockstep(ctxs, n, reps, codes)
        elif adjoint:
            rc = lib.wavi_misfit_gradient_batch(ctxs, n, threads, n_picard, 1e-13, 200, up, vp, W._lib.dptr(Js),
                                                gp, reps, codes)
        else:
            rc = lib.wavi_update_velocities_batch(ctxs, n, threads, reps, codes)
        dt = time.perf_counter() - t0
        if rc != 0:
            bad = [k for k in range(n) if codes[k] != 0]
            raise RuntimeError("batch failed (%d) at members %s: %s"
                               % (rc, bad, lib.wavi_last_error(ctxs[bad[0]]) if bad else ""))
        return dt

    for k in range(warmup):
        log("[bench] ensemble warmup %d: %.3f s" % (k, run(nthreads)))
    times = []
    for k in range(steps):
        times.append(run(nthreads))
        log("[bench] ensemble step %d: %.3f s (%d threads)" % (k, times[-1], nthreads))
    def one_at_a_time():
        for m in members:
            assert lib.wavi_state_restore(m.parallel_spec.ctx) == 0
        t0 = time.perf_counter()
        for k, m in enumerate(members):
            ctx = m.parallel_spec.ctx
            if adjoint:
                rc = lib.wavi_misfit_gradient(ctx, n_picard, 1e-13, 200, up[k], vp[k], ctypes.byref(ctypes.c_double()),
                                              gp[k], ctypes.byref(reps[k]))
            else:
                rc = lib.wavi_update_velocities(ctx, ctypes.byref(reps[k]))
            assert rc == 0, lib.wavi_last_error(ctx)
        return time.perf_counter() - t0

    serial = None if compact else one_at_a_time()
    if serial is not None:
        log("[bench] ensemble one member at a time: %.3f s" % serial)
    total = float(sum(times))

    def velocities():
        out = []
        for sp, m in zip(specs, members):
            u = np.empty((sp["nx"] + 1) * sp["ny"])
            assert lib.wavi_get_field(m.parallel_spec.ctx, W._lib.FIELD["U_U"], W._lib.dptr(u), u.size) == 0
            out.append(u)
        return out

    lock = None
    if not adjoint and precision == "mixed" and not compact:
        # the same solves with every member in one lockstep group (the
        # batch itself puts only the small members in one)
        ref = velocities()
        run(nthreads, lockstep=True)
        same = all(np.array_equal(a, b) for a, b in zip(ref, velocities()))
        lt = []
        for k in range(steps):
            lt.append(run(nthreads, lockstep=True))
            log("[bench] ensemble lockstep step %d: %.3f s" % (k, lt[-1]))
        lock = {"members_per_s": n * steps / float(sum(lt)), "ms_per_step": 1e3 * float(sum(lt)) / steps,
                "over_threaded_batch": (total / steps) / (float(sum(lt)) / steps),
                "bit_identical_to_threaded_batch": bool(same)}
    unknowns = sum(2 * sp["nx"] * sp["ny"] + sp["nx"] + sp["ny"] for sp in specs)
    cfg = {"workload": "ensemble of %d channel %s, grids from %s, nsigma 4 or 8, A log-uniform in [1e-17, 4e-17], "
                       "C = 1e4 exp(GRF std 0.3)"
                       % (n, "forward + adjoint gradients (%d Picard iterations, solves to 1e-13)" % n_picard
                          if adjoint else "solves (20 Picard iterations)",
                          ["%dx%d" % s for s in W.workloads.ENSEMBLE_SIZES if not max_nx or s[0] <= max_nx]),
           "precision": precision, "host_threads": nthreads, "unknowns_total": unknowns}
    if serial is not None:
        cfg["serial_members_per_s"] = n / serial
        cfg["batched_over_serial"] = serial / (total / steps)
    if adjoint:
        cfg["unconverged_solves"] = int(sum(reps[k].unconverged for k in range(n)))
        cfg["max_rel_residual"] = float(max(max(reps[k].max_fwd_rel, reps[k].max_adj_rel) for k in range(n)))
    else:
        picard = [reps[k].picard_iters for k in range(n)]
        cfg["picard_iters_min_max"] = [min(picard), max(picard)]
        cfg["all_converged"] = all(reps[k].converged for k in range(n))
        if lock is not None:
            cfg["all_members_lockstep"] = lock
    return {
        "metric": "ensemble members per second (%s)" % ("forward + adjoint gradients" if adjoint
                                                        else "forward velocity solves"),
        "value": n * steps / total, "unit": "members/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * total / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if precision == "mixed" else "f64",
        "data": "synthetic (seeded ensemble of channel setups)", "config": cfg,
    }


def adjoint_run(W, nx, ny, steps, warmup, ndir=16, eps=1e-6, n_picard=10, tol=1e-13, maxit=200, tol_fd=1e-14,
                fd_at_solve_tol=True):
    """The basal-drag inversion gradient: observations from an FP64 forward
    at the true log C (log 1e4 + GRF(0.5, 10 km)) times (1 + 0.01 N(0, 1));
    dJ/dlog C at the prior log C = log 1e4 timed in FP64 and mixed precision
    (linear solves to tol), and checked against central differences (eps)
    along ndir seeded directions with each forward.  The differences use
    forward solves to tol_fd (about the FP64 residual floor), so that the
    rounding left in J resolves 1e-6 of the derivative; with fd_at_solve_tol
    they are also taken with solves to tol, to show that noise.  Returns the
    JSON line."""
    lib = W._lib.load()
    lt = W.workloads.inversion_logc_true(nx, ny, seed=4)
    log("[bench] adjoint: observations from the FP64 forward at the true log C")
    m = W.workloads.inversion_model(lt, "fp64", nσ=8, n_picard=n_picard)
    W.update_geometry(m)
    W.misfit_gradient(m, n_picard, np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1)), tol, maxit, gradient=False)
    uo, vo = W.workloads.inversion_observations(m.fields.gu.u, m.fields.gv.v, seed=4)
    uo = np.ascontiguousarray(uo.ravel(order="F"))
    vo = np.ascontiguousarray(vo.ravel(order="F"))
    del m
    l0 = np.log(1e4) * np.ones((nx, ny))
    rng = np.random.default_rng(16)
    dirs = [rng.standard_normal((nx, ny)) for _ in range(ndir)]
    out, grads, fds = {}, {}, {}
    for prec in ("fp64", "mixed"):
        m = W.workloads.inversion_model(l0, prec, nσ=8, n_picard=n_picard)
        W.update_geometry(m)
        spec = m.parallel_spec
        assert lib.wavi_state_save(spec.ctx) == 0
        J = ctypes.c_double(0.0)
        g = np.zeros(nx * ny)
        rep = W._lib.AdjointReport()

        def run(logc, grad, solve_tol=tol):
            assert lib.wavi_state_restore(spec.ctx) == 0
            c = np.ascontiguousarray(np.exp(logc).ravel(order="F"))
            assert lib.wavi_set_field(spec.ctx, W._lib.FIELD["P_WEERTMAN_C"], W._lib.dptr(c), c.size) == 0
            assert lib.wavi_update_geometry(spec.ctx) == 0
            t0 = time.perf_counter()
            rc = lib.wavi_misfit_gradient(spec.ctx, n_picard, solve_tol, maxit, W._lib.dptr(uo), W._lib.dptr(vo),
                                          ctypes.byref(J), W._lib.dptr(g) if grad else None, ctypes.byref(rep))
            dt = time.perf_counter() - t0
            if rc != 0:
                raise RuntimeError("wavi_misfit_gradient: %s" % lib.wavi_last_error(spec.ctx))
            return J.value, dt

        for k in range(warmup):
            log("[bench] adjoint %s warmup %d: %.3f s" % (prec, k, run(l0, True)[1]))
        times = []
        for k in range(steps):
            times.append(run(l0, True)[1])
            r = rep.as_dict()
            log("[bench] adjoint %s step %d: %.3f s (V-cycles %d forward, %d adjoint; max rel %.1e / %.1e)"
                % (prec, k, times[-1], r["fwd_vcycles"], r["adj_vcycles"], r["max_fwd_rel"], r["max_adj_rel"]))
        J0, r0 = J.value, rep.as_dict()
        grads[prec] = g.reshape((nx, ny), order="F").copy()
        _, tf = run(l0, False)
        fds[prec] = {}
        for t in (tol_fd, tol) if fd_at_solve_tol else (tol_fd,):
            fds[prec][t] = []
            for k, d in enumerate(dirs):
                fds[prec][t].append((run(l0 + eps * d, False, t)[0] - run(l0 - eps * d, False, t)[0]) / (2 * eps))
                if k % 4 == 3:
                    log("[bench] adjoint %s FD (forward solves to %.0e): %d of %d directions" % (prec, t, k + 1, ndir))
        out[prec] = {"ms": 1e3 * float(np.median(times)), "forward_ms": 1e3 * tf, "J": J0,
                     "grad_norm": float(np.linalg.norm(grads[prec])), "report": r0}
        del m

    def fd_err(gprec, fprec, t):
        errs = [abs(fd - float(np.sum(grads[gprec] * d))) / abs(float(np.sum(grads[gprec] * d)))
                for fd, d in zip(fds[fprec][t], dirs)]
        return {"max_rel": max(errs), "median_rel": float(np.median(errs))}

    for prec in ("fp64", "mixed") if ndir > 0 else ():
        for t in fds[prec]:
            out[prec]["fd_solves_%.0e" % t] = fd_err(prec, prec, t)
            log("[bench] adjoint %s FD check (own forward, eps %.0e): solves to %.0e max rel %.2e"
                % (prec, eps, t, out[prec]["fd_solves_%.0e" % t]["max_rel"]))
    if ndir > 0:
        out["mixed"]["gradient_vs_fp64_forward_fd"] = fd_err("mixed", "fp64", tol_fd)
    out["mixed"]["gradient_rel_diff_to_fp64"] = float(np.linalg.norm(grads["mixed"] - grads["fp64"]) /
                                                      np.linalg.norm(grads["fp64"]))
    cfg = {"workload": "basal-drag inversion gradient dJ/dlogC at %dx%d, nsigma 8, %d Picard iterations with "
                       "linear solves to %.0e (flexible CG, V-cycle preconditioner)" % (nx, ny, n_picard, tol),
           "fd_check": ("central differences at eps %.0e along %d seeded N(0,1) directions with each precision's "
                        "own forward, its solves to %s; the mixed gradient also against the FP64 "
                        "forward's differences" % (eps, ndir, " and to ".join("%.0e" % t for t in fds["fp64"])))
                       if ndir > 0 else None,
           "fp64": out["fp64"], "mixed": out["mixed"],
           "mixed_over_fp64": out["fp64"]["ms"] / out["mixed"]["ms"]}
    return {
        "metric": "adjoint ms: one dJ/dlogC (forward + reverse sweep) at %dx%dx8" % (nx, ny),
        "value": out["mixed"]["ms"], "unit": "ms", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": out["mixed"]["ms"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic (seeded true log C field and observation noise)", "config": cfg,
    }


def host_cpu():
    """The host CPU the CPU legs run on: model name, logical CPUs, and the
    CPUs this process may use."""
    name = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = None
    return {"cpu_model": name, "nproc": os.cpu_count(), "usable_cpus": usable}


def cpu_baseline(W, nx, ny, gpu_report, runs=5, caps=(20, 100)):
    """The oracle (a single-threaded C restatement of the reference algorithm)
    on a bounded sample of the same solve: the first Picard iteration with the
    coarse CG capped at caps[0] and at caps[1] iterations, each timed `runs`
    times from a fresh oracle state (about 30 s of CPU work on the MI355X
    host).  The medians of the two separate the per-Picard work (inner
    update, RHS, residual, diagonal probes, 42 smoothing sweeps, transfers)
    from the per-CG-iteration cost, and the full solve is projected with the
    FP64 GPU run's own Picard and CG counts (which the oracle reproduces
    exactly at the capped sizes and on every grid where it can run whole)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    med = {}
    for cap in caps:
        m = W.workloads.channel_solve(nx=nx, ny=ny, maxiter_picard=1)
        m.solver_params.maxiter_coarse = cap
        ts = []
        for r in range(runs):
            o = pyoracle.OracleModel(m)
            o.update_geometry()
            t0 = time.perf_counter()
            rep = o.update_velocities()
            ts.append(time.perf_counter() - t0)
            del o
        assert rep["coarse_cg_iters"][0] == cap
        med[cap] = (float(np.median(ts)), ts)
        log("[bench] cpu baseline: CG cap %d, %d runs, median %.2f s" % (cap, runs, med[cap][0]))
    (t1, _), (t2, _) = med[caps[0]], med[caps[1]]
    per_cg = (t2 - t1) / (caps[1] - caps[0])
    per_picard = t1 - caps[0] * per_cg
    total = gpu_report["picard_iters"] * per_picard + gpu_report["coarse_cg_total"] * per_cg
    return {
        "value": 1.0 / total, "unit": "solves/s", "cores": 1, "kind": "port", "host": host_cpu(),
        "runs": runs, "statistic": "median",
        "sample": "oracle (C restatement, FP64, one thread), first Picard iteration at %dx%dx8 with the coarse CG "
                  "capped at %d and %d iterations, %d runs each (medians %.2f s and %.2f s); per-Picard %.2f s and "
                  "per-CG-iteration %.4f s projected to the FP64 GPU run's %d Picard / %d CG iterations = %.0f s "
                  "per solve" % (nx, ny, caps[0], caps[1], runs, t1, t2, per_picard, per_cg,
                                 gpu_report["picard_iters"], gpu_report["coarse_cg_total"], total),
        "sample_times_s": {str(c): [round(t, 4) for t in med[c][1]] for c in caps},
    }


def transient_run(W, runs=5, nsteps=200):
    """The 80 x 10 channel transient (200 steps of 0.5 a, one Picard iteration
    per step, coarse CG to 1e-6): the whole run resident on the GPU
    (wavi_run_steps from a freshly created context, one warm-up run first) in
    FP64 and mixed precision, and the same run by the oracle on one host
    core; medians of `runs` runs.  The FP64 GPU thickness must equal the
    oracle's bit for bit."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    out, h = {}, {}
    for prec in ("fp64", "mixed"):
        ts = []
        for r in range(runs + 1):
            m = W.workloads.channel_transient(prec)
            m.params.dt = 0.5
            W.update_preconditioner(m)
            W.push_fields(m)
            spec = m.parallel_spec
            rep = W._lib.SolveReport()
            t0 = time.perf_counter()
            rc = spec._lib.wavi_run_steps(spec.ctx, nsteps, ctypes.byref(rep))
            dt = time.perf_counter() - t0
            if rc != 0:
                raise RuntimeError("wavi_run_steps: %s" % spec._lib.wavi_last_error(spec.ctx))
            if r > 0:
                ts.append(dt)
            hh = np.empty(m.grid.nx * m.grid.ny)
            assert spec._lib.wavi_get_field(spec.ctx, W._lib.FIELD["H_H"], W._lib.dptr(hh), hh.size) == 0
            h[prec] = hh
            spec.close()
        out["gpu_%s_s" % prec] = float(np.median(ts))
        log("[bench] transient 80x10 x %d steps, GPU %s: median %.4f s" % (nsteps, prec, out["gpu_%s_s" % prec]))
    ts = []
    for r in range(runs):
        m = W.workloads.channel_transient("fp64")
        m.params.dt = 0.5
        o = pyoracle.OracleModel(m)
        o.set_dt(0.5)
        t0 = time.perf_counter()
        o.run_steps(nsteps)
        ts.append(time.perf_counter() - t0)
    ho = o.field("H_H").ravel(order="F")
    out["cpu_oracle_s"] = float(np.median(ts))
    log("[bench] transient 80x10 x %d steps, CPU oracle (1 core): median %.4f s" % (nsteps, out["cpu_oracle_s"]))
    out["gpu_fp64_over_cpu"] = out["cpu_oracle_s"] / out["gpu_fp64_s"]
    out["gpu_mixed_over_cpu"] = out["cpu_oracle_s"] / out["gpu_mixed_s"]
    out["fp64_thickness_bit_equal_to_oracle"] = bool(np.array_equal(h["fp64"], ho))
    out["mixed_thickness_rel_l2_to_oracle"] = float(np.linalg.norm(h["mixed"] - ho) / np.linalg.norm(ho))
    out["workload"] = ("channel transient 80x10 at 8 km, nsigma 4, %d steps of 0.5 a, 1 Picard iteration per "
                       "step, coarse CG to 1e-6; medians of %d runs (GPU: after one warm-up run; CPU: the "
                       "oracle on one core)" % (nsteps, runs))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--nx", type=int, default=1280)
    ap.add_argument("--ny", type=int, default=160)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline solves only: no adjoint or transient legs (profiling runs)")
    ap.add_argument("--picard", type=int, default=None,
                    help="profiling runs only: cap the Picard iterations")
    ap.add_argument("--maxiter-coarse", type=int, default=None,
                    help="profiling runs only: cap the coarse CG iterations")
    ap.add_argument("--ensemble", type=int, default=0,
                    help="time the seeded ensemble of this many members instead of the single solve")
    ap.add_argument("--threads", type=int, default=8, help="host threads of the ensemble batch")
    ap.add_argument("--ensemble-max-nx", type=int, default=None,
                    help="ensemble of members at most this many cells wide (same seeded draw)")
    ap.add_argument("--with-adjoint", action="store_true",
                    help="with --ensemble: forward + adjoint gradient per member")
    ap.add_argument("--adjoint", action="store_true",
                    help="time the basal-drag gradient (FP64 and mixed) and check it by finite differences")
    ap.add_argument("--transient", action="store_true",
                    help="time only the 80x10 channel transient (200 steps, GPU FP64 and mixed, CPU oracle)")
    ap.add_argument("--fd-directions", type=int, default=16)
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="profiling runs only: issue one device-to-device copy of this size first")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the velocity solve here runs on a single MI355X; "
              "--gpus %d is not supported (multi-GPU sharding is a design note only)" % args.gpus)
        sys.exit(2)
    W = _package()
    if args.adjoint:
        print(json.dumps(adjoint_run(W, args.nx, args.ny, args.steps, args.warmup, args.fd_directions)), flush=True)
        return
    if args.transient:
        print(json.dumps(transient_run(W, runs=max(args.steps, 1))), flush=True)
        return
    if args.ensemble:
        prec = "fp64" if args.precision == "fp64" else "mixed"
        print(json.dumps(ensemble_run(W, args.ensemble, args.threads, args.steps, args.warmup, prec,
                                      args.with_adjoint, max_nx=args.ensemble_max_nx)), flush=True)
        return
    nx, ny = args.nx, args.ny
    if args.calibration_copy_mib:
        calibration_copy(args.calibration_copy_mib)
    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    res = {}
    for p in precs:
        res[p] = timed_solves(W, nx, ny, p, args.steps, args.warmup, args.picard, args.maxiter_coarse)
    head = "mixed" if "mixed" in res else "fp64"
    times, rep, vel, roof = res[head]
    total = float(sum(times))
    value = args.steps / total
    cfg = {"workload": "channel single velocity solve %dx%d, nsigma 8, %g m" % (nx, ny, 640000.0 / nx),
           "grid": [nx, ny, 8], "precision": head,
           "picard_iters": rep["picard_iters"], "coarse_cg_total": rep["coarse_cg_total"],
           "rel_resid": rep["rel_resid"], "median_s_per_solve": float(np.median(times)),
           "timed_region": "wavi_update_velocities only, no event timing (the reset of the state precedes it)"}
    if head == "mixed":
        cfg["precision"] = "mixed: FP32 V-cycle (operator, smoother, coarse CG), FP64 residual refinement"
        cfg["refinements"] = rep["refinements"]
    if "fp64" in res and "mixed" in res:
        t64, rep64, vel64, roof64 = res["fp64"]
        v64 = args.steps / float(sum(t64))
        err = vel - vel64
        cfg["fp64_solves_per_s"] = v64
        cfg["fp64_median_s_per_solve"] = float(np.median(t64))
        cfg["mixed_over_fp64"] = value / v64
        cfg["error_vs_fp64"] = {
            "rel_l2": float(np.linalg.norm(err) / np.linalg.norm(vel64)),
            "max_abs": float(np.max(np.abs(err))),
            "max_rel_to_max": float(np.max(np.abs(err)) / np.max(np.abs(vel64))),
            "bar": ERROR_BAR,
        }
        cfg["fp64_roofline"] = roof64
    valid = True
    if "error_vs_fp64" in cfg:
        e = cfg["error_vs_fp64"]
        valid = e["rel_l2"] <= ERROR_BAR and e["max_rel_to_max"] <= ERROR_BAR
    if not args.no_extras:
        log("[bench] adjoint gradient at %dx%dx8" % (nx, ny))
        adj = adjoint_run(W, nx, ny, 2, 1, ndir=4, fd_at_solve_tol=False)
        c = adj["config"]
        cfg["adjoint"] = {
            "workload": c["workload"], "fd_check": c["fd_check"],
            "mixed_ms": c["mixed"]["ms"], "fp64_ms": c["fp64"]["ms"], "mixed_over_fp64": c["mixed_over_fp64"],
            "fd_max_rel": {p: c[p]["fd_solves_1e-14"]["max_rel"] for p in ("fp64", "mixed")},
            "mixed_gradient_vs_fp64_forward_fd_max_rel": c["mixed"]["gradient_vs_fp64_forward_fd"]["max_rel"],
            "mixed_gradient_rel_diff_to_fp64": c["mixed"]["gradient_rel_diff_to_fp64"],
            "bar": 1e-5,
        }
        log("[bench] transient 80x10, 200 steps")
        cfg["transient_80x10"] = transient_run(W)
        log("[bench] ensemble of 64: forward solves, then forward + adjoint gradients")
        ef = ensemble_run(W, 64, 8, 2, 1, compact=True)
        ea = ensemble_run(W, 64, 8, 2, 1, adjoint=True, compact=True)
        cfg["ensemble_64"] = {
            "workload": ef["config"]["workload"] + "; then the same members' forward + adjoint gradients "
                        "(10 Picard iterations, solves to 1e-13); mixed precision, one batched call per step, "
                        "2 steps after 1 warm-up",
            "forward_members_per_s": ef["value"], "forward_s_per_batch": ef["ms_per_step"] / 1e3,
            "picard_iters_min_max": ef["config"]["picard_iters_min_max"],
            "gradient_members_per_s": ea["value"], "gradient_s_per_batch": ea["ms_per_step"] / 1e3,
            "gradient_unconverged_solves": ea["config"]["unconverged_solves"],
            "unknowns_total": ef["config"]["unknowns_total"],
        }
    out = {
        "metric": metric(nx, ny),
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic (seeded channel bed with Gaussian random-field roughness, analytic thickness)",
        "config": cfg,
        "roofline": roof,
        "valid": valid,
    }
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle sample)")
        out["cpu_baseline"] = cpu_baseline(W, nx, ny, res["fp64"][1] if "fp64" in res else rep)
    print(json.dumps(out), flush=True)
    if not valid:
        log("[bench] INVALID: mixed velocities beyond %.0e of FP64 (%s)" % (ERROR_BAR, cfg["error_vs_fp64"]))
        sys.exit(1)


if __name__ == "__main__":
    main()
