"""Benchmark: velocity solves per second of WAVI's update_velocities! on one
MI355X at 1280 x 160 x 8 (the channel single-solve workload: 20 Picard
iterations to 1e-8, coarse CG to 1e-8, 3 Haar levels), with the inputs
resident in HBM.

A step is one update_velocities! from a fixed reset state (velocities,
viscosity, drag, quadratures, warm start and wavelet masks restored on the
device before every solve, outside the timed interval).  Prints one JSON
line.  --precision mixed runs the FP32 V-cycle inside FP64 refinement and
also reports its velocity error against an FP64 solve of the same state.
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


def cg_apply_bytes(nx, ny, w=8):
    """Algorithmic HBM bytes of one k_cg_apply launch (A then R on a tile):
    t_u, t_v in; c_u, c_v out; u_u, u_v in for dot(u, c); beta_eff on u, v;
    thickness on u, v; h*eta_av on h; dimp on h; h*eta_av on c (w bytes
    each); masks (one byte each): inner u, v; wavelet u, v; u, v masks; h and
    c masks."""
    nu, nv, nh, nc = (nx + 1) * ny, nx * (ny + 1), nx * ny, (nx - 1) * (ny - 1)
    return w * (5 * (nu + nv) + 2 * nh + nc) + 3 * (nu + nv) + nh + nc


def cgm_iter_bytes(nx, ny, w=4):
    """Algorithmic HBM bytes of one k_cgm_iter launch (one whole iteration of
    the mixed-precision coarse CG): r, w, p, s, x on the tile and the images
    P r, P s, P w in; p, s, x, r, w and the three images out; beta_eff and
    thickness on u and v, h*eta_av on h and c, dimp on h (w bytes each);
    masks (one byte each): inner u, v; wavelet u, v; u, v masks; h and c
    masks."""
    nu, nv, nh, nc = (nx + 1) * ny, nx * (ny + 1), nx * ny, (nx - 1) * (ny - 1)
    return w * (18 * (nu + nv) + 2 * nh + nc) + 3 * (nu + nv) + nh + nc


DOMINANT = {  # precision -> (kernel, timing id, bytes per launch)
    "fp64": ("k_cg_apply", "cg_apply", lambda nx, ny: cg_apply_bytes(nx, ny, 8)),
    "mixed": ("k_cgm_iter", "cg_iter_mixed", lambda nx, ny: cgm_iter_bytes(nx, ny, 4)),
}


def measured_traffic(nx, ny, precision):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary for this workload, scaled by the
    copy-kernel calibration recorded there; None when no summary matches."""
    kernel = DOMINANT[precision][0]
    path = os.path.join(ROOT, "profiles", "%s_%dx%d_%s_traffic.md" % (kernel, nx, ny, precision))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch %s:" % kernel):
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


def solve(spec, lib):
    rep = W_lib.SolveReport()
    rc = lib.wavi_update_velocities(spec.ctx, ctypes.byref(rep))
    if rc != 0:
        raise RuntimeError("wavi_update_velocities failed: %s" % lib.wavi_last_error(spec.ctx))
    return rep.as_dict()


def velocities(W, m, spec):
    """u and v of the model's device state."""
    lib = spec._lib
    nx, ny = m.grid.nx, m.grid.ny
    u = np.empty((nx + 1) * ny)
    v = np.empty(nx * (ny + 1))
    assert lib.wavi_get_field(spec.ctx, W._lib.FIELD["U_U"], W._lib.dptr(u), u.size) == 0
    assert lib.wavi_get_field(spec.ctx, W._lib.FIELD["V_V"], W._lib.dptr(v), v.size) == 0
    return u, v


def cpu_baseline(W, nx, ny, gpu_report):
    """The oracle (a single-threaded C restatement of the reference algorithm)
    on a bounded sample of the same solve: the first Picard iteration with the
    coarse CG capped at 60 and at 120 iterations.  The two timings separate
    the per-Picard work (inner update, RHS, residual, diagonal probes, 42
    smoothing sweeps, transfers) from the per-CG-iteration cost, and the full
    solve is projected with the GPU run's own Picard and CG counts (which the
    oracle reproduces exactly)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    times = {}
    for cap in (60, 120):
        m = W.workloads.channel_solve(nx=nx, ny=ny, maxiter_picard=1)
        m.solver_params.maxiter_coarse = cap
        o = pyoracle.OracleModel(m)
        o.update_geometry()
        t0 = time.perf_counter()
        rep = o.update_velocities()
        times[cap] = (time.perf_counter() - t0, rep["coarse_cg_iters"][0])
    (t1, n1), (t2, n2) = times[60], times[120]
    per_cg = (t2 - t1) / max(n2 - n1, 1)
    per_picard = t1 - n1 * per_cg
    total = gpu_report["picard_iters"] * per_picard + gpu_report["coarse_cg_total"] * per_cg
    return {
        "value": 1.0 / total, "unit": "solves/s", "cores": 1, "kind": "port",
        "sample": "oracle, first Picard iteration at %dx%dx8 with the coarse CG capped at 60 and 120 "
                  "iterations (%.1f s + %.1f s); per-Picard %.2f s and per-CG-iteration %.4f s "
                  "projected to %d Picard / %d CG iterations = %.0f s per solve"
                  % (nx, ny, t1, t2, per_picard, per_cg, gpu_report["picard_iters"],
                     gpu_report["coarse_cg_total"], total),
    }


def main():
    global W_lib
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--nx", type=int, default=1280)
    ap.add_argument("--ny", type=int, default=160)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--picard", type=int, default=None,
                    help="profiling runs only: cap the Picard iterations")
    ap.add_argument("--maxiter-coarse", type=int, default=None,
                    help="profiling runs only: cap the coarse CG iterations")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="profiling runs only: issue one device-to-device copy of this size first")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the velocity solve here runs on a single MI355X; "
              "--gpus %d is not supported (multi-GPU sharding is a design note only)" % args.gpus)
        sys.exit(2)
    W = _package()
    W_lib = W._lib
    nx, ny = args.nx, args.ny
    log("[bench] setup %dx%dx8 precision=%s" % (nx, ny, args.precision))
    m, spec, lib = setup(W, nx, ny, args.precision, args.picard, args.maxiter_coarse)
    if args.calibration_copy_mib:
        calibration_copy(args.calibration_copy_mib)
    for k in range(args.warmup):
        assert lib.wavi_state_restore(spec.ctx) == 0
        t0 = time.perf_counter()
        rep = solve(spec, lib)
        log("[bench] warmup %d: %.3f s, picard %d, cg %d, rel %.3e" % (
            k, time.perf_counter() - t0, rep["picard_iters"], rep["coarse_cg_total"], rep["rel_resid"]))
    lib.wavi_profile_enable(spec.ctx, 1)
    times = []
    for k in range(args.steps):
        assert lib.wavi_state_restore(spec.ctx) == 0
        t0 = time.perf_counter()
        rep = solve(spec, lib)
        dt = time.perf_counter() - t0
        times.append(dt)
        log("[bench] step %d: %.3f s, picard %d, cg %d, rel %.3e" % (
            k, dt, rep["picard_iters"], rep["coarse_cg_total"], rep["rel_resid"]))
    ms = ctypes.c_double(0.0)
    n = ctypes.c_int64(0)
    kernel, kid, nbytes = DOMINANT[args.precision]
    lib.wavi_profile_read(spec.ctx, W._lib.KT[kid], ctypes.byref(ms), ctypes.byref(n))
    lib.wavi_profile_enable(spec.ctx, 0)
    avg_s = ms.value / 1e3 / max(n.value, 1)
    B = nbytes(nx, ny)
    achieved = B / avg_s / 1e9
    total = float(sum(times))
    value = args.steps / total
    out = {
        "metric": "velocity solves/s at 1280x160x8 (update_velocities!, 20 Picard iterations, coarse CG tol 1e-8)",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded channel bed with Gaussian random-field roughness, analytic thickness)",
        "config": {"workload": "channel single velocity solve %dx%d, nsigma 8, 500 m" % (nx, ny),
                   "grid": [nx, ny, 8], "precision": args.precision,
                   "picard_iters": rep["picard_iters"], "coarse_cg_total": rep["coarse_cg_total"],
                   "rel_resid": rep["rel_resid"]},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(nx, ny, args.precision),
                     "bytes_per_launch": B, "avg_launch_us": avg_s * 1e6, "launches": n.value},
    }
    if args.precision == "mixed":
        log("[bench] FP64 solve of the same state for the error of the mixed solve")
        mu, mv = velocities(W, m, spec)
        m64, spec64, lib64 = setup(W, nx, ny, "fp64", args.picard, args.maxiter_coarse)
        solve(spec64, lib64)
        ru, rv = velocities(W, m64, spec64)
        ref = np.concatenate([ru.ravel(), rv.ravel()])
        got = np.concatenate([mu.ravel(), mv.ravel()])
        out["config"]["error_vs_fp64"] = {
            "rel_l2": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
            "max_abs": float(np.max(np.abs(got - ref))),
            "max_rel_to_max": float(np.max(np.abs(got - ref)) / np.max(np.abs(ref))),
        }
        out["config"]["refinements"] = rep["refinements"]
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle sample)")
        out["cpu_baseline"] = cpu_baseline(W, nx, ny, rep)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
