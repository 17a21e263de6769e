"""Benchmark: P2-P1 Navier-Stokes Jacobian assembly throughput on one MI355X.

Workload: FEDDLib's FE::assemblyNavierStokes(..., reAssemble=false, "Jacobian") --
all blocks A(0,0), B^T(0,1), B(1,0) (C(1,1) is empty) -- on the unit cube with
40^3 x 6 = 384,000 tets (P2 velocity / P1 pressure), nu = 0.01, rho = 1,
Linearization "Newton", velocity iterate = 8 seeded sine modes of amplitude 1.
A "step" is one such assembly with the inputs resident in HBM.

Second leg (the metric's other half): Neo-Hookean Newton time-to-||r||<1e-8 on the
clamped unit cube 24^3 x 6 tets (P2, skr2 law, E = 1000, nu = 0.3), uniform traction
(40, 0, 60) on z = 1 ramped in 5 load steps, with the exact tangent and with the
learned surrogate (trained on the device beforehand; training time reported apart).

Prints ONE JSON line (rank 0).  Usage:
  python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--M 40] [--cpu-sample 50000]
                  [--newton-M 24] [--no-newton]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
COPY_BYTES = 1 << 30
NU, RHO = 0.01, 1.0


TRAFFIC_FILE = "profiles/ns_vel_rows_M40_hbm_counters.txt"


def profiled_traffic():
    """HBM bytes per launch of k_ns_vel_rows<double> at 40^3x6 tets, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


FLOP_FILE = "profiles/assembly_valu_flop_counters.txt"


def profiled_fp64_flops(kernel="fedd::k_ns_vel_rows<double>"):
    """FP64 flops per launch of `kernel`, from the committed rocprofv3
    SQ_INSTS_VALU_FLOPS_FP64 summary (third column of the kernel's line)."""
    path = os.path.join(ROOT, FLOP_FILE)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.lstrip("# ").split()
        if f and f[0] == kernel:
            return float(f[2])
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # Untimed steps before the timed ones (after the clock-settling batches of settle()).
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--M", type=int, default=40)
    ap.add_argument("--cpu-sample", type=int, default=50000,
                    help="elements of the CPU-baseline sample (0 disables it)")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--newton-M", type=int, default=24)
    ap.add_argument("--no-newton", action="store_true")
    ap.add_argument("--no-scatter", action="store_true", help="skip the coloured-scatter comparison leg")
    ap.add_argument("--newton-exact-only", action="store_true",
                    help="run only the exact-tangent solves of the Newton leg (profiling runs)")
    ap.add_argument("--large-M", type=int, default=64,
                    help="cells per edge of the graph-replayed FP32-local Jacobian + residual leg (0: skip)")
    ap.add_argument("--elas-M", type=int, default=32,
                    help="cells per edge of the hyperelastic assembly leg, jittered P2 cube (0: skip)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the STREAM copy kernel (1 GiB read + 1 GiB write, 20 times): the "
                         "attainable bandwidth and the calibration of rocprofv3 byte counters")
    ap.add_argument("--tangent-probe", action="store_true",
                    help="only run the learned-tangent assembly at 24^3x6 five times (surrogate inference "
                         "k_mlp_tangent on 414,720 points, no multigrid graphs): for rocprofv3 counter passes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (the elements would be "
              "partitioned into sub-cubes, one per GPU, with one interface-row exchange per assembly); "
              "run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from feddlib_amd import native as N
    from feddlib_amd import workloads as W

    if args.copy_probe:
        with N.Context(0) as ctx:
            ms = ctx.stream_copy(COPY_BYTES, 20)
        print(json.dumps({"probe": "stream_copy", "bytes_read": COPY_BYTES, "bytes_written": COPY_BYTES,
                          "launches": 21, "ms_per_copy": ms, "GBs": 2 * COPY_BYTES / (ms * 1e-3) / 1e9}))
        return

    if args.tangent_probe:
        print(json.dumps(tangent_probe(N)))
        return

    M = args.M
    t0 = time.time()
    p2, e2, p1, e1, u, p = W.ns_problem(M, 0)
    E = e2.shape[0]
    ctx = N.Context(0)
    ctx.mesh_upload(N.MESH_PRIMARY, p2, e2)
    ctx.mesh_upload(N.MESH_PRESSURE, p1, e1)
    ctx.graph_build(N.GRAPH_NS)
    ctx.vector_upload(N.VEC_U, u)
    ctx.vector_upload(N.VEC_P, p)
    setup_s = time.time() - t0
    nnzA = ctx.block_info(N.BLOCK_A)[2]
    nnzBT = ctx.block_info(N.BLOCK_BT)[2]
    nnzB = ctx.block_info(N.BLOCK_B)[2]
    print("# setup %.1fs  E=%d  nnz A=%d BT=%d B=%d" % (setup_s, E, nnzA, nnzBT, nnzB), file=sys.stderr,
          flush=True)

    def step(precision):
        ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN, reassemble=False, precision=precision)

    def settle(precision, batch=10, max_steps=400):
        """Untimed assemblies until the GPU has left its idle clocks: batches of `batch`
        steps until one takes within 1% of the batch before (at least 30 steps).  The first
        launches after idle run 10-20% slower (kernel trace: 570 -> 460 us over about 50
        launches); the steady state is what continuous assembly sees."""
        prev, n = None, 0
        while n < max_steps:
            ctx.synchronize()
            ctx.event_record(0)
            for _ in range(batch):
                step(precision)
            ctx.event_record(1)
            ctx.synchronize()
            t = ctx.event_elapsed(0, 1)
            n += batch
            if prev is not None and n >= 30 and abs(t - prev) <= 0.01 * t:
                break
            prev = t
        return n

    def timed(precision, steps, warmup):
        for _ in range(warmup):
            step(precision)
        ctx.synchronize()
        ctx.event_record(0)
        for _ in range(steps):
            step(precision)
        ctx.event_record(1)
        ctx.synchronize()
        ms = ctx.event_elapsed(0, 1)
        # Per-kernel durations from a second pass with kernel timing on (events around
        # each launch; the rows of B run beside the velocity rows as in the timed pass).
        ctx.kernel_timing(True)
        for _ in range(steps):
            step(precision)
        ctx.synchronize()
        kv = ctx.kernel_time("ns_vel_rows")
        kp = ctx.kernel_time("ns_pres_rows")
        ctx.kernel_timing(False)
        return ms / steps, kv, kp

    settle_steps = settle(N.FP64)
    print("# clocks settled after %d untimed steps" % settle_steps, file=sys.stderr, flush=True)
    ms64, kv64, kp64 = timed(N.FP64, args.steps, args.warmup)
    blocks = (("A", N.BLOCK_A), ("BT", N.BLOCK_BT), ("B", N.BLOCK_B))
    vals64 = {k: ctx.block_values(b) for k, b in blocks} if not args.no_check else None
    settle32 = settle(N.FP32_LOCAL)  # the FP64 values' download above left the GPU idle
    ms32, kv32, kp32 = timed(N.FP32_LOCAL, args.steps, args.warmup)
    fp32_errors = None
    gatherA = None
    if vals64 is not None:
        fp32_errors = {k: fp32_error_figures(ctx.block_values(b), vals64[k], ctx.block_graph(b)[0])
                       for k, b in blocks}
        gatherA = vals64["A"]
        del vals64
        print("# fp32-local vs fp64 (row floor / per matrix): " + ", ".join(
            "%s %.2e / %.2e" % (k, e["row_floor"], e["per_matrix"]) for k, e in fp32_errors.items()),
            file=sys.stderr, flush=True)
    print("# fp64 %.3f ms/step  fp32-local %.3f ms/step" % (ms64, ms32), file=sys.stderr, flush=True)

    # Algorithmic bytes (unique data, one pass): values written once, connectivity,
    # iterate and vertex coordinates read once.
    n2, n1 = p2.shape[0], p1.shape[0]
    bytes_vel = 8 * (nnzA + nnzBT) + 4 * 10 * E + 8 * 3 * n2 + 24 * n1
    bytes_step = 8 * (nnzA + nnzBT + nnzB) + 4 * (10 + 4) * E + 8 * 3 * n2 + 24 * n1
    vel_avg_ms = kv64[0] / max(kv64[1], 1)
    achieved = bytes_vel / (vel_avg_ms * 1e-3) / 1e9

    traffic = profiled_traffic() if M == 40 else None
    flops = profiled_fp64_flops() if M == 40 else None

    cpu = None
    if args.cpu_sample > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        n = min(args.cpu_sample, E)
        secs1 = O.time_ns_jacobian(p2, e2, e1, n1, u, p, NU, RHO, n)
        threads, allowed, share = host_threads()
        secs = O.time_ns_jacobian_mt(p2, e2, e1, n1, u, p, NU, RHO, threads)
        cpu = {"value": E / secs, "unit": "elems/s", "cores": threads, "kind": "port",
               "cores_note": "%d threads: the %d cores in this process's affinity mask%s" % (
                   threads, allowed, "" if share is None else ", limited to the host's CPU share for one GPU "
                                                              "(OMP_NUM_THREADS=%d)" % share),
               "sample": "all %d elements of the same workload on %d host threads (element partitions "
                         "assembled concurrently, then merged row by row like the reference's ranks + "
                         "export-Add): FEDDLib's element loop restated in C++ (per-element quadrature "
                         "tables, 34x34 element matrix, row-wise insertion, Tpetra-style fillComplete); "
                         "%.1f s on %s" % (E, threads, secs, cpu_model()),
               "single_core": {"value": n / secs1, "cores": 1,
                               "sample": "first %d elements, %.1f s" % (n, secs1)}}
        print("# cpu baseline: %.0f elems/s on %d threads, %.0f on 1" % (E / secs, threads, n / secs1),
              file=sys.stderr, flush=True)

    out = {
        "metric": "P2-P1 NS Jacobian elems/s; Neo-Hookean Newton time-to-1e-8, MLP vs exact tangent",
        "value": E / (ms64 * 1e-3),
        "unit": "elems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle_steps": settle_steps,
        "ms_per_step": ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "FE::assemblyNavierStokes Jacobian (A, B^T, B; Newton linearisation), P2-P1 "
                        "Taylor-Hood, unit cube %d^3x6 = %d tets, nu=0.01 rho=1, u = 8 seeded sine modes"
                        % (M, E),
            "tets": E, "velocity_dofs": 3 * n2, "pressure_dofs": n1,
            "nnz": {"A": nnzA, "BT": nnzBT, "B": nnzB, "C": 0},
        },
        "roofline": {
            "bound": "hbm", "kernel": "k_ns_vel_rows<double>", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_source": TRAFFIC_FILE if traffic else None,
            "algorithmic_bytes_per_launch": bytes_vel, "avg_launch_ms": vel_avg_ms,
            "fp64_vector": None if flops is None else {
                "flop_per_launch": flops, "tflops": flops / (vel_avg_ms * 1e-3) / 1e12,
                "peak_tflops": FP64_PEAK_TFLOPS, "frac": flops / (vel_avg_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                "source": FLOP_FILE},
        },
        "step": {
            "algorithmic_bytes": bytes_step,
            "achieved_GBs": bytes_step / (ms64 * 1e-3) / 1e9,
            "frac": bytes_step / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "pres_rows_avg_ms": kp64[0] / max(kp64[1], 1),
        },
        "fp32_local": {
            "value": E / (ms32 * 1e-3), "ms_per_step": ms32, "settle_steps": settle32,
            "errors_vs_fp64": fp32_errors,
            "errors_note": FP32_NOTE,
            "vel_rows_avg_ms": kv32[0] / max(kv32[1], 1),
        },
        "cpu_baseline": cpu,
    }
    if not args.no_scatter:
        out["colour_scatter"] = scatter_leg(ctx, N, (u, p), E, ms64, bytes_step, gatherA, args.steps, args.warmup)
    del gatherA
    if not args.no_newton:
        out["newton"] = newton_leg(ctx, N, args.newton_M, exact_only=args.newton_exact_only)
    if args.large_M > 0:
        out["ns_large"] = large_leg(ctx, N, W, args.large_M, check=not args.no_check)
    if args.elas_M > 0:
        out["elasticity"] = elasticity_leg(ctx, N, W, args.elas_M)
    ctx.close()
    print(json.dumps(out))


def host_threads():
    """Threads for the CPU baseline: every core in the affinity mask, limited to the CPU
    share the host gives one GPU's jobs when it states one (OMP_NUM_THREADS)."""
    allowed = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS")
    share = int(share) if share and share.isdigit() and int(share) > 0 else None
    return (min(allowed, share) if share else allowed), allowed, share


FP32_NOTE = ("per block against the FP64 assembly of the same iterate (equal to the oracle to 1e-12): "
             "row_floor = max|a32-a64| / max(|a64|, 1e-3 row max); per_matrix = max|a32-a64| / max|a64|; "
             "rowwise = max over rows of max|a32-a64| / row max.  Rounding the exact element matrices to "
             "FP32 alone gives row_floor 3.3e-5..3.9e-5 on A (4^3..16^3 x 6), per_matrix < 1e-7")


def fp32_error_figures(v32, v64, rowptr, chunk=1 << 25):
    """Row-floor, per-matrix and row-wise relative errors of v32 against v64 on one CRS
    pattern, over row ranges of about `chunk` entries."""
    nr = len(rowptr) - 1
    scale = float(np.max(np.abs(v64))) if v64.size else 1.0
    rf = pm = rw = 0.0
    r0 = 0
    while r0 < nr:
        r1 = int(min(nr, np.searchsorted(rowptr, rowptr[r0] + chunk, side="right")))
        r1 = max(r1, r0 + 1)
        a, b = rowptr[r0], rowptr[r1]
        if b > a:
            x = np.abs(v64[a:b])
            d = np.abs(v32[a:b] - v64[a:b])
            lens = np.diff(rowptr[r0:r1 + 1])
            nz = lens > 0
            rmax = np.zeros(r1 - r0)
            rmax[nz] = np.maximum.reduceat(x, (rowptr[r0:r1][nz] - a))
            rm = np.repeat(rmax, lens)
            den = np.maximum(x, 1e-3 * rm)
            den[den == 0] = 1.0
            rf = max(rf, float(np.max(d / den)))
            pm = max(pm, float(np.max(d)) / (scale if scale > 0 else 1.0))
            rm[rm == 0] = 1.0
            rw = max(rw, float(np.max(d / rm)))
        r0 = r1
    return {"row_floor": rf, "per_matrix": pm, "rowwise": rw}


def tangent_probe(N, M=24, reps=5):
    """The learned-tangent assembly (exact residual, surrogate tangent) on the Newton leg's
    24^3x6 P2 cube at a smooth displacement, with a seeded random network: every launch of
    k_mlp_tangent infers CC at all 5 points of every element.  Graph-free, so rocprofv3's
    counter passes can serialise every launch."""
    pts, el, _ = _nh_problem(N, M)
    with N.Context(0) as ctx:
        ctx.mlp_init(1, EMOD, POISSON, 0)
        ctx.mlp_set_params(np.random.default_rng(0).normal(0.0, 0.1, N.MLP_PARAMS).astype(np.float32))
        ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
        ctx.graph_build(N.GRAPH_VECTOR)
        x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
        ctx.vector_upload(N.VEC_U, 0.1 * np.stack([z * x, -0.5 * z * y, 0.3 * x * y], axis=1).ravel())
        ctx.assembly_nonlinear_elasticity_tangent(1, EMOD, POISSON, N.FP64, tangent=1)
        ctx.kernel_timing(True)
        for _ in range(reps):
            ctx.assembly_nonlinear_elasticity_tangent(1, EMOD, POISSON, N.FP64, tangent=1)
        ctx.synchronize()
        t, n = ctx.kernel_time("mlp_tangent")
        ctx.kernel_timing(False)
    nqp = 5 * el.shape[0]
    ms = t / max(n, 1)
    return {"probe": "mlp_tangent", "points_per_launch": nqp, "launches": n, "avg_ms": ms,
            "tflops": 12032.0 * nqp / (ms * 1e-3) / 1e12}


def scatter_leg(ctx, N, iterate, E, ms_gather, bytes_step, gatherA, steps, warmup):
    """The same FP64 Jacobian (A, B^T, B) by the coloured element scatter: the NS graph
    rebuilt with FEDD_METHOD_COLOUR_SCATTER (meshes as uploaded), timed like the main
    metric; values against the row gather's."""
    u, p = iterate
    ctx.set_assembly_method(N.METHOD_COLOUR_SCATTER)
    t0 = time.time()
    ctx.graph_build(N.GRAPH_NS)
    build_s = time.time() - t0
    ctx.vector_upload(N.VEC_U, u)
    ctx.vector_upload(N.VEC_P, p)
    for _ in range(warmup):
        ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN, reassemble=False)
    ctx.synchronize()
    ctx.event_record(0)
    for _ in range(steps):
        ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN, reassemble=False)
    ctx.event_record(1)
    ctx.synchronize()
    ms = ctx.event_elapsed(0, 1) / steps
    diff = None
    if gatherA is not None:
        diff = float(np.max(np.abs(ctx.block_values(N.BLOCK_A) - gatherA)) / np.max(np.abs(gatherA)))
    ncol = ctx.colour_build()
    ctx.set_assembly_method(N.METHOD_ROW_GATHER)
    print("# coloured scatter %.3f ms/step (%d colours), row gather %.3f" % (ms, ncol, ms_gather), file=sys.stderr,
          flush=True)
    return {"workload": "the main metric's FP64 Jacobian (A, B^T, B) by the coloured element scatter",
            "colours": ncol, "ms_per_step": ms, "value": E / (ms * 1e-3), "unit": "elems/s",
            "scatter_over_row_gather_time": ms / ms_gather,
            "achieved_GBs_on_algorithmic_bytes": bytes_step / (ms * 1e-3) / 1e9,
            "schedule_build_s": build_s, "max_abs_diff_A_vs_row_gather_over_max": diff}


def elasticity_leg(ctx, N, W, M, reps=10):
    """FE::assemblyNonLinearElasticity (skr2 law, tangent + residual) on the jittered P2 cube
    with M^3 x 6 tets and the smoothed seeded displacement, FP64 and FP32-local: elements/s
    of the whole assembly (quadrature-point kernel + row gather), inputs resident."""
    pts, _, el = W.jittered_p2_mesh(M, 0)
    u = W.smoothed_displacement(pts, el, 1)
    E = el.shape[0]
    ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    ctx.graph_build(N.GRAPH_VECTOR)
    ctx.vector_upload(N.VEC_U, u.ravel())
    nnz = ctx.block_info(N.BLOCK_A)[2]
    n = pts.shape[0]
    # values written once, connectivity (10 ids), coordinates and displacement read once,
    # residual written once
    alg = 8 * nnz + 4 * 10 * E + 24 * n + 24 * n + 24 * n
    res = {"workload": "FE::assemblyNonLinearElasticity, Neo-Hooke (skr2) tangent + residual, unit cube %d^3x6 = "
                       "%d P2 tets, interior vertices jittered by up to 0.2 h, displacement 0.05 N(0,1) smoothed once, "
                       "E=1000 nu=0.3" % (M, E), "tets": E, "nnz": nnz, "algorithmic_bytes": alg}
    vals = {}
    for name, prec in (("fp64", N.FP64), ("fp32_local", N.FP32_LOCAL)):
        for _ in range(2):
            ctx.assembly_nonlinear_elasticity(1, EMOD, POISSON, prec)
        ctx.synchronize()
        ctx.event_record(0)
        for _ in range(reps):
            ctx.assembly_nonlinear_elasticity(1, EMOD, POISSON, prec)
        ctx.event_record(1)
        ctx.synchronize()
        ms = ctx.event_elapsed(0, 1) / reps
        # the reference element (AceGen skr2: law and B^T CC B at 5 points) costs about
        # 5e4 flop per element; against the FP64 vector peak
        res[name] = {"ms": ms, "elems_per_s": E / (ms * 1e-3),
                     "hbm_frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "fp64_tflops_at_5e4_flop_per_elem": 5e4 * E / (ms * 1e-3) / 1e12,
                     "fp64_frac": 5e4 * E / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
        vals[name] = (ctx.block_values(N.BLOCK_A), ctx.vector_download(N.VEC_RES_U, 3 * n))
        if name == "fp32_local":
            chunks, pairs = ctx.elasticity_fp64_counts()
            res[name]["chunks_reassembled_in_fp64"] = chunks
            res[name]["pairs_of_flagged_elements_in_fp64"] = pairs
            res[name]["pairs_total"] = 10 * E
    rp, _ = ctx.block_graph(N.BLOCK_A)
    v64, r64 = vals["fp64"]
    v32, r32 = vals["fp32_local"]
    fin = np.isfinite(v64)
    res["fp32_local"]["errors_vs_fp64"] = dict(
        fp32_error_figures(v32, v64, rp), nonfinite_fp32=int(np.sum(~np.isfinite(v32[fin]))),
        residual=float(np.max(np.abs(r32 - r64)) / np.max(np.abs(r64))))
    res["fp32_local"]["errors_note"] = ("A against the FP64 assembly of the same iterate (row_floor, per_matrix, "
                                        "rowwise as for the NS blocks); the residual is FP64 in both paths.  At "
                                        "this iterate grad u reaches 3 and many quadrature points have det F <= 0; "
                                        "elements with a point of kappa(F^T F) > 1e4 are computed in FP64 "
                                        "(pairs_of_flagged_elements_in_fp64), which the FP32-local time includes")
    del vals, v64, v32
    print("# elasticity %d tets: %.3f ms fp64, %.3f ms fp32-local" % (E, res["fp64"]["ms"], res["fp32_local"]["ms"]),
          file=sys.stderr, flush=True)
    return res


def large_leg(ctx, N, W, M, iters=10, check=True):
    """FP32-local Jacobian + FP64 residual of the NS problem at M^3 x 6 tets in one fused
    pass (FEDD_NS_JACOBIAN_RHS), captured once into a HIP graph and replayed for `iters`
    Newton-like iterate updates (the iterate is updated between replays by an upload)."""
    p2, e2, p1, e1, u, p = W.ns_problem(M, 0)
    E = e2.shape[0]
    ctx.mesh_upload(N.MESH_PRIMARY, p2, e2)
    ctx.mesh_upload(N.MESH_PRESSURE, p1, e1)
    ctx.graph_build(N.GRAPH_NS)
    ctx.vector_upload(N.VEC_U, u)
    ctx.vector_upload(N.VEC_P, p)
    for mode in (N.NS_JACOBIAN, N.NS_RHS, N.NS_JACOBIAN_RHS):
        ctx.assembly_navier_stokes(NU, RHO, True, mode, reassemble=False, precision=N.FP32_LOCAL)
    ctx.capture_begin()
    ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN_RHS, reassemble=False, precision=N.FP32_LOCAL)
    nodes = ctx.capture_end()
    ctx.graph_launch()
    ctx.synchronize()
    ms = 0.0
    for it in range(iters):
        ctx.vector_upload(N.VEC_U, u * (1.0 + 0.05 * it))
        ctx.event_record(0)
        ctx.graph_launch()
        ctx.event_record(1)
        ctx.synchronize()
        ms += ctx.event_elapsed(0, 1)
    ms /= iters
    eager = {}
    for name, mode in (("jacobian_ms", N.NS_JACOBIAN), ("residual_ms", N.NS_RHS),
                       ("fused_jacobian_residual_ms", N.NS_JACOBIAN_RHS)):
        ctx.event_record(0)
        for _ in range(3):
            ctx.assembly_navier_stokes(NU, RHO, True, mode, reassemble=False, precision=N.FP32_LOCAL)
        ctx.event_record(1)
        ctx.synchronize()
        eager[name] = ctx.event_elapsed(0, 1) / 3
    blocks = (("A", N.BLOCK_A), ("BT", N.BLOCK_BT), ("B", N.BLOCK_B))
    nnz = sum(ctx.block_info(b)[2] for _, b in blocks)
    n2, n1 = p2.shape[0], p1.shape[0]
    alg = 8 * nnz + 4 * 14 * E + 8 * (3 * n2 + n1) + 24 * (n2 + n1) + 8 * (3 * n2 + n1)
    print("# large %d tets: %.3f ms per replay" % (E, ms), file=sys.stderr, flush=True)
    # FP32-local errors at the last iterate: the eager fused pass above left its values
    errors = None
    if check:
        v32 = {k: ctx.block_values(b) for k, b in blocks}
        ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN, reassemble=False, precision=N.FP64)
        errors = {}
        for k, b in blocks:
            errors[k] = fp32_error_figures(v32.pop(k), ctx.block_values(b), ctx.block_graph(b)[0])
        print("# large fp32-local vs fp64 (row floor / per matrix): " + ", ".join(
            "%s %.2e / %.2e" % (k, e["row_floor"], e["per_matrix"]) for k, e in errors.items()),
            file=sys.stderr, flush=True)
    return {"workload": "FP32-local NS Jacobian (A, B^T, B) + FP64 residual in one fused pass, %d^3x6 = %d tets, "
                        "HIP graph (%d nodes) replayed over %d iterate updates" % (M, E, nodes, iters),
            "tets": E, "ms_per_iteration": ms, "elems_per_s": E / (ms * 1e-3), "eager": eager,
            "algorithmic_bytes": alg, "achieved_GBs": alg / (ms * 1e-3) / 1e9,
            "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "fp32_local_errors_vs_fp64": errors}


EMOD, POISSON = 1000.0, 0.3
TRACTION = np.array([40.0, 0.0, 60.0])


def _nh_problem(N, M):
    pts, _, el = N.mesh_structured(M, 2)
    el = np.ascontiguousarray(el)
    N.correct_orientation(pts, el)
    base = np.where(np.abs(pts[:, 2]) < 1e-12)[0]
    dofs = np.sort((3 * base[:, None] + np.arange(3)).ravel()).astype(np.int32)
    return pts, el, dofs


def _nh_setup(ctx, N, prob):
    pts, el, dofs = prob
    ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    ctx.graph_build(N.GRAPH_VECTOR)
    ctx.vector_upload(N.VEC_U, np.zeros(3 * pts.shape[0]))
    ctx.dirichlet_set(dofs)
    ctx.traction_load(2, 1.0, TRACTION)


def _nh_solve(ctx, N, prob, tangent, collect=False, precond=1):
    _nh_setup(ctx, N, prob)
    st, hist = ctx.newton_elasticity(model=1, E=EMOD, nu=POISSON, load_steps=5, abs_tol=1e-8, max_newton=40,
                                     tangent=tangent, collect_samples=collect, preconditioner=precond,
                                     mg_reuse=precond == 1)
    return st, ctx.vector_download(N.VEC_U, 3 * prob[0].shape[0])


def newton_leg(ctx, N, M, exact_only=False):
    """Time-to-tolerance of the load-stepped Newton solve, exact vs learned tangent."""
    prob = _nh_problem(N, M)
    nelem, nnode = prob[1].shape[0], prob[0].shape[0]
    keys = ("seconds", "newton_iters", "krylov_iters", "assembly_ms", "solve_ms", "final_residual", "converged",
            "fallbacks", "rejected")
    _nh_solve(ctx, N, prob, 0)  # warm-up: allocations, code objects
    ctx.kernel_timing(True)
    st_e, u_e = _nh_solve(ctx, N, prob, 0)
    k_qp, k_rows = ctx.kernel_time("elas_qp"), ctx.kernel_time("elas_rows")
    ctx.kernel_timing(False)
    print("# newton exact %.3fs (%d Newton, %d Krylov)" % (st_e["seconds"], st_e["newton_iters"],
                                                          st_e["krylov_iters"]), file=sys.stderr, flush=True)
    nnz = ctx.block_info(N.BLOCK_A)[2]
    _nh_solve(ctx, N, prob, 0, precond=0)
    st_bj, u_bj = _nh_solve(ctx, N, prob, 0, precond=0)
    print("# newton exact, block-Jacobi CG %.3fs (%d Krylov)" % (st_bj["seconds"], st_bj["krylov_iters"]),
          file=sys.stderr, flush=True)

    if exact_only:
        return {"exact": {k: st_e[k] for k in keys}, "exact_block_jacobi_cg": {k: st_bj[k] for k in keys}}

    # surrogate: 10^6 samples F = I + N(0, 0.1) plus the iterates of the same
    # problem solved on an 8^3 x 6 mesh, then Adam on the device
    t0 = time.time()
    ctx.mlp_init(1, EMOD, POISSON, 0)
    ctx.mlp_add_samples_random(1000000, 0.1, 2)
    _nh_solve(ctx, N, _nh_problem(N, 8), 0, collect=True)
    nsamp = ctx.mlp_num_samples()
    loss = ctx.mlp_train(20000, 8192, 1e-2, 0, history=True)
    train_s = time.time() - t0
    print("# surrogate trained on %d samples in %.2fs (loss %.2e)" % (nsamp, train_s, loss[-50:].mean()),
          file=sys.stderr, flush=True)

    _nh_solve(ctx, N, prob, 1)  # warm-up
    ctx.kernel_timing(True)
    st_m, u_m = _nh_solve(ctx, N, prob, 1)
    k_mlp = ctx.kernel_time("mlp_tangent")
    ctx.kernel_timing(False)
    print("# newton mlp %.3fs (%d Newton, %d Krylov, %d fallbacks)" % (
        st_m["seconds"], st_m["newton_iters"], st_m["krylov_iters"], st_m["fallbacks"]), file=sys.stderr,
        flush=True)

    # HBM bytes of the fine-level passes each iteration makes (a lower bound: the coarse
    # levels of the hierarchy, 1/14 of the fine matrix at this size, and their Galerkin
    # rebuilds are not counted):
    #   CG iteration: block SpMV over the matrix (block-Jacobi CG: FP64, 8 B per value;
    #     multigrid CG inside Newton: the smoother's FP32 copy, 4 B per value; 4 B column per
    #     3x3 block, p read and A p written) and the two vector kernels (x, r, z, p and A p
    #     read or written, the 3x3 block inverse read by the update): 112 B per dof;
    #   V-cycle (multigrid CG): two passes over the FP32 copy of the fine matrix (residual
    #     after the sweep from zero, post-smoothing sweep; 4 B per value, 4 B column per
    #     block) with their vectors, restriction and prolongation: 48 B per dof;
    #   Newton step: one assembly (values written, connectivity, coordinates and
    #     displacement read, residual written) and, for multigrid, the FP32 copy of the
    #     matrix (8 B read, 4 B written per value).
    n = 3 * nnode
    spmv_bytes = 8 * nnz + 4 * nnz // 9 + 112 * n
    spmv32_bytes = 4 * nnz + 4 * nnz // 9 + 112 * n
    vcycle_bytes = 2 * (4 * nnz + 4 * nnz // 9) + 48 * n
    asm_bytes = 8 * nnz + 4 * 10 * nelem + 24 * nnode + 8 * n + 8 * n

    def frac(st, mg=True):
        b = st["krylov_iters"] * ((spmv32_bytes + vcycle_bytes) if mg else spmv_bytes) + \
            st["newton_iters"] * (asm_bytes + (12 * nnz if mg else 0))
        return b / st["seconds"] / 1e9 / HBM_PEAK_GBS
    nqp = 5 * nelem
    mlp_ms = k_mlp[0] / max(k_mlp[1], 1)
    return {
        "workload": "Neo-Hookean (skr2) Newton, unit cube %d^3x6 = %d P2 tets, %d dofs, z=0 clamped, traction "
                    "(40,0,60) on z=1 in 5 load steps, ||r||_2 < 1e-8, CG preconditioned by a geometric "
                    "multigrid V-cycle (Galerkin P2->P1->coarser P1; coarse levels kept until the CG count grows, finest level in FP32 for smoothing and for the CG products), Eisenstat-Walker forcing; timing includes the multigrid setup" % (M, nelem, n),
        "metric": "Neo-Hookean Newton time-to-1e-8, MLP vs exact tangent",
        "unit": "s", "higher_is_better": False,
        "roofline_note": "fine-level passes only (lower bound): per CG iteration one block SpMV over the "
                         "FP32 copy of the matrix (%d B; block-Jacobi CG: FP64, %d B) + multigrid V-cycle "
                         "with two FP32 matrix passes (%d B), per Newton step one assembly (%d B) + FP32 "
                         "matrix copy; bytes / seconds / 8 TB/s"
                         % (spmv32_bytes, spmv_bytes, vcycle_bytes, asm_bytes),
        "exact": dict({k: st_e[k] for k in keys}, roofline_frac=frac(st_e)),
        "exact_block_jacobi_cg": dict({k: st_bj[k] for k in keys}, roofline_frac=frac(st_bj, mg=False),
                                      solution_max_rel_diff=float(np.max(np.abs(u_bj - u_e)) / np.max(np.abs(u_e)))),
        "mlp": dict({k: st_m[k] for k in keys}, roofline_frac=frac(st_m)),
        "mlp_over_exact_time": st_m["seconds"] / st_e["seconds"],
        "solution_max_rel_diff": float(np.max(np.abs(u_m - u_e)) / np.max(np.abs(u_e))),
        "train": {"seconds": train_s, "samples": nsamp, "adam_steps": 20000, "batch": 8192,
                  "final_loss": float(loss[-50:].mean())},
        "elas_kernels_avg_ms": {"qp": k_qp[0] / max(k_qp[1], 1), "rows": k_rows[0] / max(k_rows[1], 1)},
        "mlp_tangent": {"avg_ms": mlp_ms, "qp_per_launch": nqp,
                        "tflops": 12032.0 * nqp / (mlp_ms * 1e-3) / 1e12 if mlp_ms > 0 else None,
                        "peak_bf16_tflops": 2500.0},
    }


if __name__ == "__main__":
    main()
