"""Benchmark of the RSVD-Delta-t hot path on one MI355X.

Workload (2D flow operator): 3 coupled real fields on a 512 x 256 periodic
grid, N = 393,216, 11 nonzeros per row, complex FP64 state, k = 16 test
vectors, Nw = 128 frequencies (Nw_eff = 64 for the real operator), w = 2 pi/20,
dt = 0.5 x the RK4 stability limit, transient removal on, q = 1.

A "step" is one RK4 time step of the whole N x k block of test vectors: four
fused SpMM stage launches plus its share of the forcing-chunk GEMM.  `value`
is the step's bytes in the reference formulation divided by the step time:
    bytes/step = 4 S_A + 20 V + (2 Nw_eff/64 + 2) V
(4 SpMM with the stage AXPYs; forcing chunks of 64 half-step samples),
S_A = nnz (8 + 4) + 4 (N + 1) (real FP64 values, int32 columns), V = N k 16.
The default stepping form (Horner stages with the forcing folded into one
term per step) needs 4 S_A + 12 V + (Nw_eff/S + 1) V; that figure and its
rate are in `config`.  The sketch is generated on the device and every input
is resident in HBM before the timed region; the folded forcing terms (once
per action) are computed before it.  The forcing-chunk GEMM runs on the stage
kernels' stream (engine option overlap_forcing = 0; --overlap runs it
concurrently on a second stream, as the engine does by default), so the
stage kernel's event-timed launch duration is its own.

--full additionally times the whole main.c sequence (time to modes) in FP64
and in the FP32-state form and reports the sigma_1 error of the latter.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rsvd-delta-t_amd"))


def load_pkg():
    import importlib.util
    spec = importlib.util.spec_from_file_location("rsvdt_amd", os.path.join(ROOT, "rsvd-delta-t_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rsvdt_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_baseline(A, k, Nw, w, TL, dt, bytes_per_step, budget_s=12.0):
    """The CPU oracle (reference algorithm restated in C, OpenMP over rows):
    a bounded sample of forced RK4 vector-steps of the same operator, scaled to
    the GB/s-equivalent of the GPU metric (bytes of one block step / time of k
    vector-steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = min(os.cpu_count() or 1, 16)
    o = oracle.Oracle(A, k=k, q=0, Nw=Nw, w=w, TL=TL, dt=dt, transient_removal=False, nthreads=threads)
    rng = np.random.default_rng(3)
    Fk = rng.standard_normal((A.shape[0], o.Nw_eff)) + 1j * rng.standard_normal((A.shape[0], o.Nw_eff))
    t1 = o.time_steps(Fk, 1, 2)
    nsteps = max(2, int(budget_s / max(t1 / 2, 1e-6)))
    nsteps = min(nsteps, 400)
    t = o.time_steps(Fk, 1, nsteps)
    per_vector_step = t / nsteps
    value = bytes_per_step / (k * per_vector_step) / 1e9
    return {"value": round(value, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"{nsteps} forced RK4 steps of 1 test vector (4 SpMV + 2 forcing GEMV + 7 AXPY each, "
                      f"as SourceCode/TSRK4.c), {per_vector_step * 1e3:.1f} ms per vector-step; "
                      f"scaled to k={k} vectors per block step",
            "vector_step_ms": round(per_vector_step * 1e3, 4)}


def c0_full_run(pkg):
    """The whole main.c sequence on the CPU-runnable Ginzburg-Landau case
    (N=500, U=2+0.2i, k=10, Nw=64, w=0.05, dt=T/4096, TL=10, transient
    removal, q=0) with the same seeded F_hat: the CPU oracle on one core
    against the engine's default form on the GPU (second run, after the
    autotuner), and the largest relative difference of their gains."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import workloads
    A = workloads.gl(500)
    k, Nw, w, TL = 10, 64, 0.05, 10.0
    dt = 2 * np.pi / w / 4096
    rng = np.random.default_rng(2024)
    o = oracle.Oracle(A, k=k, q=0, Nw=Nw, w=w, TL=TL, dt=dt, transient_removal=True, nthreads=1)
    F = rng.standard_normal((500, k * o.Nw_eff)) + 1j * rng.standard_normal((500, k * o.Nw_eff))
    t0 = time.time()
    _, _, S_cpu = o.run(F)
    t_cpu = time.time() - t0
    s = pkg.Session(A, k=k, q=0, Nw=Nw, w=w, TL=TL, dt=dt, transient_removal=True)
    s.run(F)
    pkg.lib().rsvdt_hip_sync()
    t0 = time.time()
    _, _, S_gpu = s.run(F)
    pkg.lib().rsvdt_hip_sync()
    t_gpu = time.time() - t0
    s.close()
    return {"workload": "Ginzburg-Landau N=500, k=10, Nw=64, TL=10, transient removal, q=0 (main.c sequence)",
            "cpu_oracle_s": round(t_cpu, 3), "cpu_cores": 1, "gpu_s": round(t_gpu, 4),
            "speedup": round(t_cpu / t_gpu, 1),
            "sigma_max_rel_diff": float(np.max(np.abs(S_gpu - S_cpu) / np.abs(S_cpu)))}


def calib_copy(pkg, gib=1, reps=5):
    """A plain device copy of known size (gib GiB read + gib GiB written per
    launch, reps launches of rsvdt_hip_zcopy3): the reference reading for
    scaling the profiler's FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes
    L = pkg.lib()
    if L.rsvdt_hip_init(0) != 0:
        raise RuntimeError("no HIP device")
    n = (gib << 30) // 16
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if L.rsvdt_hip_malloc(ctypes.byref(p), ctypes.c_size_t(n * 16)) != 0:
            raise RuntimeError("device allocation failed")
    L.rsvdt_hip_zero(a, ctypes.c_size_t(n * 16), None)
    for _ in range(reps):
        if L.rsvdt_hip_zcopy3(ctypes.c_int64(1), ctypes.c_int64(n), ctypes.c_int64(1), a, ctypes.c_int64(0),
                              ctypes.c_int64(1), ctypes.c_int64(0), b, ctypes.c_int64(0), ctypes.c_int64(1),
                              ctypes.c_int64(0), None) != 0:
            raise RuntimeError("copy failed")
    L.rsvdt_hip_sync()
    L.rsvdt_hip_free(a); L.rsvdt_hip_free(b)
    print(json.dumps({"calib_copy": {"bytes_read_per_launch": n * 16, "bytes_written_per_launch": n * 16,
                                     "launches": reps, "kernel": "k_zcopy3"}}))


def stream_ceiling(pkg, gib=1, reps=10):
    """The HBM rate a plain streaming kernel with the stage kernel's mix
    reaches on this GPU: y += a x over two gib-GiB complex arrays
    (rsvdt_hip_zaxpy: two reads and one write per element), timed with HIP
    events; GB/s of bytes moved."""
    import ctypes
    L = pkg.lib()
    V = ctypes.c_void_p
    L.rsvdt_hip_event_create.argtypes = [ctypes.POINTER(V)]
    L.rsvdt_hip_event_record.argtypes = [V, V]
    L.rsvdt_hip_event_elapsed_ms.argtypes = [V, V, ctypes.POINTER(ctypes.c_float)]
    L.rsvdt_hip_zaxpy.argtypes = [ctypes.c_int64, ctypes.c_double, V, V, V]
    n = (gib << 30) // 16
    x, y, e0, e1 = V(), V(), V(), V()
    for p in (x, y):
        if L.rsvdt_hip_malloc(ctypes.byref(p), ctypes.c_size_t(n * 16)) != 0:
            raise RuntimeError("device allocation failed")
        L.rsvdt_hip_zero(p, ctypes.c_size_t(n * 16), None)
    L.rsvdt_hip_event_create(ctypes.byref(e0)); L.rsvdt_hip_event_create(ctypes.byref(e1))
    L.rsvdt_hip_zaxpy(n, 0.5, x, y, None)
    L.rsvdt_hip_event_record(e0, None)
    for _ in range(reps):
        L.rsvdt_hip_zaxpy(n, 0.5, x, y, None)
    L.rsvdt_hip_event_record(e1, None)
    L.rsvdt_hip_sync()
    ms = ctypes.c_float()
    L.rsvdt_hip_event_elapsed_ms(e0, e1, ctypes.byref(ms))
    L.rsvdt_hip_event_destroy(e0); L.rsvdt_hip_event_destroy(e1)
    L.rsvdt_hip_free(x); L.rsvdt_hip_free(y)
    return round(3 * 16 * n * reps / (ms.value * 1e-3) / 1e9, 1)


def time_to_modes_gl(pkg, nmembers):
    """Time to sigma_1 relative error <= 1e-6 on held-out members of the
    Ginzburg-Landau family (N=500, k=16, Nw=64, transient removal): the plain
    FP64 GPU form (reference stage order, one launch per stage, random
    sketch) against the optimised form (Horner stages, LDS-resident chunks,
    learned sketch columns from the trained network), each with the fewest
    power iterations that reach the tolerance."""
    import learned
    if not os.path.exists(learned.WEIGHTS):
        return {"skipped": f"no trained weights at {os.path.relpath(learned.WEIGHTS, ROOT)}"}
    net = learned.SketchNet(seed=0).load(learned.WEIGHTS)
    members = learned.family_sample(np.random.default_rng(777), nmembers)
    rows, summ = learned.time_to_modes(pkg, net, members)
    summ = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in summ.items()}
    summ["members"] = [{"mu0": round(r["mu0"], 4), "ReU": round(r["Ur"], 4),
                        "plain": {"q": r["plain"]["q"], "s": round(r["plain"]["wall_s"], 4),
                                  "err": r["plain"]["sigma1_err"]},
                        "fast": {"q": r["fast"]["q"], "s": round(r["fast"]["wall_s"], 4),
                                 "err": r["fast"]["sigma1_err"]}} for r in rows]
    summ["workload"] = ("Ginzburg-Landau family (N=500, k=16 = 8 learned + 8 random columns, Nw=64, w=0.05, "
                        "dt=T/4096, TL=10, transient removal), held-out members; tolerance 1e-6 on sigma_1 "
                        "against q=4 of the optimised form")
    return summ


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--precision", type=int, default=64, choices=(64, 32))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--full", action="store_true")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--overlap", action="store_true",
                    help="forcing GEMM on a second stream, concurrent with the stage kernels (about 1 %% faster "
                         "per step; the concurrent GEMM's traffic then stretches the stage kernels' event-timed "
                         "durations, so the default keeps it on the stage kernels' stream)")
    ap.add_argument("--workload", default="flow2d", choices=("flow2d", "flow3d"),
                    help="flow2d: the metric's workload; flow3d: the 3D operator (k=32, Nw=64)")
    ap.add_argument("--rk4-form", type=int, default=0, choices=(0, 1),
                    help="0: Horner stages with the folded forcing (default); 1: the reference's stage order")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a 1 GiB device copy 5 times (profiler counter calibration)")
    ap.add_argument("--no-ttm", action="store_true",
                    help="skip the time-to-modes comparison on held-out Ginzburg-Landau family members")
    ap.add_argument("--ttm-members", type=int, default=4)
    ap.add_argument("--graph", action="store_true",
                    help="replay each forcing chunk's stage launches from a captured HIP graph")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU design shards test vectors "
              "and frequencies and is not built", file=sys.stderr)
        sys.exit(2)

    # the library's progress messages go to stderr: stdout carries only the JSON line
    sys.stdout.flush()
    json_fd = os.dup(1)
    os.dup2(2, 1)
    pkg = load_pkg()
    if args.calib_copy:
        calib_copy(pkg)
        return
    import workloads
    if args.workload == "flow2d":
        A = workloads.flow2d()
        k, Nw, w = 16, 128, 2 * np.pi / 20
        dt = workloads.flow2d_dt()
        wl = "flow2d N=393216 (3 fields, 512x256 periodic, 11 nnz/row)"
    else:
        A = workloads.flow3d()
        k, Nw, w = 32, 64, 2 * np.pi / 20
        dt = workloads.flow3d_dt()
        wl = "flow3d N=1966080 (5 fields, 128x64x48, periodic x/z, Dirichlet y, 11 nnz/row)"
    N = A.shape[0]
    T = 2 * np.pi / w
    TL = T / 2
    s = pkg.Session(A, k=k, q=1, Nw=Nw, w=w, TL=TL, dt=dt, transient_removal=True, precision=args.precision,
                    chunk_steps=args.chunk, overlap_forcing=1 if args.overlap else 0, verbose=1,
                    rk4_form=args.rk4_form, use_graph=int(args.graph))
    s.random_sketch(seed=1373, gaussian=True)
    ms, rk_ms, fr_ms = s.bench_steps(args.steps, warmup=args.warmup)
    st = pkg.action_stats()
    S = st["chunk_steps"]
    form = st["rk4_form"]
    esz = 16 if args.precision == 64 else 8
    V = N * k * esz
    # the matrix as the kernels read it (FP64 real values, int32 columns), and
    # as the reference formulation of the same precision would store it
    S_A = A.nnz * (8 + 4) + 4 * (N + 1)
    S_A_ref = A.nnz * ((8 if args.precision == 64 else 4) + 4) + 4 * (N + 1)
    Nwe = s.Nw_eff
    # the metric's byte count: the reference formulation of one step (4 SpMM
    # with the stage AXPYs, 20 block reads/writes, plus the forcing samples of
    # a 64-sample chunk), whatever form the engine runs
    bytes_step_ref = 4 * S_A_ref + 20 * V + (2 * Nwe / 64 + 2) * V
    # what the form that ran needs per step
    if form == 0:
        bytes_step_alg = 4 * S_A + 12 * V + (Nwe / S + 1) * V
        kernel = {6: "k_horner_group", 7: "k_horner_pair", 8: "k_horner_pair"}.get(st["spmm_variant"], "k_horner_stage")
        bytes_launch = S_A + 3 * V
    else:
        bytes_step_alg = 4 * S_A + 20 * V + (2 * Nwe / (2 * S + 1) + 2) * V
        kernel, bytes_launch = "k_rk4_stage", S_A + 5 * V
    # HBM traffic per launch of the stage kernel, from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE runs of this workload and precision
    traffic, traffic_src = None, None
    tj = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tj) and form == 0:
        ent = json.load(open(tj)).get(f"{args.workload}_fp{args.precision}")
        if ent:
            traffic, traffic_src = ent["traffic_bytes"], "profiles/" + ent["profile"]
    ms_step = ms / args.steps
    value = bytes_step_ref / (ms_step * 1e-3) / 1e9
    # dominant kernel: the fused SpMM stage kernel, 4 launches per step
    launch_ms = rk_ms / (4 * args.steps)
    achieved = bytes_launch / (launch_ms * 1e-3) / 1e9
    out = {
        "metric": "time-to-modes at sigma1 rel.err<=1e-6 vs plain FP64 GPU; SpMM/RK4 HBM GB/s",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if args.precision == 64 else "c64",
        "data": "synthetic",
        "config": {"workload": f"{wl} RK4 time stepping, k={k}, Nw={Nw} (Nw_eff={Nwe}), transient removal, "
                               + ("FP64" if args.precision == 64 else "FP32 state"),
                   "N": N, "k": k, "Nw": Nw, "dt": dt, "chunk_steps": S, "precision": args.precision,
                   "hip_graph": bool(args.graph), "forcing_overlapped": bool(args.overlap),
                   "rk4_form": "horner+folded forcing" if form == 0 else "reference stage order",
                   "bytes_per_step_reference_form": int(bytes_step_ref),
                   "bytes_per_step_algorithmic": int(bytes_step_alg),
                   "alg_GBps": round(bytes_step_alg / (ms_step * 1e-3) / 1e9, 2),
                   "spmm_shape": {"lanes_per_row": st["spmm_lanes"], "variant": st["spmm_variant"],
                                  "xcd_remap": st["spmm_xcd_remap"]}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": kernel, "bytes_per_launch": int(bytes_launch),
                     "avg_launch_ms": round(launch_ms, 5)},
        "breakdown_ms_per_step": {"rk4": round(rk_ms / args.steps, 5), "forcing": round(fr_ms / args.steps, 5)},
    }
    # one action's time stepping: Nt + Ns + RR steps with the transient removal
    rend = s.Nt + s.Ns + s.RR
    out["action_stepping_s"] = round(ms_step * rend / 1e3, 3)
    if not args.no_cpu:
        cb = cpu_baseline(A, k, Nw, w, TL, dt, bytes_step_ref)
        # the same action's stepping on the host cores, extrapolated from the
        # sample (k test vectors, one after the other as TSActionRK4.c:89 does)
        cb["action_stepping_s_extrapolated"] = round(cb["vector_step_ms"] * 1e-3 * k * rend, 1)
        if args.workload == "flow2d":
            cb["ginzburg_landau_full_run"] = c0_full_run(pkg)
        out["cpu_baseline"] = cb
    if args.full:
        rng = None
        res = {}
        for prec in (64, 32):
            s2 = pkg.Session(A, k=k, q=1, Nw=Nw, w=w, TL=TL, dt=dt, transient_removal=True, precision=prec)
            s2.random_sketch(seed=1373, gaussian=True)
            t0 = time.time()
            for name, fn in (("direct", s2.direct), ("power iteration", s2.power_iteration),
                             ("svd", s2.svd_before_adjoint), ("store U", s2.store_u), ("adjoint", s2.adjoint),
                             ("svd all", s2.svd_allfreqs)):
                fn()
                pkg.lib().rsvdt_hip_sync()
                print(f"[bench] FP{prec} {name} done at {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
            res[prec] = (time.time() - t0, pkg.mat_to_dense(pkg._vp(s2.Res.S_hat)).real)
            s2.close()
        err = float(np.max(np.abs(res[32][1][0] - res[64][1][0]) / res[64][1][0]))
        out["time_to_modes_s"] = {"fp64": round(res[64][0], 3), "fp32_state": round(res[32][0], 3),
                                  "speedup": round(res[64][0] / res[32][0], 3), "sigma1_max_rel_err_fp32": err}
    s.close()
    sc = stream_ceiling(pkg)
    out["roofline"]["stream_2r1w_GBps"] = sc
    out["roofline"]["frac_of_stream"] = round(achieved / sc, 4)
    if not args.no_ttm and args.workload == "flow2d":
        out["time_to_modes"] = time_to_modes_gl(pkg, args.ttm_members)
    sys.stdout.flush()
    os.write(json_fd, (json.dumps(out) + "\n").encode())


if __name__ == "__main__":
    main()
