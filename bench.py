"""Benchmark of the RSVD-Delta-t hot path on one MI355X.

Headline workload (3D flow operator): 5 coupled real fields on a 128 x 64 x 48
grid (N = 1,966,080, 11 nonzeros per row), complex FP64 state, k = 32 test
vectors, Nw = 64 frequencies (Nw_eff = 32 for the real operator),
w = 2 pi/20, dt = 0.5 x the RK4 stability limit, transient removal on, q = 1.

A "step" is one forcing chunk of the resolvent action's time stepping
(SourceCode/TSActionRK4.c:89-122 -> TSRK4.c:18-47 for all k test vectors at
once): S = 64 RK4 steps of the N x k block -- four fused SpMM stage launches
each -- plus the coefficient GEMM that forms the chunk's S forcing samples
(CreateForcingOnFly.c:17-19).  Whole chunks are timed, so the time per step
does not depend on --steps.  `value` is the algorithmic bytes of the form that
runs divided by the measured time:
    bytes/chunk = S (4 S_A + 12 V) + (Nw_eff + S) V
(Horner stages: stage 1 reads y and writes v1, stages 2-3 gather v and read
y, stage 4 gathers v3 and reads y and the forcing term; the GEMM reads the
Nw_eff folded forcing blocks once per chunk and writes S samples),
S_A = nnz (8 + 4) + 4 (N + 1) (real FP64 values, int32 columns), V = N k 16.
The sketch is generated on the device and every input is resident in HBM
before the timed region; the folded forcing terms (once per action) are
computed before it.  The forcing GEMM runs on the stage kernels' stream, so
the stage kernel's event-timed launch duration is its own.

Extra keys: the same measurement on the 2D flow operator (N = 393,216,
k = 16, Nw = 128) and in the FP32-state form, the FP64-MFMA rates of the
forcing GEMM, the DFT GEMM and the CholeskyQR Gram/apply kernels at the 3D
shape, a tuned streaming probe, the CPU baseline, and the time to modes on the
Ginzburg-Landau family (learned sketch against two baselines).
"""
import argparse
import ctypes
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rsvd-delta-t_amd"))

PEAK_HBM = 8000.0          # GB/s, MI355X
PEAK_FP64_MFMA = 78.6      # TFLOP/s, MI355X dense FP64 matrix


def load_pkg():
    import importlib.util
    spec = importlib.util.spec_from_file_location("rsvdt_amd", os.path.join(ROOT, "rsvd-delta-t_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rsvdt_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(name):
    """(A, k, Nw, w, dt, description) of a named benchmark workload."""
    import workloads
    if name == "flow3d":
        return (workloads.flow3d(), 32, 64, 2 * np.pi / 20, workloads.flow3d_dt(),
                "flow3d N=1966080 (5 fields, 128x64x48, periodic x/z, Dirichlet y, 11 nnz/row)")
    return (workloads.flow2d(), 16, 128, 2 * np.pi / 20, workloads.flow2d_dt(),
            "flow2d N=393216 (3 fields, 512x256 periodic, 11 nnz/row)")


STAGE_KERNELS = {6: "k_horner_group", 7: "k_horner_pair", 8: "k_horner_pair", 9: "k_horner_tile", 10: "k_horner_tile"}


def stepping(pkg, name, steps, warmup, precision=64, chunk=64, rk4_form=0, graph=False):
    """Time `steps` whole forcing chunks (after `warmup` chunks) of the
    action's time stepping on a named workload.  Returns the figures."""
    A, k, Nw, w, dt, desc = workload(name)
    N = A.shape[0]
    T = 2 * np.pi / w
    s = pkg.Session(A, k=k, q=1, Nw=Nw, w=w, TL=T / 2, dt=dt, transient_removal=True, precision=precision,
                    chunk_steps=chunk, overlap_forcing=0, verbose=1, rk4_form=rk4_form, use_graph=int(graph))
    s.random_sketch(seed=1373, gaussian=True)
    ms, rk_ms, fr_ms = s.bench_steps(steps * chunk, warmup=warmup * chunk)
    st = pkg.action_stats()
    S = st["chunk_steps"]
    assert S == chunk, (S, chunk)
    Nwe = s.Nw_eff
    rend = s.Nt + s.Ns + s.RR
    s.close()
    pkg.lib().RSVDT_Finalize()      # return the chunk buffers (2 x 64 blocks) to the device
    esz = 16 if precision == 64 else 8
    V = N * k * esz
    S_A = A.nnz * (8 + 4) + 4 * (N + 1)
    nsteps = steps * S
    if st["rk4_form"] == 0:
        bytes_chunk = S * (4 * S_A + 12 * V) + (Nwe + S) * V
        bytes_launch = S_A + 3 * V
        kernel = STAGE_KERNELS.get(st["spmm_variant"], "k_horner_stage")
    else:
        # the reference's stage order: 20 block reads/writes per step, 2S+1 forcing samples per chunk
        bytes_chunk = S * (4 * S_A + 20 * V) + (Nwe + 2 * S + 1) * V
        bytes_launch = S_A + 5 * V
        kernel = "k_rk4_stage"
    launch_ms = rk_ms / (4 * nsteps)
    ms_chunk = ms / steps
    fr_chunk = fr_ms / steps
    return {
        "workload": desc, "N": N, "k": k, "Nw": Nw, "Nw_eff": Nwe, "dt": dt, "chunk_steps": S,
        "precision": precision, "rk4_form": "horner+folded forcing" if st["rk4_form"] == 0 else "reference stage order",
        "nnz": int(A.nnz), "S_A": int(S_A), "V": int(V),
        "ms_per_chunk": ms_chunk, "ms_per_rk4_step": ms_chunk / S,
        "bytes_per_chunk": int(bytes_chunk), "GBps": bytes_chunk / (ms_chunk * 1e-3) / 1e9,
        "stage_kernel": kernel, "stage_bytes_per_launch": int(bytes_launch), "stage_avg_launch_ms": launch_ms,
        "stage_GBps": bytes_launch / (launch_ms * 1e-3) / 1e9,
        "forcing_ms_per_chunk": fr_chunk,
        "forcing_TFLOPs": 8.0 * N * k * Nwe * S / (fr_chunk * 1e-3) / 1e12,
        "spmm_shape": {"lanes": st["spmm_lanes"], "variant": st["spmm_variant"], "xcd_remap": st["spmm_xcd_remap"]},
        "action_stepping_s": ms_chunk / S * rend / 1e3, "rend": rend,
    }


def host_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except Exception:
        nproc = None
    return {"cpu_model": model, "nproc": nproc, "os_cpu_count": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)), "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def cpu_threads():
    """Threads for the CPU baseline: the cores this process may run on,
    capped by OMP_NUM_THREADS when the environment sets it (the host's share
    for this GPU)."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        n = min(n, int(omp))
    return max(1, n)


CPU_LABEL = "reference-algorithm CPU restatement (PETSc unavailable offline)"


def cpu_baseline(name, bytes_per_chunk, chunk, budget_s=15.0):
    """The CPU oracle (the reference algorithm restated in C: per test vector,
    per step 4 CSR SpMV + 2 forcing GEMV + 7 AXPY as SourceCode/TSRK4.c,
    OpenMP over rows) on a bounded sample of forced RK4 vector-steps of the
    same operator; the GB/s-equivalent of the GPU metric (bytes of one chunk
    of the block / time of k x S vector-steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    A, k, Nw, w, dt, _ = workload(name)
    threads = cpu_threads()
    T = 2 * np.pi / w
    o = oracle.Oracle(A, k=k, q=0, Nw=Nw, w=w, TL=T / 2, dt=dt, transient_removal=False, nthreads=threads)
    rng = np.random.default_rng(3)
    Fk = rng.standard_normal((A.shape[0], o.Nw_eff)) + 1j * rng.standard_normal((A.shape[0], o.Nw_eff))
    t1 = o.time_steps(Fk, 1, 2)
    nsteps = int(min(400, max(2, budget_s / max(t1 / 2, 1e-6))))
    t = o.time_steps(Fk, 1, nsteps)
    per = t / nsteps
    value = bytes_per_chunk / (k * chunk * per) / 1e9
    out = {"value": round(value, 3), "unit": "GB/s", "cores": threads, "kind": "port", "label": CPU_LABEL,
           "sample": f"{nsteps} forced RK4 steps of 1 test vector of {name} (4 SpMV + 2 forcing GEMV + 7 AXPY "
                     f"each, as SourceCode/TSRK4.c), {per * 1e3:.2f} ms per vector-step; scaled to k={k} vectors x "
                     f"{chunk} steps per chunk",
           "vector_step_ms": round(per * 1e3, 4)}
    out.update(host_info())
    return out


def c0_full_run(pkg):
    """The whole main.c sequence on the CPU-runnable Ginzburg-Landau case
    (N=500, U=2+0.2i, k=10, Nw=64, w=0.05, dt=T/4096, TL=10, transient
    removal, q=0) with the same seeded F_hat: the CPU oracle given all host
    cores (its OpenMP row loops engage above 8192 rows, so at N=500 it runs
    serially, as a row decomposition of 500 rows would gain nothing) against
    the engine's default form on the GPU (second run, after the autotuner),
    and the largest relative difference of their gains."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import workloads
    A = workloads.gl(500)
    k, Nw, w, TL = 10, 64, 0.05, 10.0
    dt = 2 * np.pi / w / 4096
    rng = np.random.default_rng(2024)
    threads = cpu_threads()
    o = oracle.Oracle(A, k=k, q=0, Nw=Nw, w=w, TL=TL, dt=dt, transient_removal=True, nthreads=threads)
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
            "label": CPU_LABEL, "cpu_oracle_s": round(t_cpu, 3), "cpu_threads": threads,
            "cpu_parallel": "serial at N=500 (OpenMP row loops engage above 8192 rows)", "gpu_s": round(t_gpu, 4),
            "speedup": round(t_cpu / t_gpu, 1),
            "sigma_max_rel_diff": float(np.max(np.abs(S_gpu - S_cpu) / np.abs(S_cpu)))}


class Ev:
    """HIP event pair on the null stream (where the kernels below run)."""

    def __init__(self, L):
        self.L = L
        V = ctypes.c_void_p
        L.rsvdt_hip_event_create.argtypes = [ctypes.POINTER(V)]
        L.rsvdt_hip_event_record.argtypes = [V, V]
        L.rsvdt_hip_event_elapsed_ms.argtypes = [V, V, ctypes.POINTER(ctypes.c_float)]
        self.a, self.b = V(), V()
        L.rsvdt_hip_event_create(ctypes.byref(self.a))
        L.rsvdt_hip_event_create(ctypes.byref(self.b))

    def time(self, fn, reps):
        fn()
        self.L.rsvdt_hip_sync()
        self.L.rsvdt_hip_event_record(self.a, None)
        for _ in range(reps):
            if fn():
                raise RuntimeError(self.L.rsvdt_hip_last_error())
        self.L.rsvdt_hip_event_record(self.b, None)
        self.L.rsvdt_hip_sync()
        ms = ctypes.c_float()
        self.L.rsvdt_hip_event_elapsed_ms(self.a, self.b, ctypes.byref(ms))
        return ms.value / reps

    def close(self):
        self.L.rsvdt_hip_event_destroy(self.a)
        self.L.rsvdt_hip_event_destroy(self.b)


def dev_alloc(L, nbytes):
    p = ctypes.c_void_p()
    if L.rsvdt_hip_malloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
        raise RuntimeError(f"device allocation of {nbytes} bytes failed")
    L.rsvdt_hip_zero(p, ctypes.c_size_t(nbytes), None)
    return p


def mfma_kernels(pkg, N, k, Nwe, Nw, reps=5):
    """FP64-MFMA rates of the DFT GEMM (DFT.c:26-38: Y_hat_m = RR sum_c
    Y_all_c e^{-2 pi i m c/Nw}), on one eighth of the 3D workload's N*k
    columns (the kernel streams the columns; the shape of the coefficient
    side is the workload's), and of the CholeskyQR Gram and apply passes
    (QRAllFreqs.c:22-32) on all Nw_eff frequency blocks of N x k."""
    L = pkg.lib()
    V, I64 = ctypes.c_void_p, ctypes.c_int64
    L.rsvdt_hip_coef_gemm.argtypes = [I64, I64, I64, V, I64, I64, I64, ctypes.c_int, I64, ctypes.c_double, V, I64,
                                      V, I64, ctypes.c_int, ctypes.c_double, V]
    L.rsvdt_hip_bgram_work_bytes.restype = ctypes.c_size_t
    L.rsvdt_hip_bgram_work_bytes.argtypes = [I64] * 4
    L.rsvdt_hip_bgram.argtypes = [I64] * 4 + [V, I64, I64, I64, V, I64, I64, I64, V, V, V]
    L.rsvdt_hip_bapply.argtypes = [I64] * 4 + [V, I64, I64, I64, V, ctypes.c_double, V, I64, I64, I64, V, I64, I64,
                                               I64, V]
    ev = Ev(L)
    out = {}
    # DFT slice: P = Nw_eff outputs, K = Nw snapshots, Ls columns
    Ls = N * k // 8
    rng = np.random.default_rng(5)
    coef = rng.standard_normal((Nw, Nwe)) + 1j * rng.standard_normal((Nw, Nwe))
    dcoef = dev_alloc(L, coef.nbytes)
    L.rsvdt_hip_memcpy_h2d.argtypes = [V, V, ctypes.c_size_t]
    L.rsvdt_hip_memcpy_h2d(dcoef, coef.ctypes.data_as(V), coef.nbytes)
    din = dev_alloc(L, Nw * Ls * 16)
    dout = dev_alloc(L, Nwe * Ls * 16)
    ms = ev.time(lambda: L.rsvdt_hip_coef_gemm(Nwe, Nw, Ls, dcoef, Nwe, 1, 0, 1, Nwe, 1.0, din, Ls, dout, Ls, 64, 0.0,
                                               None), reps)
    fl = 8.0 * Nwe * Nw * Ls
    out["dft_gemm"] = {"kernel": "k_coef_gemm", "shape": f"P={Nwe} K={Nw} L={Ls} (N k / 8)", "avg_ms": round(ms, 4),
                       "TFLOPs": round(fl / (ms * 1e-3) / 1e12, 2),
                       "frac_fp64_mfma": round(fl / (ms * 1e-3) / 1e12 / PEAK_FP64_MFMA, 4),
                       "GBps": round((Nw + Nwe) * Ls * 16 / (ms * 1e-3) / 1e9, 1)}
    for p in (din, dout, dcoef):
        L.rsvdt_hip_free(p)
    # CholeskyQR passes on Nw_eff blocks of N x k (row-major blocks)
    dX = dev_alloc(L, Nwe * N * k * 16)
    L.rsvdt_hip_fill_random.argtypes = [I64, I64, I64, ctypes.c_uint64, ctypes.c_int, V, V]
    L.rsvdt_hip_fill_random(N, k, Nwe, 11, 1, dX, None)
    dG = dev_alloc(L, Nwe * k * k * 16)
    dW = dev_alloc(L, L.rsvdt_hip_bgram_work_bytes(Nwe, N, k, k))
    dO = dev_alloc(L, Nwe * N * k * 16)
    ms = ev.time(lambda: L.rsvdt_hip_bgram(Nwe, N, k, k, dX, N * k, k, 1, dX, N * k, k, 1, dG, dW, None), reps)
    fl = 8.0 * Nwe * N * k * k
    out["gram"] = {"kernel": "k_bgram_mfma + k_bgram_reduce", "shape": f"{Nwe} blocks of {N} x {k}",
                   "avg_ms": round(ms, 4), "TFLOPs": round(fl / (ms * 1e-3) / 1e12, 2),
                   "frac_fp64_mfma": round(fl / (ms * 1e-3) / 1e12 / PEAK_FP64_MFMA, 4),
                   "GBps": round(Nwe * N * k * 16 / (ms * 1e-3) / 1e9, 1)}
    ms = ev.time(lambda: L.rsvdt_hip_bapply(Nwe, N, k, k, dX, N * k, k, 1, dG, 0.0, None, 0, 0, 0, dO, N * k, k, 1,
                                            None), reps)
    out["apply"] = {"kernel": "k_bapply_mfma", "shape": f"{Nwe} blocks of {N} x {k} times {k} x {k}",
                    "avg_ms": round(ms, 4), "TFLOPs": round(fl / (ms * 1e-3) / 1e12, 2),
                    "frac_fp64_mfma": round(fl / (ms * 1e-3) / 1e12 / PEAK_FP64_MFMA, 4),
                    "GBps": round(2 * Nwe * N * k * 16 / (ms * 1e-3) / 1e9, 1),
                    "frac_hbm": round(2 * Nwe * N * k * 16 / (ms * 1e-3) / 1e9 / PEAK_HBM, 4)}
    for p in (dX, dG, dW, dO):
        L.rsvdt_hip_free(p)
    ev.close()
    return out


def stream_ceiling(pkg, gib=1, reps=10):
    """The HBM rate of a plain streaming kernel with the stage kernel's mix:
    y += a x over two gib-GiB complex arrays (rsvdt_hip_zaxpy: two reads and
    one write per element, nontemporal), HIP events, GB/s moved."""
    L = pkg.lib()
    V = ctypes.c_void_p
    L.rsvdt_hip_zaxpy.argtypes = [ctypes.c_int64, ctypes.c_double, V, V, V]
    n = (gib << 30) // 16
    x, y = dev_alloc(L, n * 16), dev_alloc(L, n * 16)
    ev = Ev(L)
    ms = ev.time(lambda: L.rsvdt_hip_zaxpy(n, 0.5, x, y, None), reps)
    ev.close()
    L.rsvdt_hip_free(x)
    L.rsvdt_hip_free(y)
    return round(3 * 16 * n / (ms * 1e-3) / 1e9, 1)


def calib_copy(pkg, gib=1, reps=5):
    """A plain device copy of known size (gib GiB read + gib GiB written per
    launch, reps launches of rsvdt_hip_zcopy3): the reference reading for
    scaling the profiler's FETCH_SIZE / WRITE_SIZE counters to bytes."""
    L = pkg.lib()
    if L.rsvdt_hip_init(0) != 0:
        raise RuntimeError("no HIP device")
    n = (gib << 30) // 16
    a, b = dev_alloc(L, n * 16), dev_alloc(L, n * 16)
    I = ctypes.c_int64
    for _ in range(reps):
        if L.rsvdt_hip_zcopy3(I(1), I(n), I(1), a, I(0), I(1), I(0), b, I(0), I(1), I(0), None) != 0:
            raise RuntimeError("copy failed")
    L.rsvdt_hip_sync()
    L.rsvdt_hip_free(a)
    L.rsvdt_hip_free(b)
    return {"calib_copy": {"bytes_read_per_launch": n * 16, "bytes_written_per_launch": n * 16, "launches": reps,
                           "kernel": "k_zcopy3"}}


def full_run(pkg, name="flow3d", precision=64):
    """The whole main.c sequence (SourceCode/main.c:134-146: direct action,
    one power iteration, SVD, StoreU, adjoint action, SVD) on a named workload
    with transient removal at TL = T/2 (the stepping workload's setting), on
    a seeded device sketch: wall time and, per action, the device time of the
    stage kernels, the forcing GEMMs and the transient removal."""
    A, k, Nw, w, dt, desc = workload(name)
    T = 2 * np.pi / w
    s = pkg.Session(A, k=k, q=1, Nw=Nw, w=w, TL=T / 2, dt=dt, transient_removal=True, precision=precision)
    s.random_sketch(seed=1373, gaussian=True)
    pkg.lib().rsvdt_hip_sync()
    t0 = time.time()
    actions = []
    for nm, fn in (("direct", s.direct), ("power iteration", s.power_iteration), ("svd", s.svd_before_adjoint),
                   ("store U", s.store_u), ("adjoint", s.adjoint), ("svd all", s.svd_allfreqs)):
        fn()
        pkg.lib().rsvdt_hip_sync()
        if nm in ("direct", "adjoint"):
            st = pkg.action_stats()
            actions.append({"action": nm, "stage_kernels_s": round(st["rk4_ms"] / 1e3, 3),
                            "forcing_s": round(st["forcing_ms"] / 1e3, 3),
                            "transient_removal_s": round(st["tr_ms"] / 1e3, 3),
                            "action_device_s": round(st["total_ms"] / 1e3, 3)})
        log(f"{name} FP{precision} {nm} done at {time.time() - t0:.1f} s")
    wall = time.time() - t0
    S = pkg.mat_to_dense(pkg._vp(s.Res.S_hat)).real
    s.close()
    pkg.lib().RSVDT_Finalize()
    return {"workload": f"{desc}, k={k}, Nw={Nw}, q=1, transient removal, TL=T/2, "
                        + ("FP64" if precision == 64 else "FP32 state"),
            "time_to_modes_s": round(wall, 2), "actions": actions, "sigma1_max": float(S[0].max())}


def subspace_gap(U1, U2):
    """sin of the largest principal angle between span(U1) and span(U2)
    (orthonormal columns)."""
    return float(np.linalg.norm(U2 - U1 @ (U1.conj().T @ U2), 2))


def fp32_form_c2(pkg, TLT=6.0, nmodes=3):
    """The precision form on the 2D flow operator (N = 393,216, k = 16,
    Nw = 128 -> 64 frequencies, q = 1, transient removal, TL = TLT periods,
    the same for both forms): the whole main.c sequence with the FP32 state
    (the Nt steps before the first snapshot in complex FP32, the snapshot
    period, DFT, CholeskyQR, SVD and the removal in FP64) against the FP64
    form, on the same seeded sketch.  Per frequency: the largest relative
    error of sigma_1..sigma_nmodes and the gap between the leading nmodes
    response (U) and forcing (V) subspaces.  The FP32 rounding accumulated
    over the transient steps is a perturbation of the state when the
    snapshots start, i.e. part of the transient, which the FP64 snapshot
    period carries and the removal takes out; with the whole run in FP32 the
    removal's correction carried the rounding into the gains (1.6e-5 at
    TL = 6 T on this operator).  At TL = 6 T the FP64 form is converged; at
    TL = T/2 the FP64 removal itself is accurate to ~5e-3 on this operator
    and the two forms differ within that (3e-3)."""
    A, k, Nw, w, dt, desc = workload("flow2d")
    T = 2 * np.pi / w
    res = {}
    for prec in (64, 32):
        s = pkg.Session(A, k=k, q=1, Nw=Nw, w=w, TL=TLT * T, dt=dt, transient_removal=True, precision=prec)
        s.random_sketch(seed=1373, gaussian=True)
        pkg.lib().rsvdt_hip_sync()
        t0 = time.time()
        for fn in (s.direct, s.power_iteration, s.svd_before_adjoint, s.store_u, s.adjoint, s.svd_allfreqs):
            fn()
        pkg.lib().rsvdt_hip_sync()
        wall = time.time() - t0
        U = pkg.mat_to_dense(pkg._vp(s.Res.U_hat))
        Vh = pkg.mat_to_dense(pkg._vp(s.Res.V_hat))
        S = pkg.mat_to_dense(pkg._vp(s.Res.S_hat)).real
        nwe = s.Nw_eff
        s.close()
        pkg.lib().RSVDT_Finalize()
        res[prec] = (U, Vh, S, wall)
        log(f"C2 FP{prec}: {wall:.1f} s")
    U64, V64, S64, t64 = res[64]
    U32, V32, S32, t32 = res[32]
    es = np.abs(S32[:nmodes] - S64[:nmodes]) / S64[:nmodes]
    gu = [subspace_gap(U64[:, m * k:m * k + nmodes], U32[:, m * k:m * k + nmodes]) for m in range(nwe)]
    gv = [subspace_gap(V64[:, m * k:m * k + nmodes], V32[:, m * k:m * k + nmodes]) for m in range(nwe)]
    return {"workload": f"{desc}, k={k}, Nw={Nw} (Nw_eff={nwe}), q=1, transient removal, TL={TLT:g}T (both forms)",
            "fp32_steps_fraction": round(TLT / (TLT + 1.0 + 1.0 / Nw), 3),
            "fp64_s": round(t64, 2), "fp32_s": round(t32, 2), "time_ratio_fp64_over_fp32": round(t64 / t32, 3),
            "sigma_rel_err_max": float(es.max()), "sigma1_rel_err_max": float(es[0].max()),
            "sigma1_rel_err_median": float(np.median(es[0])), "within_1e-6": bool(es[0].max() <= 1e-6),
            f"sigma1_{nmodes}_rel_err_per_freq": [float(x) for x in es.max(axis=0)],
            f"U_top{nmodes}_subspace_gap_per_freq": [float(x) for x in gu],
            f"V_top{nmodes}_subspace_gap_per_freq": [float(x) for x in gv],
            "U_gap_max": float(max(gu)), "V_gap_max": float(max(gv))}


def save_results_c3(pkg, folder, N=1966080, k=32, nwe=32, opt=2):
    """SaveResults (SourceCode/SaveResults.c:32-92) of results of the 3D
    workload's size -- U_hat and V_hat of N x k*Nw_eff complex FP64 (32 GB
    each) resident on the device -- into `folder`, timed end to end (file
    bytes / wall time), the files removed afterwards."""
    import shutil
    L = pkg.lib()
    if L.RSVDT_Initialize(None, None) != 0:
        raise RuntimeError("RSVDT_Initialize failed")
    res = pkg.Resolvent_matrices()
    handles = []
    for name in ("U_hat", "V_hat"):
        m = ctypes.c_void_p()
        pkg.check(L.RSVDT_MatCreateDenseBlocked(N, k * nwe, k, None, ctypes.byref(m)), "RSVDT_MatCreateDenseBlocked")
        setattr(res, name, m.value)
        handles.append(m.value)
    res.S_hat = pkg.mat_from_dense(np.ones((k, nwe), np.complex128)).value
    handles.append(res.S_hat)
    V = pkg.RSVDt_vars()
    V.RSVD.k, V.RSVD.Nw_eff, V.SaveResultsOpt = k, nwe, opt
    d = pkg.Directories()
    os.makedirs(folder, exist_ok=True)
    d.FolderDir = (os.path.abspath(folder) + "/").encode()
    pkg.lib().rsvdt_hip_sync()
    t0 = time.time()
    rc = L.SaveResults(ctypes.byref(res), ctypes.byref(V), ctypes.byref(d))
    wall = time.time() - t0
    nbytes = sum(os.path.getsize(os.path.join(folder, f)) for f in os.listdir(folder))
    nfiles = len(os.listdir(folder))
    shutil.rmtree(folder, ignore_errors=True)
    for h in handles:
        pkg.destroy(h)
    pkg.check(rc, "SaveResults")
    return {"opt": opt, "files": nfiles, "file_bytes": nbytes, "seconds": round(wall, 2),
            "GBps": round(nbytes / wall / 1e9, 2), "folder": folder,
            "data": f"U_hat, V_hat: {N} x {k * nwe} complex FP64 each ({N * k * nwe * 16 / 1e9:.1f} GB), "
                    "on the device"}


def time_to_modes_gl(pkg, nmembers):
    """Time to sigma_1 relative error <= 1e-6 on held-out members of the
    Ginzburg-Landau family (N=500, k=16, Nw=64, transient removal): the
    optimised form with the learned sketch against (a) the plain FP64 GPU
    form (reference stage order, one launch per stage, random sketch) and
    (b) the engine's best FP64 form with a random sketch (Horner stages,
    LDS-resident chunks), each with the fewest power iterations that reach
    the tolerance."""
    import learned
    if not os.path.exists(learned.WEIGHTS):
        return {"skipped": f"no trained weights at {os.path.relpath(learned.WEIGHTS, ROOT)}"}
    net = learned.SketchNet(seed=0).load(learned.WEIGHTS)
    members = learned.family_sample(np.random.default_rng(777), nmembers)
    rows, summ = learned.time_to_modes(pkg, net, members,
                                       log=lambda s: print(f"[ttm] {s}", file=sys.stderr, flush=True))
    summ = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in summ.items()}
    summ["speedup_vs_best_random_fp64"] = round(summ["fast_random"]["s"] / summ["fast_s"], 3)
    summ["members_evaluated"] = len(rows)
    summ["max_sigma1_err"] = {n: float(max(r[n]["sigma1_err"] for r in rows))
                              for n in ("plain", "fast", "fast_fp32", "fast_random")}
    summ["workload"] = ("Ginzburg-Landau family (N=500, k=16 = 8 learned + 8 random columns, Nw=64, w=0.05, "
                        "dt=T/4096, TL=10, transient removal), held-out members; tolerance 1e-6 on sigma_1 "
                        "against q=4 of the optimised form")
    return summ


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed forcing chunks (64 RK4 steps each)")
    ap.add_argument("--warmup", type=int, default=3, help="untimed forcing chunks before the timed ones")
    ap.add_argument("--precision", type=int, default=64, choices=(64, 32))
    ap.add_argument("--workload", default="flow3d", choices=("flow3d", "flow2d"))
    ap.add_argument("--rk4-form", type=int, default=0, choices=(0, 1),
                    help="0: Horner stages with the folded forcing (default); 1: the reference's stage order")
    ap.add_argument("--graph", action="store_true",
                    help="replay each forcing chunk's stage launches from a captured HIP graph")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline workload (no flow2d, FP32, MFMA, streaming, time-to-modes keys)")
    ap.add_argument("--no-ttm", action="store_true")
    ap.add_argument("--no-c2", action="store_true", help="skip the FP32-form measurement on the 2D operator")
    ap.add_argument("--ttm-members", type=int, default=50)
    ap.add_argument("--c2-tlt", type=float, default=6.0,
                    help="transient length of the FP32-form comparison on the 2D operator, in periods")
    ap.add_argument("--only", choices=("mfma", "stream", "save", "ttm", "c2", "full3d"), default=None,
                    help="run one extra measurement alone (profiling; save: SaveResults of 3D-size results; "
                         "ttm: time to modes on --ttm-members members; c2: the FP32 form on the 2D operator)")
    ap.add_argument("--save-dir", default="/dev/shm/rsvdt_save_bench",
                    help="folder for --only save (removed afterwards)")
    ap.add_argument("--save-nwe", type=int, default=32, help="frequency blocks of the --only save results")
    ap.add_argument("--save-opt", type=int, default=2, choices=(1, 2))
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a 1 GiB device copy 5 times (profiler counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU design shards test vectors "
              "and frequencies and is not built", file=sys.stderr)
        sys.exit(2)

    # the library's progress messages go to stderr: stdout carries only the JSON line
    sys.stdout.flush()
    json_fd = os.dup(1)
    os.dup2(2, 1)

    def emit(obj):
        sys.stdout.flush()
        os.write(json_fd, (json.dumps(obj) + "\n").encode())

    pkg = load_pkg()
    if args.calib_copy:
        emit(calib_copy(pkg))
        return
    if args.only == "mfma":
        L = pkg.lib()
        if L.rsvdt_hip_init(0) != 0:
            raise RuntimeError("no HIP device")
        emit({"mfma_kernels": mfma_kernels(pkg, 1966080, 32, 32, 64)})
        return
    if args.only == "ttm":
        emit({"time_to_modes": time_to_modes_gl(pkg, args.ttm_members)})
        return
    if args.only == "full3d":
        emit({"full_run_flow3d": full_run(pkg, "flow3d")})
        return
    if args.only == "c2":
        emit({"fp32_form_c2": fp32_form_c2(pkg, TLT=args.c2_tlt)})
        return
    if args.only == "save":
        emit({"save_results": save_results_c3(pkg, args.save_dir, nwe=args.save_nwe, opt=args.save_opt)})
        return
    if args.only == "stream":
        L = pkg.lib()
        if L.rsvdt_hip_init(0) != 0:
            raise RuntimeError("no HIP device")
        emit({"stream_2r1w_GBps": stream_ceiling(pkg)})
        return

    t_start = time.time()
    h = stepping(pkg, args.workload, args.steps, args.warmup, precision=args.precision, rk4_form=args.rk4_form,
                 graph=args.graph)
    log(f"headline {args.workload}: {h['ms_per_chunk']:.3f} ms per chunk, {h['GBps']:.1f} GB/s, stage "
        f"{h['stage_avg_launch_ms'] * 1e3:.1f} us ({time.time() - t_start:.0f} s)")
    # HBM traffic per launch of the stage kernel, from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE runs of this workload, precision and kernel
    traffic, traffic_src = None, None
    tj = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tj):
        ent = json.load(open(tj)).get(f"{args.workload}_fp{args.precision}")
        if ent and ent.get("stage_kernel") == h["stage_kernel"] and h["rk4_form"].startswith("horner"):
            traffic, traffic_src = ent["traffic_bytes"], "profiles/" + ent["profile"]
    achieved = h["stage_GBps"]
    out = {
        "metric": "time-to-modes at sigma1 rel.err<=1e-6 vs plain FP64 GPU; SpMM/RK4 HBM GB/s",
        "value": round(h["GBps"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["ms_per_chunk"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if args.precision == 64 else "c64",
        "data": "synthetic",
        "config": {"workload": f"{h['workload']} resolvent-action time stepping, k={h['k']}, Nw={h['Nw']} "
                               f"(Nw_eff={h['Nw_eff']}), transient removal, "
                               + ("FP64" if args.precision == 64 else "FP32 state"),
                   "step": f"one forcing chunk: {h['chunk_steps']} RK4 steps of the N x k block (4 stage launches "
                           "each) + the chunk's forcing GEMM",
                   "N": h["N"], "k": h["k"], "Nw": h["Nw"], "dt": h["dt"], "chunk_steps": h["chunk_steps"],
                   "precision": args.precision, "hip_graph": bool(args.graph), "rk4_form": h["rk4_form"],
                   "bytes_per_step": h["bytes_per_chunk"],
                   "bytes_per_step_formula": "S (4 S_A + 12 V) + (Nw_eff + S) V" if args.rk4_form == 0
                   else "S (4 S_A + 20 V) + (Nw_eff + 2S + 1) V",
                   "S_A": h["S_A"], "V": h["V"], "ms_per_rk4_step": round(h["ms_per_rk4_step"], 5),
                   "spmm_shape": h["spmm_shape"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": h["stage_kernel"], "bytes_per_launch": h["stage_bytes_per_launch"],
                     "bytes_per_launch_formula": "S_A + 3 V (mean of the 4 Horner stages)",
                     "avg_launch_ms": round(h["stage_avg_launch_ms"], 5)},
        "breakdown_ms_per_step": {"stage_kernels": round(h["stage_avg_launch_ms"] * 4 * h["chunk_steps"], 4),
                                  "forcing_gemm": round(h["forcing_ms_per_chunk"], 4)},
        "forcing_gemm": {"kernel": "k_coef_gemm", "TFLOPs": round(h["forcing_TFLOPs"], 2),
                         "frac_fp64_mfma": round(h["forcing_TFLOPs"] / PEAK_FP64_MFMA, 4),
                         "flops_per_chunk": 8 * h["N"] * h["k"] * h["Nw_eff"] * h["chunk_steps"]},
        "action_stepping_s": round(h["action_stepping_s"], 3),
    }
    if not args.no_extras:
        try:
            out["roofline"]["stream_2r1w_GBps"] = stream_ceiling(pkg)
            out["roofline"]["frac_of_stream"] = round(achieved / out["roofline"]["stream_2r1w_GBps"], 4)
        except Exception as e:   # a failed extra must not lose the headline
            out["roofline"]["stream_error"] = str(e)
        try:
            out["mfma_kernels"] = mfma_kernels(pkg, h["N"], h["k"], h["Nw_eff"], h["Nw"]) \
                if args.workload == "flow3d" else None
            log(f"mfma kernels done ({time.time() - t_start:.0f} s)")
        except Exception as e:
            out["mfma_kernels"] = {"error": str(e)}
        other = "flow2d" if args.workload == "flow3d" else "flow3d"
        extra = {}
        for nm, prec in ((other, 64), (args.workload, 32), (other, 32)):
            try:
                r = stepping(pkg, nm, 8, 2, precision=prec)
                extra[f"{nm}_fp{prec}"] = {
                    "ms_per_chunk": round(r["ms_per_chunk"], 4), "GBps": round(r["GBps"], 1),
                    "stage_kernel": r["stage_kernel"], "stage_avg_launch_ms": round(r["stage_avg_launch_ms"], 5),
                    "stage_GBps": round(r["stage_GBps"], 1), "stage_frac": round(r["stage_GBps"] / PEAK_HBM, 4),
                    "forcing_TFLOPs": round(r["forcing_TFLOPs"], 2), "spmm_shape": r["spmm_shape"]}
                log(f"{nm} FP{prec}: {r['ms_per_chunk']:.3f} ms per chunk ({time.time() - t_start:.0f} s)")
            except Exception as e:
                extra[f"{nm}_fp{prec}"] = {"error": str(e)}
        out["other_workloads"] = extra
    if not args.no_cpu:
        cb = cpu_baseline(args.workload, h["bytes_per_chunk"], h["chunk_steps"])
        # one action's stepping on the host cores, extrapolated from the sample
        # (k test vectors one after the other, as TSActionRK4.c:89 does)
        cb["action_stepping_s_extrapolated"] = round(cb["vector_step_ms"] * 1e-3 * h["k"] * h["rend"], 1)
        cb["ginzburg_landau_full_run"] = c0_full_run(pkg)
        out["cpu_baseline"] = cb
        log(f"cpu baseline done ({time.time() - t_start:.0f} s)")
    if not args.no_ttm and not args.no_extras:
        out["time_to_modes"] = time_to_modes_gl(pkg, args.ttm_members)
        log(f"time to modes done ({time.time() - t_start:.0f} s)")
    if not args.no_c2 and not args.no_extras:
        out["fp32_form_c2"] = fp32_form_c2(pkg, TLT=args.c2_tlt)
        log(f"FP32 form on the 2D operator done ({time.time() - t_start:.0f} s)")
    out["bench_wall_s"] = round(time.time() - t_start, 1)
    emit(out)


if __name__ == "__main__":
    main()
