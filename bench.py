"""Benchmark of the Taylor-Green RK3 step on one MI355X.

Workload: Taylor-Green vortex on a 512^3 grid, nu = 1/1600, dt = 5e-4, FP32
storage and arithmetic with FP64 reductions (the chosen FP32 form); the same
steps are also timed in FP64 (the plain FP64 path) and the dissipation of the
two runs is compared after the timed steps.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512]

Prints one JSON line. `value` is Mpts/s = n^3 / (seconds per RK3 step) / 1e6.
`roofline` describes the dominant pass kernel: algorithmic HBM bytes per launch
divided by its mean duration inside real steps (HIP events between direct
launches on the solver's stream; the isolated relaunch figure is reported
beside it). `cpu_baseline` times the CPU oracle (a C++ restatement of the
reference's CPU path) on one full RK3 step of the same 512^3 FP32 workload on
the host cores, and on the reference's 64^3 FP64 configuration on 1 thread and
on all cores.
"""
from __future__ import annotations

import argparse
import json

import numpy as np
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PUBLISHED_MPTS_512_FP32 = 512 ** 3 / 0.5694 / 1e6  # 569.4 ms per RK3 step
PASS_NAMES = ("inverse_z", "inverse_y", "x_products", "forward_y", "forward_z_update")
KERNEL_NAMES = ("k_pinv_z", "k_plines_inv", "k_px_products_rd", "k_plines_fwd", "k_pfwd_z_rk_inv")
KERNEL_NAMES_TRANSPOSED = ("k_pinv_z", "k_plines_t_inv", "k_px_products_rd_xt", "k_plines_t_fwd", "k_pfwd_z_rk_inv")
KERNEL_NAMES_BLOCKED = ("k_pinv_z", "k_plines_b_inv", "k_px_products_rd", "k_plines_b_fwd", "k_pfwd_z_rk_inv")
LAYOUT_NAMES = ("line", "transposed", "ky-blocked")
ENSEMBLE_SIZES = [32] * 16 + [48] * 16 + [64] * 16 + [96] * 16


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def step_bytes(n: int, s: int) -> float:
    """Algorithmic HBM bytes of one RK3 step in the five-pass model, 3 x 87 C:
    per stage inverse z 9C, inverse y 12C, x pass 18C, forward y 24C, forward
    z + update 24C, C = n^2 (n/2+1) complex values of s-byte parts. The
    solver's stages run four passes (the inverse z is fused into the previous
    stage's forward-z pass, whose state read-back is served from L2): 84C
    per stage, implemented_step_bytes()."""
    C = n * n * (n // 2 + 1) * 2 * s
    return 261.0 * C


def implemented_step_bytes(n: int, s: int) -> float:
    """3 x (12C + 18C + 24C) + (27C + 30C + 27C): the forward-z pass of
    stage 0 reads no g and that of stage 2 stores none."""
    C = n * n * (n // 2 + 1) * 2 * s
    return 246.0 * C


def time_steps(sph, n, precision, steps, warmup, dt):
    """Host clock around `steps` graph-launched steps after `warmup` ones,
    with a device synchronisation on both sides; also the median of the
    timed steps' GPU times (HIP events around each step's graph replay)."""
    s = sph.TaylorGreenSolver(n, 1600.0, precision)
    for _ in range(warmup):
        s.step_rk3(dt)
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step_rk3(dt)
    s.synchronize()
    sec = (time.perf_counter() - t0) / steps
    ev = s.step_times(steps)
    s.event_median_ms = float(np.median(ev)) if len(ev) else None
    return s, sec


def ensemble_leg(sph, steps: int, warmup: int):
    """64 independent solves (16 each of 32^3, 48^3, 64^3, 96^3), FP32,
    nu = 1e-2, seeded random-phase solenoidal initial fields (seed 1000 + j,
    E0 = 0.5), dt fixed per member from CFL 0.5 at t = 0."""
    import math
    sizes = ENSEMBLE_SIZES
    ens = sph.Ensemble(sizes, 1e-2, [1000 + j for j in range(len(sizes))], precision="fp32")
    dt = [0.5 * (2 * math.pi / n) / ens.max_velocity(j) for j, n in enumerate(sizes)]
    for _ in range(warmup):
        ens.step_rk3(dt)
    ens.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ens.step_rk3(dt)
    ens.synchronize()
    sec = (time.perf_counter() - t0) / steps
    pts = sum(n ** 3 for n in sizes)
    alg = sum(step_bytes(n, 4) for n in sizes)
    # the 32^3 and 48^3 groups' working sets (18C each: 40 and 133 MB) fit in
    # the 256 MB last-level cache; the 64^3 and 96^3 groups' do not
    alg_hbm = sum(step_bytes(n, 4) for n in sizes if n >= 64)
    mode = ens.launch_mode()
    passes = ens.time_passes(3)
    ens.close()
    return {"workload": "64 solves: 16 each of 32^3, 48^3, 64^3, 96^3, FP32, nu=1e-2, CFL 0.5 dt, "
                        f"{steps} RK3 steps, one hipGraph per step",
            "launch_structure": ("per_size_chains", "one_launch_per_pass", "one_launch_per_stage")[mode[0]],
            "timed_step_ms": {"per_size_chains": round(mode[1][0], 4), "one_launch_per_pass": round(mode[1][1], 4),
                              "one_launch_per_stage": round(mode[1][2], 4)},
            "mpts_per_s": round(pts / sec / 1e6, 3), "ms_per_step": round(sec * 1e3, 4),
            "points": pts, "algorithmic_bytes_per_step": alg,
            "roofline_frac": round(alg / sec / 1e9 / HBM_PEAK_GBS, 4),
            "roofline_frac_hbm_groups": round(alg_hbm / sec / 1e9 / HBM_PEAK_GBS, 4),
            "roofline_note": "roofline_frac counts every group's algorithmic bytes, including the 32^3 and 48^3 groups "
                             "whose working sets stay in the 256 MB last-level cache; roofline_frac_hbm_groups counts "
                             "the 64^3 and 96^3 groups only, over the whole step time",
            "passes_ms": {k: round(v[0], 4) for k, v in passes.items()}}


def traffic_from_profiles(kernel: str, n: int):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), if recorded."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{n}_fp32.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"hbm_bytes_per_launch\s*=\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def fft_vs_rocfft(sph, n: int, reps: int = 10):
    """One 3D real-to-complex forward and complex-to-real backward transform
    of an n^3 FP32 field on device pointers: the library's FFT handle
    (sph_fft_forward_device / backward_device, three pass kernels each)
    against hipFFT/rocFFT plans of the same transforms (R2C, C2R), both on one
    stream created here, timed with HIP events on that stream over `reps`
    launches after two warm-ups. rocFFT is only timed here; nothing in the
    library calls it."""
    import ctypes
    import numpy as np
    hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    fftlib = ctypes.CDLL("libhipfft.so.0")
    vp = ctypes.c_void_p
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    hip.hipFree.argtypes = [vp]
    hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
    hip.hipEventRecord.argtypes = [vp, vp]
    hip.hipEventSynchronize.argtypes = [vp]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
    hip.hipEventDestroy.argtypes = [vp]
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
    hip.hipStreamDestroy.argtypes = [vp]
    hip.hipStreamSynchronize.argtypes = [vp]
    fftlib.hipfftSetStream.argtypes = [vp, vp]
    fftlib.hipfftPlan3d.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    fftlib.hipfftExecR2C.argtypes = [vp, vp, vp]
    fftlib.hipfftExecC2R.argtypes = [vp, vp, vp]
    fftlib.hipfftDestroy.argtypes = [vp]
    HIPFFT_R2C, HIPFFT_C2R = 0x2A, 0x2C
    rbytes = n ** 3 * 4
    cbytes = n * n * (n // 2 + 1) * 8
    bufs = [vp(), vp(), vp()]
    for b, sz in zip(bufs, (rbytes, cbytes, rbytes)):
        if hip.hipMalloc(ctypes.byref(b), sz) != 0:
            raise RuntimeError("hipMalloc failed")
    ev = [vp(), vp()]
    plans = [vp(), vp()]
    st = vp()
    try:
        if hip.hipStreamCreate(ctypes.byref(st)) != 0:
            raise RuntimeError("hipStreamCreate failed")
        for e in ev:
            hip.hipEventCreate(ctypes.byref(e))
        u = np.random.default_rng(1).standard_normal(n ** 3).astype(np.float32)
        hip.hipMemcpy(bufs[0], u.ctypes.data, rbytes, 1)

        def timed(fn):
            for _ in range(2):
                fn()
            hip.hipEventRecord(ev[0], st)
            for _ in range(reps):
                fn()
            hip.hipEventRecord(ev[1], st)
            hip.hipEventSynchronize(ev[1])
            ms = ctypes.c_float()
            hip.hipEventElapsedTime(ctypes.byref(ms), ev[0], ev[1])
            return ms.value / reps

        fft = sph.HipFft3D((n, n, n), (2 * 3.141592653589793,) * 3, precision="fp32")
        ours_f = timed(lambda: fft.forward_device(bufs[0].value, bufs[1].value, st.value))
        ours_b = timed(lambda: fft.backward_device(bufs[1].value, bufs[2].value, st.value))
        hip.hipStreamSynchronize(st)
        fft.close()
        for q, kind in ((0, HIPFFT_R2C), (1, HIPFFT_C2R)):
            if fftlib.hipfftPlan3d(ctypes.byref(plans[q]), n, n, n, kind) != 0 or \
                    fftlib.hipfftSetStream(plans[q], st) != 0:
                raise RuntimeError("hipFFT plan set-up failed")
        roc_f = timed(lambda: fftlib.hipfftExecR2C(plans[0], bufs[0], bufs[1]))
        roc_b = timed(lambda: fftlib.hipfftExecC2R(plans[1], bufs[1], bufs[2]))
    finally:
        if st.value:
            hip.hipStreamSynchronize(st)
        for p in plans:
            if p.value:
                fftlib.hipfftDestroy(p)
        for e in ev:
            if e.value:
                hip.hipEventDestroy(e)
        for b in bufs:
            hip.hipFree(b)
        if st.value:
            hip.hipStreamDestroy(st)
    gbs = lambda ms: round((rbytes + cbytes) / (ms * 1e-3) / 1e9, 1)
    return {"workload": f"one 3D R2C / C2R transform of a {n}^3 FP32 field, device pointers", "reps": reps,
            "ours_forward_ms": round(ours_f, 4), "rocfft_forward_ms": round(roc_f, 4),
            "ours_backward_ms": round(ours_b, 4), "rocfft_backward_ms": round(roc_b, 4),
            "min_traffic_gbs": {"ours_forward": gbs(ours_f), "rocfft_forward": gbs(roc_f),
                                "ours_backward": gbs(ours_b), "rocfft_backward": gbs(roc_b)},
            "note": "min_traffic_gbs = (one real + one half-spectrum field) / time; the backward includes "
                    "the 1/N^3 scaling in ours, not in rocFFT's"}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n: int, dt: float):
    """The CPU oracle (C++ restatement of the reference's AVA(CPU)+FFTW path,
    C2C storage, its own Stockham FFT in place of FFTW) timed on this host:
    steady_clock around step_rk3 only, as the reference's step timer
    (src/ava_taylor_green_impl.hpp:980,1043), after one untimed step.

    value: one full RK3 step of the benchmark workload (n^3 FP32) on all
    cores. Beside it, the reference path's own configuration: 64^3 FP64 on 1
    thread (the reference's serial build) and on all cores, next to the
    published 64^3 FP64 CPU steps, and one 256^3 FP64 step on all cores."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def per_step(size, single, nthreads, steps):
        oracle.set_threads(nthreads)
        o = oracle.TaylorGreenOracle(size, 1600.0, single=single)
        o.step_rk3(dt)
        t0 = o.step_seconds()
        for _ in range(steps):
            o.step_rk3(dt)
        sec = (o.step_seconds() - t0) / steps
        o.close()
        return sec

    log(f"cpu baseline: oracle on {threads} threads ({cpu_model()}, nproc {os.cpu_count()})")
    sec64_1 = per_step(64, False, 1, 3)
    sec64_all = per_step(64, False, threads, 5)
    log(f"cpu baseline: 64^3 FP64 {sec64_1 * 1e3:.1f} ms/step (1 thread), {sec64_all * 1e3:.1f} ms/step "
        f"({threads} threads)")
    sec256 = per_step(256, False, threads, 1)
    log(f"cpu baseline: 256^3 FP64 {sec256:.2f} s/step ({threads} threads)")
    sec = per_step(n, True, threads, 1)
    log(f"cpu baseline: {n}^3 FP32 {sec:.2f} s/step ({threads} threads)")
    oracle.set_threads(threads)
    mp = lambda size, t: round(size ** 3 / t / 1e6, 4)
    return {"value": mp(n, sec), "unit": "Mpts/s", "cores": threads, "kind": "port",
            "sample": f"one full RK3 step of the {n}^3 FP32 Taylor-Green workload after one untimed step, "
                      f"{sec:.2f} s (C2C oracle: the reference's algorithm with its own FFT in place of FFTW)",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "fp64_64cube": {"ms_per_step_1_thread": round(sec64_1 * 1e3, 2),
                            f"ms_per_step_{threads}_threads": round(sec64_all * 1e3, 2),
                            "mpts_per_s_1_thread": mp(64, sec64_1),
                            "published_reference_ms_per_step": {
                                "i9-12900H, serial build, FFTW": 188.1, "i5-13500, serial build, FFTW": 106.3},
                            "note": "the published figures are the reference's AVA(CPU)+FFTW path (Benchmark.md:33,"
                                    "42,172,280); the oracle's FFT is a plain mixed-radix Stockham, not FFTW"},
            "fp64_256cube": {f"s_per_step_{threads}_threads": round(sec256, 3), "steps_timed": 1}}


def accuracy_run(sph, n: int, dt: float, t_end: float, every: float):
    """FP32 form against plain FP64 on the same GPU over a whole run: both
    solvers advance in lock step from t = 0 to t_end with fixed dt, and E, the
    dissipation and the enstrophy of both are recorded every `every`."""
    s32 = sph.TaylorGreenSolver(n, 1600.0, "fp32")
    s64 = sph.TaylorGreenSolver(n, 1600.0, "fp64")
    steps = int(round(t_end / dt))
    per = max(1, int(round(every / dt)))
    rows = []
    t0 = time.perf_counter()
    for k in range(steps + 1):
        if k % per == 0 or k == steps:
            d32, d64 = s32.diagnostics(k * dt), s64.diagnostics(k * dt)
            rows.append([k * dt, d64.kinetic_energy, d64.dissipation, d64.enstrophy,
                         d32.kinetic_energy, d32.dissipation, d32.enstrophy])
            log(f"t={k * dt:.4f} E64={d64.kinetic_energy:.12f} eps64={d64.dissipation:.10e} "
                f"eps32 rel err={abs(d32.dissipation - d64.dissipation) / d64.dissipation:.3e} "
                f"{time.perf_counter() - t0:.0f} s")
        if k < steps:
            s32.step_rk3(dt)
            s64.step_rk3(dt)
    r = [list(map(float, x)) for x in rows]
    rel = lambda a, b: abs(a - b) / abs(b)
    out = {"workload": f"Taylor-Green {n}^3, nu=1/1600, dt={dt}, t=0..{t_end}, FP32 form vs plain FP64",
           "steps": steps, "seconds": round(time.perf_counter() - t0, 1),
           "max_rel_err": {"kinetic_energy": max(rel(x[4], x[1]) for x in r),
                           "dissipation": max(rel(x[5], x[2]) for x in r),
                           "enstrophy": max(rel(x[6], x[3]) for x in r)},
           "columns": ["t", "E64", "eps64", "Z64", "E32", "eps32", "Z32"], "rows": r}
    s32.close()
    s64.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--reps", type=int, default=5, help="launches per pass for the isolated kernel timing")
    ap.add_argument("--instep-steps", type=int, default=3,
                    help="real steps whose pass kernels are timed with events for the roofline")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-isolated", action="store_true",
                    help="skip the isolated relaunch timing of the passes (profiled runs: kernel averages then "
                         "cover in-step launches only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-fft-compare", action="store_true",
                    help="skip the timed comparison of one 3D FFT with hipFFT/rocFFT")
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--accuracy-to", type=float, default=None, metavar="T",
                    help="instead of timing: FP32 vs FP64 run to t=T, diagnostics every --accuracy-every")
    ap.add_argument("--accuracy-every", type=float, default=0.1)
    ap.add_argument("--check-n", type=int, default=None,
                    help="after the timed FP32 steps, compare E / dissipation / enstrophy with an FP64 solve "
                         "of this size advanced by the same steps")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the workload is one {args.n}^3 grid per GPU and this "
              f"build has no multi-GPU path (--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    import spectral_solver_amd as sph
    n, K, W = args.n, args.steps, args.warmup
    if args.accuracy_to is not None:
        print(json.dumps(accuracy_run(sph, n, args.dt, args.accuracy_to, args.accuracy_every)), flush=True)
        return

    log(f"fp32 {n}^3: warmup {W}, timing {K} steps")
    s32, sec32 = time_steps(sph, n, "fp32", K, W, args.dt)
    s32_event_median = s32.event_median_ms
    # diagnostics after the timed steps (the in-step pass timing below
    # advances the solver further)
    d32_timed = s32.diagnostics((W + K) * args.dt)
    check = None
    if args.check_n:
        # the Taylor-Green field stays spectrally resolved on a coarser grid
        # for short times: compare with an FP64 solve of n = check_n
        t = (W + K) * args.dt
        d32 = s32.diagnostics(t)
        ref = sph.TaylorGreenSolver(args.check_n, 1600.0, "fp64")
        for _ in range(W + K):
            ref.step_rk3(args.dt)
        d64 = ref.diagnostics(t)
        ref.close()
        check = {"t": t, "against": f"{args.check_n}^3 FP64, same dt and steps",
                 "kinetic_energy": [d32.kinetic_energy, d64.kinetic_energy],
                 "dissipation": [d32.dissipation, d64.dissipation], "enstrophy": [d32.enstrophy, d64.enstrophy],
                 "rel_err": {k: abs(getattr(d32, k) - getattr(d64, k)) / abs(getattr(d64, k))
                             for k in ("kinetic_energy", "dissipation", "enstrophy")}}
        log("check:", check)
    # pass kernels inside real steps (direct launches, HIP events between
    # them on the solver's stream), then the same passes relaunched in
    # isolation (sph_solver_time_passes) for comparison
    instep = s32.time_step_passes(args.dt, args.instep_steps)
    log("in-step pass ms:", {k: round(v[0], 4) if isinstance(v, tuple) else round(v, 4) for k, v in instep.items()})
    passes = None
    if not args.no_isolated:
        passes = s32.time_passes(args.reps)
        log("isolated pass ms:", {k: round(v[0], 4) for k, v in passes.items()})
    tuning = s32.tuning()
    log("autotuner (variant, lines per tile, ms per variant):", tuning)
    layout = s32.layout()
    log("work layout (kept, ms of passes 1-3 line / transposed, 1-4 ky-blocked / best other):", layout)

    fp64 = None
    eps_err = None
    if not args.no_fp64:
        log(f"fp64 {n}^3: warmup {W}, timing {K} steps")
        s64, sec64 = time_steps(sph, n, "fp64", K, W, args.dt)
        t = (W + K) * args.dt
        d32, d64 = d32_timed, s64.diagnostics(t)
        eps_err = abs(d32.dissipation - d64.dissipation) / d64.dissipation
        e_err = abs(d32.kinetic_energy - d64.kinetic_energy) / d64.kinetic_energy
        z_err = abs(d32.enstrophy - d64.enstrophy) / d64.enstrophy
        fp64 = {"ms_per_step": round(sec64 * 1e3, 4), "mpts_per_s": round(n ** 3 / sec64 / 1e6, 3),
                "ms_per_step_gpu_median": round(s64.event_median_ms, 4) if s64.event_median_ms else None,
                "fp32_speedup": round(sec64 / sec32, 3),
                "roofline_frac": round(step_bytes(n, 8) / sec64 / 1e9 / HBM_PEAK_GBS, 4)}
        s64.close()

    in_step = [k for k in range(1, len(PASS_NAMES))]  # the standalone inverse z runs only after set-up
    dom = max(in_step, key=lambda k: instep[PASS_NAMES[k]][0])
    kernel = (KERNEL_NAMES, KERNEL_NAMES_TRANSPOSED, KERNEL_NAMES_BLOCKED)[layout[0]][dom]
    if PASS_NAMES[dom] == "x_products" and tuning["x_products"][0] == 2 and layout[0] != 1:
        kernel = "k_px_products"  # the LDS-staged form
    ms, by = instep[PASS_NAMES[dom]]
    iso_ms = passes[PASS_NAMES[dom]][0] if passes else None
    achieved = by / (ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(kernel, n),
                "kernel": f"{PASS_NAMES[dom]} ({kernel})", "launch_ms": round(ms, 4),
                "launch_ms_source": f"mean over the {3 * args.instep_steps} stages of {args.instep_steps} real steps, "
                                    "HIP events between direct launches on the solver's stream",
                "isolated_launch_ms": round(iso_ms, 4) if iso_ms else None,
                "isolated_frac": round(by / (iso_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if iso_ms else None,
                "algorithmic_bytes": by,
                "algorithmic_bytes_note": ("bytes per launch of this pass, each array read or written once; the "
                                           "forward z + update pass moves 27C, 30C and 27C in the three stages "
                                           "(stage 0 reads no g, stage 2 stores none): the mean, 28C"
                                           if PASS_NAMES[dom] == "forward_z_update" else
                                           "bytes per launch of this pass, each array read or written once"),
                "step_frac": round(step_bytes(n, 4) / sec32 / 1e9 / HBM_PEAK_GBS, 4),
                "step_frac_implemented_bytes": round(implemented_step_bytes(n, 4) / sec32 / 1e9 / HBM_PEAK_GBS, 4),
                "passes_in_step_ms": {k: round(instep[k][0], 4) for k in PASS_NAMES[1:]},
                "passes_in_step_gbs": {k: round(instep[k][1] / (instep[k][0] * 1e-3) / 1e9, 1) for k in PASS_NAMES[1:]},
                "passes_isolated_gbs": ({k: round(v[1] / (v[0] * 1e-3) / 1e9, 1) for k, v in passes.items()}
                                        if passes else None),
                "autotuned": {k: {"variant": v[0], "lines": v[1]} for k, v in tuning.items()},
                "work_layout": {"layout": LAYOUT_NAMES[layout[0]], "passes_1_3_ms_line": round(layout[1][0], 4),
                                "passes_1_3_ms_transposed": round(layout[1][1], 4),
                                "passes_1_4_ms_ky_blocked": round(layout[1][2], 4),
                                "passes_1_4_ms_best_other": round(layout[1][3], 4)}}
    s32.close()

    copy_gbs = sph.copy_bandwidth(2 << 30, 10)
    log(f"copy kernel: {copy_gbs:.1f} GB/s")
    roofline["copy_ceiling_gbs"] = round(copy_gbs, 1)
    roofline["frac_of_copy"] = round(achieved / copy_gbs, 4)

    ens = None
    if not args.no_ensemble:
        log(f"ensemble: 64 members, {args.ensemble_steps} steps")
        ens = ensemble_leg(sph, args.ensemble_steps, W)
        log("ensemble:", ens)

    fftc = None
    if not args.no_fft_compare:
        try:
            fftc = fft_vs_rocfft(sph, n)
            log("fft vs rocfft:", fftc)
        except Exception as e:  # a comparison only, never fatal for the GPU number
            log("fft comparison failed:", e)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(n, args.dt)
        except Exception as e:  # reported, never fatal for the GPU number
            log("cpu baseline failed:", e)

    out = {
        "metric": "ms/RK3 step & Mpts/s at 512^3, FP32 vs FP64, % HBM roofline, dissip. err",
        "value": round(n ** 3 / sec32 / 1e6, 3),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(sec32 * 1e3, 4),
        # median GPU time of the timed steps (HIP events around each graph
        # replay); ms_per_step is the host clock over all of them
        "ms_per_step_gpu_median": round(s32_event_median, 4) if s32_event_median else None,
        "higher_is_better": True,
        "scaling": "weak",
        # the reference's published 512^3 FP32 step (cuFFT on one A100-40GB,
        # 569.4 ms/step = 235.7 Mpts/s, Benchmark.md:62,520); other sizes
        # have no published figure of this metric
        "vs_baseline": round(n ** 3 / sec32 / 1e6 / PUBLISHED_MPTS_512_FP32, 2) if n == 512 else None,
        "dtype": "f32",
        "data": "synthetic (analytic Taylor-Green initial condition)",
        "config": {"workload": f"Taylor-Green vortex {n}^3, nu=1/1600, dt={args.dt}, FP32 fields, FP64 reductions",
                   "grid": n, "dt": args.dt, "parallelism": "single GPU"},
        "fp64": fp64,
        "check": check,
        "dissipation_rel_err_fp32_vs_fp64": eps_err,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "ensemble": ens,
        "fft_vs_rocfft": fftc,
    }
    if fp64:
        out["energy_rel_err_fp32_vs_fp64"] = e_err
        out["enstrophy_rel_err_fp32_vs_fp64"] = z_err
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
