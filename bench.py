#!/usr/bin/env python3
"""Benchmark of the model-image hot path on one MI355X.

Workload: the 1024^2 single-image configuration -- Sersic (n=4, r_e=60, ell=0.2) +
Exponential (h=120, ell=0.5) + FlatSky, 51^2 Moffat PSF, data pixels 493-532
oversampled 5x with a 255^2 Moffat PSF, Poisson noise (gain 4, seed 2).  One step
is one model + chi^2 evaluation (GetFitStatistic): evaluate the model with
sub-pixel integration, convolve it (and the oversampled region) with the PSF,
form the weighted deviates against the data and reduce chi^2 in FP64.  Data,
weights and PSF spectra are resident in HBM before timing starts.

The headline `value` is in FP64, the reference's arithmetic (parity mode: pixels,
FFT and reduction in FP64); the FP32-pixel mode is timed the same way and reported
beside it.  Prints one JSON line (see the keys below).  The CPU baseline is the FP64
CPU oracle of the same evaluation, timed on this host's cores on a bounded sample.
"""
import argparse
import ctypes
import json
import re
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "model+χ² evals/s and LM fits/s, adjoint+batched vs plain GPU at equal χ²"
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6


def copy_calibration(mb):
    """A device-to-device copy of `mb` MiB (run by the runtime's blit kernel), whose
    FETCH_SIZE / WRITE_SIZE counters calibrate the ones of our kernels."""
    hip = ctypes.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log("copy calibration: %d MiB device-to-device" % mb)


def device_sync():
    hip = ctypes.CDLL("libamdhip64.so")
    if hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def fft_flops(n):
    return 5.0 * n * math.log2(n)


def smooth(n):
    """Smallest n' >= n whose only prime factors are 2, 3 and 5 (the FFT lengths)."""
    while True:
        k = n
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return n
        n += 1


def grid_geometry(W, H, P, split_multiple=1):
    """Convolution plane of one single-image grid: the evaluated extent is the output
    plus the PSF halo (W+P-1 columns), the two halves of the image (split row rounded
    up to a multiple of `split_multiple`) share one complex plane of Nr x Nc."""
    Hh = -(-((H + 1) // 2) // split_multiple) * split_multiple
    Lc, Lr = W + P - 1, Hh + P - 1
    return Lc, Lr, smooth(Lc), smooth(Lr), Hh


def kernel_model(w, prec_bytes):
    """Algorithmic HBM bytes and flops per launch of the merged kernels of one
    evaluation (main grid + every oversampled region grid), FP32 or FP64 pixels.

    k_eval_rowfft writes the Lr evaluated rows of each plane (Nc complex each);
    k_colconv reads them, reads the PSF spectrum (Nr x Nc) and writes the Hh output
    rows; k_inverse_stat reads those rows plus the data and weight images.  The
    sub-pixel tail kernel is compute work (tens of kB of output)."""
    cb = 2 * prec_bytes
    grids = [(w.width, w.height, w.psf.shape[0], 1)]
    for (psf, scale, x1, x2, y1, y2) in w.osamp:
        grids.append(((x2 - x1 + 1) * scale, (y2 - y1 + 1) * scale, psf.shape[0], scale))
    out = {"k_eval_rowfft": {"bytes": 0.0, "flops": 0.0}, "k_colconv": {"bytes": 0.0, "flops": 0.0},
           "k_inverse_stat": {"bytes": 0.0, "flops": 0.0}}
    for (W, H, P, s) in grids:
        Lc, Lr, Nc, Nr, Hh = grid_geometry(W, H, P, s)
        out["k_eval_rowfft"]["bytes"] += Lr * Nc * cb
        out["k_eval_rowfft"]["flops"] += Lr * fft_flops(Nc)
        out["k_colconv"]["bytes"] += (Lr + Nr + Hh) * Nc * cb
        out["k_colconv"]["flops"] += Nc * 2 * fft_flops(Nr) + 6 * Nr * Nc
        out["k_inverse_stat"]["bytes"] += Hh * Nc * cb
        out["k_inverse_stat"]["flops"] += Hh * fft_flops(Nc)
    out["k_inverse_stat"]["bytes"] += w.width * w.height * 2 * prec_bytes   # data + weights
    return out


def profiled_traffic(kernel, prec):
    """HBM bytes per dispatch of `kernel` ("k_colconv", ...) in precision `prec` ("float" or
    "double") from profiles/hbm_traffic_cfg2_1024.md (rocprofv3 FETCH_SIZE / WRITE_SIZE of
    this workload, scaled by a copy kernel), or None.  Kernels with compile-time variants
    are listed as <T,DISK,DERIV>; this workload runs the plain one."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_cfg2_1024.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes" and plain_variant(f[1], kernel, prec):
            return float(f[2])
    return None


def plain_variant(name, kernel, prec, flags=None):
    """Whether the profiled kernel name is `kernel` in precision `prec` with the given
    compile-time flags after the precision (default: all false -- the variant of a
    single image without the 3-D disk, derivatives, point sources or looping grid)."""
    m = re.match(r"^ifg::(\w+)<(\w+)((?:,\w+)*)>$", name)
    if not m or m.group(1) != kernel or m.group(2) != prec:
        return False
    got = [v for v in m.group(3).split(",") if v]
    if flags is None:
        return all(v == "false" for v in got)
    return got[:len(flags)] == list(flags) and all(v == "false" for v in got[len(flags):])


def profiled_valu(kernel, prec, profile="valu_cfg2_1024.md", flags=None):
    """VALU wave instructions per dispatch of `kernel` from profiles/<profile>
    (rocprofv3 SQ_INSTS_VALU of the workload; flags: the compile-time variant after the
    precision, as plain_variant), or None."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "valu_insts" and plain_variant(f[1], kernel, prec, flags):
            return float(f[2])
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, p, gpu_stats, warmup=5, timed=20):
    """The FP64 CPU oracle of the same evaluation (the reference's algorithm restated,
    with a real-input FFT in place of FFTW) on this host: `warmup` untimed calls, then
    the median of `timed` GetFitStatistic calls, with 1 thread and with all threads
    (OMP_NUM_THREADS, else os.cpu_count()); the FFT's share of the time; and, as the
    checker, the GPU statistics' relative error against the oracle's."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    res = {}
    stat_ref = None
    for thr in sorted({1, threads}):
        o = w.setup_fit(oracle.OracleModel(threads=thr))
        for _ in range(warmup):
            stat_ref = o.GetFitStatistic(p)
        oracle.conv_seconds(reset=True)
        ts = []
        for _ in range(timed):
            t0 = time.perf_counter()
            o.GetFitStatistic(p)
            ts.append(time.perf_counter() - t0)
        res[thr] = (float(np.median(ts)), oracle.conv_seconds(reset=True) / sum(ts))
        log("cpu oracle, %d thread(s): median %.4f s per evaluation (%.1f evals/s), FFT share %.2f" %
            (thr, res[thr][0], 1.0 / res[thr][0], res[thr][1]))
    out = {"value": 1.0 / res[threads][0], "unit": "evals/s", "cores": threads, "kind": "port",
           "sample": "%d warm-up + median of %d FP64 GetFitStatistic calls of the same 1024^2 workload, with 1 and %d "
                     "threads; CPU oracle restatement of the reference path (OpenMP schedule(static,10) pixel "
                     "loops, Kahan sums, serial deviates, mp_enorm) with its own vectorised real-input FFT in "
                     "place of FFTW (absent on this host)" % (warmup, timed, threads),
           "value_1_thread": 1.0 / res[1][0], "cpu_model": cpu_model(),
           "fft_share_1_thread": res[1][1], "fft_share_all_threads": res[threads][1]}
    for key, s in gpu_stats.items():
        out["gpu_%s_vs_oracle_stat_rel" % key] = abs(s - stat_ref) / abs(stat_ref)
    return out


def lm_benchmark(w, repeats):
    """LM fits/s on the 1024^2 workload, FP64, every leg with ftol 1e-10 from the same
    start: (1) mpfit (MINPACK LMDIF, the reference's caller) with its forward-difference
    Jacobian -- one GPU ComputeDeviates per free parameter, QR of the m x n Jacobian on
    the host; (2) LM on the normal equations with the same forward-difference Jacobian,
    its nFree + 1 evaluations in one batched GPU launch and J^T J, J^T r reduced on the
    GPU; (3) the same LM with the exact (analytic) Jacobian.  (2) vs (3) isolates the
    exact Jacobian's gain; (1) vs (2) the gain of the GPU normal equations."""
    import imfit_amd
    m = w.fit_model(imfit_amd.PRECISION_FP64)
    x0, lim = w.lm_start(), w.lm_limits()
    legs = {}
    for key, kw, reps in (("fd_mpfit", dict(method="mpfit", jacobian="fd"), 1),
                          ("fd_normal", dict(method="normal", jacobian="fd"), repeats),
                          ("analytic_normal", dict(method="normal", jacobian="analytic"), repeats)):
        if reps > 1:
            m.FitLM(x0, limits=lim, ftol=1e-10, **kw)   # plans and graphs (untimed)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            x, r = m.FitLM(x0, limits=lim, ftol=1e-10, **kw)
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        c = m.GetFitStatistic(x)
        legs[key] = {"s_per_fit": t, "fits_per_s": 1.0 / t, "chi2": c, "status": r.status, "niter": r.niter,
                     "nfev": r.nfev}
        log("LM %-16s %.4f s per fit (median of %d), status %d, niter %d, nfev %d, chi2 %.10f" %
            (key, t, reps, r.status, r.niter, r.nfev, c))
    c0 = legs["fd_mpfit"]["chi2"]
    for v in legs.values():
        v["chi2_rel_diff_vs_fd_mpfit"] = (v["chi2"] - c0) / c0
    return {"legs": legs, "ftol": 1e-10,
            "speedup_analytic_vs_fd_normal": legs["fd_normal"]["s_per_fit"] / legs["analytic_normal"]["s_per_fit"],
            "speedup_analytic_vs_fd_mpfit": legs["fd_mpfit"]["s_per_fit"] / legs["analytic_normal"]["s_per_fit"],
            "start": "truth x U(0.9,1.1), centre +-5 px (seed 1)"}


def gradient_benchmark(models, p, results, steps):
    """Exact gradient of the statistic on the 1024^2 workload, per precision: the
    reverse-mode (adjoint) pass against the tangent method (one derivative plane per
    parameter through the forward path), each as ms per gradient and in units of one
    evaluation (the headline's ms per step), with their largest difference relative to
    the gradient's largest entry."""
    import imfit_amd
    out = {}
    reps = max(5, steps // 10)
    for key, m in models.items():
        res, grads = {}, {}
        for method, name in ((imfit_amd.GRADIENT_ADJOINT, "adjoint"), (imfit_amd.GRADIENT_TANGENT, "tangent")):
            m.SetGradientMethod(method)
            m.GetFitStatisticGradient(p)   # plans and graphs (untimed)
            device_sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                g, _ = m.GetFitStatisticGradient(p)
            device_sync()
            dt = (time.perf_counter() - t0) / reps
            grads[name] = g
            res[name] = {"ms_per_gradient": 1e3 * dt, "evaluations": dt / (1e-3 * results[key]["ms_per_step"])}
        m.SetGradientMethod(imfit_amd.GRADIENT_ADJOINT)
        kt, _ = kernel_times(m, p, 3, 1, call=lambda: m.GetFitStatisticGradient(p))
        res["adjoint"]["kernels_us"] = {k: 1e3 * t / n for k, (t, n) in kt.items()}
        res["adjoint_vs_tangent_max_rel"] = float(np.abs(grads["adjoint"] - grads["tangent"]).max() /
                                                  np.abs(grads["tangent"]).max())
        res["speedup_adjoint_vs_tangent"] = res["tangent"]["ms_per_gradient"] / res["adjoint"]["ms_per_gradient"]
        res["reps"] = reps
        out[key] = res
        log("gradient %s: adjoint %.3f ms (%.2f evaluations), tangent %.3f ms (%.2f evaluations), max rel diff %.2e" %
            (key, res["adjoint"]["ms_per_gradient"], res["adjoint"]["evaluations"], res["tangent"]["ms_per_gradient"],
             res["tangent"]["evaluations"], res["adjoint_vs_tangent_max_rel"]))
    return out


def valu_roofline(kt, prec, profile, flags=None):
    """VALU-bound workloads: the kernel with the most event time, its average duration
    and its VALU issue fraction (4 cycles per wave64 instruction on one of the 1024
    SIMDs at 2.4 GHz, over the duration) from the committed SQ_INSTS_VALU profile
    (flags: the dominant kernel's compile-time variant, e.g. ["true"] for the 3-D disk)."""
    dom = max(kt.items(), key=lambda kv: kv[1][0])[0]
    avg_s = kt[dom][0] / kt[dom][1] * 1e-3
    valu = profiled_valu(dom, prec, profile, flags if dom in ("k_eval_rowfft", "k_tail") else None)
    return {"bound": "valu", "kernel": dom, "avg_us": avg_s * 1e6,
            "valu_issue_frac": 4.0 * valu / 1024 / 2.4e9 / avg_s if valu is not None else None,
            "kernels_us": {k: 1e3 * t / n for k, (t, n) in kt.items()}}


def aux_benchmarks(steps, fits=True):
    """Short measurements of the other single-GPU configurations, reported beside the
    headline: the 3-D disk model (cfg4), a 4096-vector population (cfg3), many
    cutouts (cfg5) and bootstrap resampling on the 256^2 image."""
    import imfit_amd
    from imfit_amd import workloads
    out = {}
    w4 = workloads.cfg4()
    p4 = w4.start_params(seed=1, frac=0.05)
    res4 = {}
    for prec, key in ((imfit_amd.PRECISION_FP32, "fp32"), (imfit_amd.PRECISION_FP64, "fp64")):
        m = w4.fit_model(prec)
        m.Autotune(p4)
        n = max(5, steps // 20)
        _, dt = time_statistic(m, p4, n, 2)
        res4["evals_per_s_" + key] = n / dt
        kt, _ = kernel_times(m, p4, n, 1)
        res4["roofline_" + key] = valu_roofline(kt, "float" if key == "fp32" else "double", "valu_cfg3_cfg4.md",
                                                ["true"])
    out["cfg4_expdisk3d_512"] = res4
    log("cfg4 (3-D disk + bulge, 512^2, PSF 25): %.1f evals/s fp32, %.1f evals/s fp64" %
        (res4["evals_per_s_fp32"], res4["evals_per_s_fp64"]))
    w3, pop = workloads.cfg3_population()
    m3 = w3.fit_model(imfit_amd.PRECISION_FP32)
    m3.Autotune(pop[:1024])   # the population grids' shapes (kept for the full population)
    m3.GetFitStatisticBatch(pop)   # untimed: builds the 4096-vector plan and its graph
    device_sync()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        chi = m3.GetFitStatisticBatch(pop)
    device_sync()
    dt = (time.perf_counter() - t0) / reps
    kt3, _ = kernel_times(m3, None, 2, 1, call=lambda: m3.GetFitStatisticBatch(pop))
    out["cfg3_population_512"] = {"vectors": int(pop.shape[0]), "s_per_batch": dt,
                                  "chi2_per_s": pop.shape[0] / dt, "finite": bool(np.isfinite(chi).all()),
                                  "roofline_fp32": valu_roofline(kt3, "float", "valu_cfg3_cfg4.md")}
    log("cfg3 (4096-vector population, 512^2, PSF 35, FP32): %.3f s per batch, %.0f chi^2/s" %
        (dt, pop.shape[0] / dt))
    if fits:
        out["cfg5_cutouts"] = cutout_benchmark()
        out["bootstrap_cfg1"] = bootstrap_benchmark()
        out["de_cfg1"] = de_benchmark()
    return out


def de_benchmark(seed=1):
    """imfit's Differential Evolution (DiffEvolnFit) on the 256^2 Sersic+Exp+sky image
    (cfg1, 12 parameters, population 96) within the LM limits: each generation's trials
    in one batched launch, against DESolver's sequential loop (one evaluation per
    trial); both with the same seed, compared with the LM minimum."""
    import imfit_amd
    from imfit_amd import workloads
    w = workloads.cfg1()
    w.make_data()
    lim = w.lm_limits()
    m = w.fit_model(imfit_amd.PRECISION_FP64)
    x0 = w.lm_start(seed=1)
    best, _ = m.FitLM(x0, limits=lim, ftol=1e-10, method="normal")
    c_lm = m.GetFitStatistic(best)
    res = {"image": "256x256, no PSF", "params": int(x0.size), "chi2_lm": c_lm}
    for mode, key in ((imfit_amd.DE_GENERATIONAL, "generational"), (imfit_amd.DE_SEQUENTIAL, "sequential")):
        m.FitDE(x0, lim, seed=seed, max_generations=2, mode=mode)   # plans (untimed)
        device_sync()
        t0 = time.perf_counter()
        x, r = m.FitDE(x0, lim, seed=seed, mode=mode)
        dt = time.perf_counter() - t0
        c = m.GetFitStatistic(x)
        res[key] = {"s_per_fit": dt, "generations": r.generations, "population": r.population,
                    "evals_per_s": r.generations * r.population / dt, "status": r.status,
                    "chi2_rel_vs_lm": (c - c_lm) / c_lm}
        log("DE (%s) on cfg1: %d generations x %d in %.2f s (%.0f evals/s), chi^2 rel. to LM %.1e" %
            (key, r.generations, r.population, dt, r.generations * r.population / dt, (c - c_lm) / c_lm))
    res["speedup_generational"] = res["sequential"]["s_per_fit"] / res["generational"]["s_per_fit"]
    return res


def bootstrap_benchmark(n=256, n_serial=4, seed=10):
    """Bootstrap error estimation on the 256^2 Sersic+Exp+sky image (cfg1): the main fit,
    then n resamples (imfit's draw rule, MT19937 seeded with `seed`) fitted from its best
    values -- all at once by the batched normal-equation LM (one model per sample),
    versus imfit's loop (one mpfit fit with a finite-difference Jacobian per sample, on
    the GPU) on the first n_serial samples; chi^2 compared on those."""
    import imfit_amd
    from imfit_amd import workloads
    w = workloads.cfg1()
    w.make_data()
    lim = w.lm_limits()
    m = w.fit_model(imfit_amd.PRECISION_FP64)
    best, _ = m.FitLM(w.lm_start(seed=1), limits=lim, ftol=1e-10, method="normal")
    models = [w.fit_model(imfit_amd.PRECISION_FP64) for _ in range(n)]
    imfit_amd.init_genrand(seed)
    models[0].UseBootstrap()
    for mk in models:
        mk.MakeBootstrapSample()
    b = imfit_amd.BatchGPU(models)
    b.GetFitStatistic([best] * n)   # plans and graphs
    device_sync()
    t0 = time.perf_counter()
    xb, rb = b.FitLM([best] * n, limits=[lim] * n, ftol=1e-10)
    dt = time.perf_counter() - t0
    imfit_amd.init_genrand(seed)
    m.UseBootstrap()
    rel = []
    t1 = time.perf_counter()
    for k in range(n_serial):
        m.MakeBootstrapSample()
        x, r = m.FitLM(best, limits=lim, ftol=1e-10, jacobian="fd")
        c_serial, c_batch = m.GetFitStatistic(x), m.GetFitStatistic(xb[k])
        rel.append(abs(c_batch - c_serial) / c_serial)
    dts = time.perf_counter() - t1
    res = {"samples": n, "image": "256x256, no PSF", "batched_fits_per_s": n / dt,
           "batched_mean_niter": float(np.mean([r.niter for r in rb])),
           "serial_fd_mpfit_fits_per_s": n_serial / dts, "serial_sample": n_serial,
           "chi2_rel_diff_max_on_sample": float(max(rel)), "speedup": (n / dt) / (n_serial / dts),
           "param_spread_rel": [float(v) for v in np.std(np.array(xb), axis=0) / np.abs(best)]}
    log("bootstrap (cfg1, %d resamples): batched %.1f fits/s, serial mpfit %.2f fits/s (x%.0f); "
        "chi2 rel diff on the sample <= %.2e" % (n, n / dt, n_serial / dts, res["speedup"], max(rel)))
    del b, models
    return res


def cutout_benchmark(n=2000, n_plain=100, precisions=("fp64", "fp32"), resident=True, resident_only=False):
    """Many-cutout fits: n cutouts (64-512 px, 1-3 components, 25^2 Moffat PSF) fitted
    all at once by the batched normal-equation LM (analytic Jacobians, J^T J and J^T r
    reduced on the GPU in FP64, every model's step in one batched evaluation; the
    default host-driven rounds, and with `resident` the same fits resident on the GPU
    beside them), versus plain GPU fits one cutout after another on a sample of
    n_plain: mpfit with finite-difference Jacobians (chi^2 compared), and the
    normal-equation LM with batched finite differences.  resident_only: the resident
    fits alone."""
    import imfit_amd
    from imfit_amd import workloads
    ws = workloads.cutouts(n)
    for w in ws:
        w.make_data()
    starts = [w.lm_start(seed=k) for k, w in enumerate(ws)]
    lims = [w.lm_limits() for w in ws]
    res = {"cutouts": n, "sizes_px": "64-512 log-uniform per side (seed 4)"}
    for prec, key in ((imfit_amd.PRECISION_FP64, "fp64"), (imfit_amd.PRECISION_FP32, "fp32")):
        if key not in precisions:
            continue
        models = [w.fit_model(prec) for w in ws]
        b = imfit_amd.BatchGPU(models)
        b.Autotune(starts)
        b.GetFitStatistic(starts)   # plans and graphs
        b.SetLMResident(resident_only)   # host-driven rounds: the default driver
        b.FitLM(starts, limits=lims, ftol=1e-10, maxiter=2)   # warm-up: the batch's Jacobian and deviate plans
        device_sync()
        t0 = time.perf_counter()
        xb, rb = b.FitLM(starts, limits=lims, ftol=1e-10)
        dt = time.perf_counter() - t0
        res["batched_fits_per_s_" + key] = n / dt
        res["batched_mean_niter_" + key] = float(np.mean([r.niter for r in rb]))
        res["batched_driver"] = "resident" if resident_only else "host rounds"
        log("cfg5 batched normal-equation LM, %s (%s): %d fits in %.2f s -> %.1f fits/s (mean niter %.1f)" %
            (res["batched_driver"], key, n, dt, n / dt, res["batched_mean_niter_" + key]))
        if resident_only or not resident:
            del b, models
            continue
        # the same fits resident on the GPU (LM steps in kernels, one host wait per 8 rounds)
        b.SetLMResident(True)
        b.FitLM(starts, limits=lims, ftol=1e-10, maxiter=2)   # warm-up, as above
        device_sync()
        t0 = time.perf_counter()
        xr, rr = b.FitLM(starts, limits=lims, ftol=1e-10)
        dtr = time.perf_counter() - t0
        b.SetLMResident(False)
        c_res, c_host = b.GetFitStatistic(xr), b.GetFitStatistic(xb)
        rel_rh = np.abs(c_res - c_host) / c_host
        res["resident_fits_per_s_" + key] = n / dtr
        res["resident_speedup_vs_host_rounds_" + key] = dt / dtr
        res["resident_matches_host_rounds_chi2_1e-9_" + key] = int(np.sum(rel_rh <= 1e-9))
        res["resident_lower_chi2_than_host_rounds_" + key] = int(np.sum((rel_rh > 1e-9) & (c_res < c_host)))
        log("cfg5 resident (%s): %.1f fits/s; resident / host rounds %.2fx; same chi^2 (1e-9) on %d of %d, "
            "resident lower on %d" % (key, n / dtr, dt / dtr, res["resident_matches_host_rounds_chi2_1e-9_" + key], n,
                                      res["resident_lower_chi2_than_host_rounds_" + key]))
        if prec == imfit_amd.PRECISION_FP64 and n_plain > 0:
            t0 = time.perf_counter()
            rel, c_plain_all = [], []
            for k in range(n_plain):
                x, r = models[k].FitLM(starts[k], limits=lims[k], ftol=1e-10, jacobian="fd")
                c_plain, c_batch = models[k].GetFitStatistic(x), models[k].GetFitStatistic(xb[k])
                rel.append(abs(c_batch - c_plain) / c_plain)
                c_plain_all.append(c_plain)
            dtp = time.perf_counter() - t0
            # control: mpfit itself with the exact Jacobian (same start, limits, ftol) against
            # mpfit with forward differences, and the batch against one model's own
            # normal-equation LM -- separates the problems' sensitivity to the Jacobian
            # (several minima, parameters pegged at limits) from anything the batching does
            rel_ctrl, same_single = [], 0
            for k in range(n_plain):
                xa, _ = models[k].FitLM(starts[k], limits=lims[k], ftol=1e-10, jacobian="analytic")
                rel_ctrl.append(abs(models[k].GetFitStatistic(xa) - c_plain_all[k]) / c_plain_all[k])
                xn, _ = models[k].FitLM(starts[k], limits=lims[k], ftol=1e-10, method="normal")
                same_single += int(models[k].GetFitStatistic(xn) == models[k].GetFitStatistic(xb[k]))
            res["control_mpfit_analytic_matched_chi2_1e-9"] = int(sum(v <= 1e-9 for v in rel_ctrl))
            res["host_rounds_equal_single_model_normal_lm"] = same_single
            res["batch_lower_chi2_than_fd_mpfit"] = int(sum(1 for k in range(n_plain) if rel[k] > 1e-9 and
                                                            models[k].GetFitStatistic(xb[k]) < c_plain_all[k]))
            # the normal-equation LM with batched forward differences, one cutout at a time
            t1 = time.perf_counter()
            for k in range(n_plain):
                models[k].FitLM(starts[k], limits=lims[k], ftol=1e-10, method="normal", jacobian="fd")
            dtn = time.perf_counter() - t1
            res["plain_fd_normal_fits_per_s"] = n_plain / dtn
            res["speedup_fp64_vs_fd_normal"] = (n / dt) / (n_plain / dtn)
            matched = [r for r in rel if r <= 1e-9]
            res["plain_fd_mpfit_fits_per_s"] = n_plain / dtp
            res["plain_sample"] = n_plain
            # A cutout whose fit has two minima can end in either, with either method: its
            # end point then depends on rounding (from starts perturbed by 1e-9, mpfit
            # itself lands in both), so matched and unmatched cases are reported apart.
            res["sample_matched_chi2_1e-9"] = len(matched)
            res["chi2_rel_diff_max_matched"] = float(max(matched)) if matched else None
            res["chi2_rel_diff_max_on_sample"] = float(max(rel))
            res["speedup_fp64"] = (n / dt) / (n_plain / dtp)
            log("cfg5 plain FD mpfit (fp64, one cutout at a time): %.2f fits/s; same minimum (chi2 within 1e-9) "
                "on %d of %d sampled cutouts (max rel diff %.2e there), max rel diff overall %.2e" %
                (n_plain / dtp, len(matched), n_plain, max(matched) if matched else float("nan"), max(rel)))
            log("cfg5 control: mpfit with the exact Jacobian matches FD mpfit on %d of %d; the host-driven batch "
                "equals the single-model normal-equation LM on %d of %d; of the batch's mismatches %d end lower" %
                (res["control_mpfit_analytic_matched_chi2_1e-9"], n_plain, same_single, n_plain,
                 res["batch_lower_chi2_than_fd_mpfit"]))
        del b, models
    return res


def time_statistic(m, p, steps, warmup):
    """`warmup` untimed GetFitStatistic calls, then `steps` timed ones between device
    synchronisations; returns (statistic, seconds)."""
    for _ in range(warmup):
        m.GetFitStatistic(p)
    device_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        stat = m.GetFitStatistic(p)
    device_sync()
    return stat, time.perf_counter() - t0


def kernel_times(m, p, steps, warmup, call=None):
    """The same calls again, eager, each kernel carrying a pair of HIP events on the
    library's stream (hipExtLaunchKernel: the dispatch's own start and end
    timestamps); returns ({kernel: (total ms, launches)}, ms per call)."""
    call = call or (lambda: m.GetFitStatistic(p))
    m.SetKernelTiming(True)
    for _ in range(warmup):
        call()
    m.SetKernelTiming(True)   # clears the warm-up totals
    t1 = time.perf_counter()
    for _ in range(steps):
        call()
    device_sync()
    ms = 1e3 * (time.perf_counter() - t1) / steps
    kt = m.GetKernelTiming()
    m.SetKernelTiming(False)
    return kt, ms


def roofline(w, kt, prec):
    """Roofline of the kernel with the most event time: algorithmic bytes per launch
    (kernel_model) over its average duration against 8 TB/s; `traffic` = HBM bytes per
    launch from the committed FETCH_SIZE / WRITE_SIZE profile of the same workload; and
    the VALU issue fraction from the committed SQ_INSTS_VALU profile."""
    tname = "float" if prec == "fp32" else "double"
    km = kernel_model(w, 4 if prec == "fp32" else 8)
    dom = max(kt.items(), key=lambda kv: kv[1][0])[0]
    avg_s = kt[dom][0] / kt[dom][1] * 1e-3
    ach = km[dom]["bytes"] / avg_s / 1e9 if dom in km else None
    # VALU issue time (4 cycles per wave64 instruction on one of the 1024 SIMDs at
    # 2.4 GHz) as a fraction of the kernel's duration
    valu = profiled_valu(dom, tname)
    return {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": ach / PEAK_HBM_GBS if ach is not None else None, "traffic": profiled_traffic(dom, tname),
            "kernel": dom, "avg_us": avg_s * 1e6, "alg_bytes": km.get(dom, {}).get("bytes"),
            "valu_issue_frac": 4.0 * valu / 1024 / 2.4e9 / avg_s if valu is not None else None}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp64",
                    help="arithmetic of the headline value (fp64: the reference's; fp32 is reported beside it)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lm", action="store_true")
    ap.add_argument("--lm-repeats", type=int, default=3)
    ap.add_argument("--copy-calibration-mb", type=int, default=0)
    ap.add_argument("--no-aux", action="store_true", help="skip the other workloads' side measurements")
    ap.add_argument("--no-gradient", action="store_true",
                    help="skip the gradient timings (counter runs of the plain evaluation alone)")
    ap.add_argument("--aux-only", action="store_true",
                    help="only the 3-D disk and population measurements (profiling runs of those kernels)")
    ap.add_argument("--cutouts-only", action="store_true",
                    help="only the many-cutout LM fits (resident and host-driven batches, plain fits)")
    ap.add_argument("--cutouts-resident-fp64", action="store_true",
                    help="with --cutouts-only: the resident FP64 batch fit alone (profiling runs)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes (profiling runs: no tuning trials in the trace)")
    ap.add_argument("--launch-shapes", default=None,
                    help="main and region shapes 'C,R,C,R' (columns, rows per workgroup) of the headline model "
                         "instead of tuning them (counter runs repeat the shapes a timed run chose)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; sharding the path over several GPUs is a design "
              "only, not built. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import imfit_amd
    from imfit_amd import workloads

    if args.aux_only:
        print(json.dumps(aux_benchmarks(args.steps, fits=False)))
        return
    if args.cutouts_only:
        if args.cutouts_resident_fp64:
            print(json.dumps(cutout_benchmark(n_plain=0, precisions=("fp64",), resident_only=True)))
        else:
            print(json.dumps(cutout_benchmark()))
        return
    log("building the 1024^2 workload (PSFs and data on the GPU)")
    w = workloads.cfg2()
    p = w.start_params()
    fixed_shapes = args.no_autotune or args.launch_shapes is not None
    other = "fp32" if args.precision == "fp64" else "fp64"
    models, results = {}, {}
    for key in (args.precision, other):
        m = w.fit_model(imfit_amd.PRECISION_FP32 if key == "fp32" else imfit_amd.PRECISION_FP64)
        trials = 0
        if key == args.precision and args.launch_shapes is not None:
            m.GetFitStatistic(p)   # builds the grids
            v = [int(t) for t in args.launch_shapes.split(",")]
            for grid in range(len(v) // 2):
                m.SetLaunchShape(grid, v[2 * grid], v[2 * grid + 1])
        elif not fixed_shapes:
            trials = m.Autotune(p)   # launch shapes for this geometry (untimed, once)
        # timed region: K evaluations, each replaying the captured graph of the call
        stat, dt = time_statistic(m, p, args.steps, args.warmup)
        kt, ms_timed = kernel_times(m, p, args.steps, args.warmup)
        shapes = {"main_cols_rows": list(m.GetLaunchShape(0)), "region_cols_rows": list(m.GetLaunchShape(1)),
                  "autotune_trials": trials}
        results[key] = {"stat": stat, "s": dt, "evals_per_s": args.steps / dt, "ms_per_step": 1e3 * dt / args.steps,
                        "ms_per_step_with_kernel_events": ms_timed, "kt": kt, "shapes": shapes}
        models[key] = m
        log("%s: chi2 = %.10e, %.4f ms/eval, %.1f evals/s (shapes %s)" %
            (key, stat, 1e3 * dt / args.steps, args.steps / dt, shapes))
        log("%s with kernel events: %.4f ms/eval" % (key, ms_timed))
        for k, (tot, n) in sorted(kt.items(), key=lambda kv: -kv[1][0]):
            log("  %-22s %9.2f us avg  x%d" % (k, 1e3 * tot / n, n))
    head, side = results[args.precision], results[other]
    rel_fp32_fp64 = abs(results["fp32"]["stat"] - results["fp64"]["stat"]) / abs(results["fp64"]["stat"])
    log("statistic fp32 vs fp64: rel diff %.3e" % rel_fp32_fp64)
    roof = roofline(w, head["kt"], args.precision)
    side_roof = roofline(w, side["kt"], other)

    grad = None if args.no_gradient else gradient_benchmark(models, p, results, args.steps)
    lm = None if args.no_lm else lm_benchmark(w, args.lm_repeats)
    aux = None if args.no_aux else aux_benchmarks(args.steps)

    cpu = None
    if not args.no_cpu:
        log("timing the CPU oracle baseline")
        cpu = cpu_baseline(w, p, {"fp64": results["fp64"]["stat"], "fp32": results["fp32"]["stat"]})

    def kernels_us(r):
        return {k: 1e3 * t / n for k, (t, n) in r["kt"].items()}

    out = {"metric": METRIC, "value": head["evals_per_s"], "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": args.precision, "data": "synthetic",
           "config": {"workload": w.name, "image": "1024x1024", "psf": "51x51 Moffat",
                      "oversampled_region": "40x40 at 5x, 255x255 PSF",
                      "precision": "%s pixels, FFT and reduction" % args.precision if args.precision == "fp64"
                      else "fp32 pixels and FFT, FP64 deviate reduction",
                      "ms_per_step_with_kernel_events": head["ms_per_step_with_kernel_events"],
                      "kernels_us": kernels_us(head), "launch_shapes": head["shapes"],
                      other: {"evals_per_s": side["evals_per_s"], "ms_per_step": side["ms_per_step"],
                              "steps": args.steps, "kernels_us": kernels_us(side), "launch_shapes": side["shapes"],
                              "roofline": side_roof,
                              "precision": "fp32 pixels and FFT, FP64 deviate reduction" if other == "fp32"
                              else "fp64 pixels, FFT and reduction"},
                      "fp32_vs_fp64_stat_rel": rel_fp32_fp64},
           "roofline": roof, "cpu_baseline": cpu, "gradient": grad, "lm_fits": lm, "other_configs": aux}
    print(json.dumps(out))
    if args.copy_calibration_mb > 0:
        copy_calibration(args.copy_calibration_mb)


if __name__ == "__main__":
    main()
