#!/usr/bin/env python3
"""SRM throughput on one MI355X: point-harmonics per second.

A step is one SRM realization on the device with its harmonic table already
resident in HBM: per-axis operand tables, the field kernel, the FP64
statistics pass and the rescale/taper epilogue (scarf_srm_run). Host
sampling of the harmonics (the reference's draw loop) is timed separately.

Default workload (the configuration the metric is quoted on): structured
512 x 512 x 512 grid, dh = 20 m, von Karman Hurst 0.1, cl = 5000 m,
sigma = 0.05, 20000 harmonics, seed 1234. The primary value is the split-bf16
(3-term, 6-product) MFMA kernel; the FP32 direct-sum kernel is measured on
the same resident table for the speed-up, and every kernel's error is
measured over the whole field against the FP64 kernel, itself checked
against the FP64 CPU oracle at sampled points.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

METRIC = "SRM Gpt·harm/s at 512³×20k: split-BF16 MFMA vs FP32 direct, err≤1e-5σ"
PEAK_MFMA_TFLOPS = 2500.0  # MI355X dense bf16 / fp16
PEAK_VALU_TFLOPS = 157.3   # MI355X FP32 vector
PEAK_VALU_F64_TFLOPS = 78.6  # MI355X FP64 vector
MFMA_PRODUCTS = {"mfma_bf16x3": 6, "mfma_f16x2": 3, "mfma_bf16x2": 3}

# Workloads. flavour: library `real` (f = libscarf3df, d = libscarf3d);
# kernel: the timed field kernel; compare: kernels also timed on the same
# resident table; every field is checked against the FP64 kernel over the
# whole field and against the FP64 CPU oracle at 256 sampled points.
WORKLOADS = {
    # 512^3 x 20k, the configuration the metric is quoted on (seed 1234)
    "precision_512": dict(n=(512, 512, 512), ds=20.0, acf=0, hurst=0.1, cl=(5000.0,) * 3, sigma=0.05, nharm=20000,
                          seed=1234, flavour="f", kernel="mfma_bf16x3",
                          compare=("direct_f32", "mfma_f16x2", "mfma_bf16x2")),
    # 128^3 x 5k with FP64 host semantics (double library); the reference's
    # CPU path on one rank is timed beside the GPU FP32 direct-sum kernel
    "plain_128": dict(n=(128, 128, 128), ds=50.0, acf=0, hurst=0.25, cl=(2000.0,) * 3, sigma=0.05, nharm=5000,
                      seed=1234, flavour="d", kernel="direct_f32", compare=("mfma_bf16x3", "mfma_f16x2", "direct_f64"),
                      cpu_one_rank=True),
    # 2,000,000 points uniform in 10 x 10 x 5 km (numpy PCG64 seed 7, rounded
    # to float), Gaussian cl = (3000, 3000, 1000) m, dh = 50 m, 10k harmonics
    "points_2m": dict(points=2_000_000, box=(10000.0, 10000.0, 5000.0), ds=50.0, acf=1, hurst=0.25,
                      cl=(3000.0, 3000.0, 1000.0), sigma=0.1, nharm=10000, seed=7, flavour="f", kernel="direct_f32",
                      compare=("direct_f64",)),
    # 1024 x 1024 x 512 x 50k, 16 POIs (numpy seed 16), mute 20 m, taper 500 m,
    # rescale to sigma (seed 1234)
    "large_1024": dict(n=(1024, 1024, 512), ds=20.0, acf=0, hurst=0.2, cl=(6000.0, 4000.0, 1500.0), sigma=0.08,
                       nharm=50000, seed=1234, rescale=1, npoi=16, mute=20.0, taper=500.0, flavour="f",
                       kernel="mfma_bf16x3", compare=("mfma_f16x2",)),
}
KERNELS = ("mfma_bf16x3", "mfma_f16x2", "mfma_bf16x2", "direct_f32", "direct_f64")
# HBM bytes per launch of the timed kernel, measured with rocprofv3 FETCH_SIZE /
# WRITE_SIZE passes scaled by a 1 GiB device copy (the counters cannot be read
# inside the timed run): 1.038e10 - 1.044e10 B on each of the x-128 tiles the
# tuner picks from (32x8, 64x4, 128x2, 256x1 columns), the mean reported:
# (workload, kernel) -> (bytes, summary file)
TRAFFIC = {
    ("precision_512", "mfma_bf16x3"): (1.041e10, "profiles/mfma_grid_bf16x3_512cube_20k_traffic.md"),
}


def _sig(v, digits=4):
    """v rounded to `digits` significant digits (errors of 1e-6 stay visible)."""
    return float(f"{v:.{digits}g}")


def ensemble_members():
    """The 64-realization ensemble: uneven grids 96^3 .. 384x384x192, dh 25-100
    m, von Karman (Hurst 0.05-0.5) and Gaussian spectra, cl 500-8000 m, 2k-16k
    harmonics, seeds 1..64. Gaussian members keep cl <= span/2 so that their
    sampling band holds spectral power (with cl far above the box the band sits
    ~1e-40 below the peak and rejection sampling never ends)."""
    out = []
    for r in range(1, 65):
        nx = 96 + 32 * ((r - 1) % 10)
        nz = min(nx, 96 + 32 * ((r - 1) % 4))
        dh = 25.0 * (1 + (r - 1) % 4)
        acf = (r + 1) % 2
        hurst = round(0.05 + 0.05 * ((r - 1) % 10), 2)
        cl = 500.0 * 2 ** ((r - 1) % 5)
        if acf == 1:
            cl = min(cl, 0.5 * (min(nx, nz) - 1) * dh)
        out.append(dict(nd=3, fs=(1, 1, 1), fe=(nx, nx, nz), ds=dh, dh=dh, acf=acf, cl=(cl, cl, cl), sigma=0.05,
                        hurst=hurst, nharm=2000 * (1 + (r - 1) % 8), seed=r))
    return out


def bench_ensemble(args):
    import scarf3d_amd
    from scarf3d_amd.hip import Hip

    lib = scarf3d_amd.load("f")
    hip = Hip()
    members = ensemble_members()
    work = float(sum(np.prod(m["fe"]) * m["nharm"] for m in members))
    bufs = [hip.empty(int(np.prod(m["fe"])), np.float32) for m in members]
    ptrs = [b.ptr for b in bufs]
    t0 = time.perf_counter()
    log(f"[bench] ensemble of {len(members)}: first run (draws, tile tuning per shape, graph capture, run)")
    stats = lib.ensemble_execute(members, ptrs, args.kernel)
    log(f"[bench] ensemble first run {time.perf_counter() - t0:.2f} s")
    for _ in range(max(0, args.warmup - 1)):
        lib.ensemble_execute(members, ptrs, args.kernel)
    dev_ms, wall = [], []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        stats = lib.ensemble_execute(members, ptrs, args.kernel)
        wall.append(time.perf_counter() - t1)
        dev_ms.append(lib.ensemble_ms())
    host_s = float(np.mean(stats[:, 4].astype(np.float64)))
    ms = float(np.mean(dev_ms))
    P = MFMA_PRODUCTS[args.kernel]
    achieved = work * 4 * P / (ms * 1e-3) / 1e12
    out = {
        "metric": "SRM ensemble Gpt·harm/s: 64 uneven realizations, one persistent grouped-GEMM launch in a hipGraph",
        "value": round(work / (ms * 1e-3) / 1e9, 2), "unit": "Gpt·harm/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": args.kernel, "data": "synthetic",
        "config": {"workload": "ensemble_64", "members": len(members), "point_harmonics": work,
                   "kernel": args.kernel, "host_draw_s_per_step_wall": round(float(np.mean(wall)) - ms * 1e-3, 3),
                   "host_draw_s_mean_member": round(host_s, 3)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 1), "peak": PEAK_MFMA_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_MFMA_TFLOPS, 4), "traffic": None,
                     "kernel": f"mfma_ensemble<{args.kernel}>", "flops_per_point_harmonic": 4 * P},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    for b in bufs:
        b.free()
    return 0


def bench_copy_calibration(args):
    """A device-to-device copy of a known size (1 GiB): the reference point
    that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters into bytes."""
    from scarf3d_amd.hip import Hip

    import ctypes as C

    hip = Hip()
    n = 1 << 28  # floats
    a = hip.empty(n, np.float32)
    b = hip.empty(n, np.float32)
    a.zero()
    for _ in range(max(1, args.steps)):
        rc = hip.L.hipMemcpy(C.c_void_p(b.ptr), C.c_void_p(a.ptr), C.c_size_t(4 * n), 3)
        assert rc == 0
    hip.synchronize()
    print(json.dumps({"copy_bytes": 4 * n, "copies": max(1, args.steps)}), flush=True)
    a.free()
    b.free()
    return 0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def init(lib, w):
    from srm_cases import init_product, structured_cfg, unstructured_cfg

    real = np.float32 if w["flavour"] == "f" else np.float64
    if "points" in w:
        cfg = unstructured_cfg(npts=w["points"], box=w["box"], dh=w["ds"], acf=w["acf"], cl=w["cl"], sigma=w["sigma"],
                               hurst=w["hurst"], nharm=w["nharm"], seed=w["seed"], real=real)
    else:
        kw = {k: w[k] for k in ("n", "ds", "acf", "hurst", "cl", "sigma", "nharm") if k in w}
        extra = {}
        if w.get("npoi"):
            rng = np.random.default_rng(16)
            ext = np.array([(w["n"][i] - 1) * w["ds"] for i in range(3)])
            extra = dict(poi=(rng.uniform(0, 1, (3, w["npoi"])) * ext[:, None]), mute=w["mute"], taper=w["taper"])
        cfg = structured_cfg(rescale=w.get("rescale", 0), real=real, **kw, **extra)
    init_product(lib, cfg)
    return cfg


def time_kernel(lib, hip, kernel, d_field, steps, warmup):
    assert lib.set_kernel(kernel) == 0
    for _ in range(warmup):
        lib.srm_run(d_field.ptr)
    hip.synchronize()
    t0 = time.perf_counter()
    field_ms, total_ms = [], []
    for _ in range(steps):
        stats = lib.srm_run(d_field.ptr)
        t = lib.timings()
        field_ms.append(t["field"])
        total_ms.append(t["total"])
    hip.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return dict(ms_per_step=wall * 1e3, field_ms=float(np.mean(field_ms)), device_ms=float(np.mean(total_ms)),
                stats=stats)


def _host_threads():
    return min(16, len(os.sched_getaffinity(0)))


def cpu_baseline(oracle, cfg, seed, flavour="f", threads=None, seconds_target=15.0):
    """The reference loop (harmonic outer, points inner, libm sin/cos at
    `real` precision) on the host cores, each thread owning a block of
    z-slabs (or of points) like an MPI rank, over a bounded subset of the
    same workload: calibrated on a short run, then sized to about
    seconds_target."""
    import ctypes as C

    from oracle_lib import _DP

    p = oracle.params(cfg)
    s = oracle.setup(p, flavour)
    harm, used, _ = oracle.draw(p, s, seed, flavour)
    threads = threads or _host_threads()
    nh = int(used.sum())
    structured = cfg.get("structured", True)
    if structured:
        nx, ny, nz = cfg["fe"][0], cfg["fe"][1], cfg["fe"][2]
        unit, total = nx * ny, nz  # points per slab, slabs
    else:
        unit, total = 1, p.npts
    real = np.float32 if flavour == "f" else np.float64

    def run(nunits, nharm_used):
        u = used.copy()
        idx = np.flatnonzero(u)
        u[idx[nharm_used:]] = 0
        out = np.zeros(unit * nunits, dtype=real)
        t0 = time.perf_counter()
        if structured:
            fn = getattr(oracle.L, f"orc_baseline_structured_{flavour}")
            fn(C.byref(p), C.byref(s), harm.ctypes.data_as(_DP), u.ctypes.data_as(C.POINTER(C.c_uint8)), 0, nunits,
               out.ctypes.data_as(C.c_void_p), threads)
        else:
            oracle.L.orc_baseline_unstructured_f(C.byref(p), C.byref(s), harm.ctypes.data_as(_DP),
                                                 u.ctypes.data_as(C.POINTER(C.c_uint8)), C.c_long(0),
                                                 C.c_long(nunits), out.ctypes.data_as(C.c_void_p), threads)
        return time.perf_counter() - t0

    probe_units = min(total, threads if structured else 200_000)
    probe_h = min(nh, 200)
    tp = run(probe_units, probe_h)
    rate = unit * probe_units * probe_h / tp
    n_units = int(seconds_target * rate / (unit * nh))
    step = threads if structured else 1
    n_units = min(total, max(step, n_units // step * step))
    hh = nh
    if unit * n_units * nh / rate > 4 * seconds_target:  # even the smallest subset is too long: cut harmonics
        hh = max(probe_h, int(seconds_target * rate / (unit * n_units)))
    t = run(n_units, hh)
    wps = unit * n_units * hh / t
    what = f"{n_units} z-slabs of {nx}x{ny}" if structured else f"{n_units} points"
    return dict(value=wps / 1e9, unit="Gpt·harm/s", cores=threads, kind="port",
                sample=f"{what} x {hh} harmonics ({t:.1f} s), reference loop order, "
                       f"{'FP32' if flavour == 'f' else 'FP64'} libm sin/cos, {threads} OpenMP thread(s) each owning "
                       f"{'z-slabs' if structured else 'a point range'}")


def oracle_points(oracle, cfg, seed, flavour, nsamp=256):
    """FP64 CPU oracle at sampled points of the workload (same draw as the
    library): returns (flat indices, values)."""
    p = oracle.params(cfg)
    s = oracle.setup(p, flavour)
    harm, used, _ = oracle.draw(p, s, seed, flavour)
    rng = np.random.default_rng(5)
    if cfg.get("structured", True):
        n = [cfg["fe"][a] - cfg["fs"][a] + 1 for a in range(3)]
        ii, jj, kk = (rng.integers(0, n[a], nsamp) for a in range(3))
        ref = oracle.field_at(p, s, harm, used, ii, jj, kk)
        return ii + n[0] * (jj + n[1] * kk), ref
    flat = rng.integers(0, p.npts, nsamp)
    sub = dict(cfg, x=cfg["x"][flat], y=cfg["y"][flat], z=cfg["z"][flat])
    p2 = oracle.params(sub)  # same parameters, the sampled points only
    return flat, oracle.field_unstructured(p2, s, harm, used)


def epilogue_at(vals, flat, cfg, mean, var):
    """Rescale (src/scarflib_srm.f90:897-912: field * sigma/sqrt(var) - mean)
    and the taper/mute of every POI in order (scarflib_common.f90:881-941) at
    the structured points `flat`, in FP64, given the FP64 field's moments."""
    out = np.array(vals, dtype=np.float64)
    if cfg.get("rescale") == 1:
        out = out * (cfg["sigma"] / np.sqrt(var)) - mean
    poi = cfg.get("poi")
    if poi is not None and cfg.get("structured", True):
        n = [cfg["fe"][a] - cfg["fs"][a] + 1 for a in range(3)]
        idx = [flat % n[0], (flat // n[0]) % n[1], flat // (n[0] * n[1])]
        pts = np.stack([(idx[a] + cfg["fs"][a] - 1) * cfg["ds"] for a in range(3)])
        mute, taper = cfg["mute"], cfg["taper"]
        for q in range(np.asarray(poi).shape[1]):
            d = np.sqrt(((pts - np.asarray(poi)[:, q:q + 1]) ** 2).sum(axis=0))
            f = np.where(d <= mute, 0.0, np.where(d <= mute + taper,
                                                  0.5 - 0.5 * np.cos(np.pi * (d - mute) / taper), 1.0))
            out = out * f
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="precision_512", choices=sorted(WORKLOADS) + ["ensemble_64", "copy_calib"])
    ap.add_argument("--kernel", default=None, choices=KERNELS, help="timed kernel (default: the workload's)")
    ap.add_argument("--no-compare", action="store_true", help="skip the comparison kernels and the FP64 reference")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the SRM shards by z-slab ranges with no exchange inside the harmonic "
              "loop, and multi-GPU runs are not implemented; run with --gpus 1", file=sys.stderr)
        return 2

    if args.workload == "ensemble_64":
        args.kernel = args.kernel or "mfma_bf16x3"
        return bench_ensemble(args)
    if args.workload == "copy_calib":
        return bench_copy_calibration(args)

    import scarf3d_amd
    from oracle_lib import Oracle
    from scarf3d_amd.hip import Hip

    w = WORKLOADS[args.workload]
    kernel = args.kernel or w["kernel"]
    flavour = w["flavour"]
    real = np.float32 if flavour == "f" else np.float64
    lib = scarf3d_amd.load(flavour)
    cfg = init(lib, w)
    structured = cfg.get("structured", True)
    npts = int(np.prod(w["n"])) if structured else int(w["points"])
    work = float(npts) * w["nharm"]
    hip = Hip()
    d_field = hip.empty(npts, real)
    n_used, host_s = lib.srm_draw(w["seed"])
    shape = w["n"] if structured else f"{npts} points"
    log(f"[bench] {args.workload}: {shape} x {w['nharm']} harmonics ({n_used} kept), host draw {host_s:.3f} s")

    res = {kernel: time_kernel(lib, hip, kernel, d_field, args.steps, args.warmup)}
    tile = lib.last_tile() if kernel.startswith("mfma") else None
    main_field = d_field.numpy()
    log(f"[bench] {kernel}: {res[kernel]['field_ms']:.3f} ms field kernel, {res[kernel]['ms_per_step']:.3f} ms/step")

    extra = {}
    if not args.no_compare:
        for k in w["compare"]:
            if k == kernel:
                continue
            slow = k in ("direct_f32", "direct_f64") and work > 1e12
            steps = max(1, min(3, args.steps)) if slow else args.steps
            res[k] = time_kernel(lib, hip, k, d_field, steps, 1)
            log(f"[bench] {k}: {res[k]['field_ms']:.3f} ms field kernel")
            res[k]["field"] = d_field.numpy()
        d_field.free()
        # FP64 reference over the whole field, on the same resident table
        t0 = time.perf_counter()
        d_ref = hip.empty(npts, np.float64)
        st64 = lib.srm_run_fp64(d_ref.ptr)  # {count, mean, variance, sqrt} of the field before the epilogue
        hip.synchronize()
        ref64 = d_ref.numpy()
        d_ref.free()
        log(f"[bench] FP64 reference field in {time.perf_counter() - t0:.1f} s")
        sigma = cfg["sigma"]
        res[kernel]["field"] = main_field
        errs = {}
        for k, r in res.items():
            dif = r["field"].astype(np.float64) - ref64
            st3 = float(r["stats"][2])
            mx, rms = float(np.abs(dif).max()), float(np.sqrt(np.mean(dif * dif)))
            errs[k] = dict(max_err_sigma=mx / sigma, rms_err_sigma=rms / sigma, max_err_stats3=mx / st3)
            del dif
            log(f"[bench] {k}: max|err| = {errs[k]['max_err_sigma']:.2e} sigma, rms {errs[k]['rms_err_sigma']:.2e}")
        # the FP64 kernel and the timed kernel against the CPU oracle at sampled points (same draw)
        flat, ref_pts = oracle_points(Oracle(), cfg, w["seed"], flavour)
        ref_pts = epilogue_at(ref_pts, flat, cfg, st64[1], st64[2])
        extra["oracle_points"] = dict(
            n=len(flat), note="FP64 CPU oracle vs the timed kernel's field at sampled points (same draw)",
            max_err_sigma=float(np.abs(main_field[flat].astype(np.float64) - ref_pts).max() / sigma))
        extra["fp64_kernel_vs_oracle_max_err_sigma"] = float(np.abs(ref64[flat] - ref_pts).max() / sigma)
        log(f"[bench] oracle points: {extra['oracle_points']['max_err_sigma']:.2e} sigma; FP64 kernel vs oracle "
            f"{extra['fp64_kernel_vs_oracle_max_err_sigma']:.2e} sigma")
        del ref64
        for k in res:
            res[k].pop("field", None)
            res[k].update(errs.get(k, {}))
    else:
        d_field.free()

    main_r = res[kernel]
    P = MFMA_PRODUCTS.get(kernel)
    if P:
        achieved = work * 4 * P / (main_r["field_ms"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": _sig(achieved), "peak": PEAK_MFMA_TFLOPS, "unit": "TFLOP/s",
                "frac": _sig(achieved / PEAK_MFMA_TFLOPS), "traffic": None,
                "kernel": f"mfma_grid<{kernel}>", "flops_per_point_harmonic": 4 * P,
                "useful_frac": _sig(work * 4 / (main_r["field_ms"] * 1e-3) / 1e12 / PEAK_MFMA_TFLOPS)}
    elif kernel == "direct_f64":
        achieved = work * 10 / (main_r["field_ms"] * 1e-3) / 1e12
        roof = {"bound": "valu", "achieved": _sig(achieved), "peak": PEAK_VALU_F64_TFLOPS, "unit": "TFLOP/s",
                "frac": _sig(achieved / PEAK_VALU_F64_TFLOPS), "traffic": None,
                "kernel": "direct_grid_f64" if structured else "direct_points<f64>", "flops_per_point_harmonic": 10}
    else:
        achieved = work * 10 / (main_r["field_ms"] * 1e-3) / 1e12
        roof = {"bound": "valu", "achieved": _sig(achieved), "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
                "frac": _sig(achieved / PEAK_VALU_TFLOPS), "traffic": None,
                "kernel": "direct_grid_f32" if structured else "direct_points<f32>", "flops_per_point_harmonic": 10}

    if (args.workload, kernel) in TRAFFIC and tile is not None and tile[0] == 128:
        roof["traffic"], roof["traffic_source"] = TRAFFIC[(args.workload, kernel)]
    cpu = None
    if not args.no_cpu:
        oracle = Oracle()
        cpu = cpu_baseline(oracle, cfg, w["seed"], flavour)
        log(f"[bench] CPU baseline {cpu['value']:.4f} Gpt.harm/s on {cpu['cores']} threads ({cpu['sample']})")
        if w.get("cpu_one_rank"):
            one = cpu_baseline(oracle, cfg, w["seed"], flavour, threads=1)
            extra["cpu_baseline_one_rank"] = one
            log(f"[bench] CPU baseline, one rank: {one['value']:.4f} Gpt.harm/s ({one['sample']})")

    value = work / (main_r["ms_per_step"] * 1e-3) / 1e9
    dtype = {"mfma_bf16x3": "bf16x3-split MFMA, fp32 accumulate", "mfma_f16x2": "fp16x2-split (scaled) MFMA, fp32 "
             "accumulate", "mfma_bf16x2": "bf16x2-split MFMA, fp32 accumulate", "direct_f32": "f32 (FP64 phase "
             "reduction)", "direct_f64": "f64"}[kernel]
    config = {"workload": args.workload, "harmonics": w["nharm"], "dh_m": w["ds"],
              "acf": "von Karman" if w["acf"] == 0 else "Gaussian", "cl_m": list(w["cl"]), "sigma": w["sigma"],
              "seed": w["seed"], "kernel": kernel, "library": "libscarf3df" if flavour == "f" else "libscarf3d",
              "host_draw_s": _sig(host_s)}
    if tile:
        config["block_tile"] = dict(zip(("x", "j", "k"), tile))
    if structured:
        config["grid"] = list(w["n"])
    else:
        config["points"] = npts
        config["box_m"] = list(w["box"])
    if w["acf"] == 0:
        config["hurst"] = w["hurst"]
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Gpt·harm/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": _sig(main_r["ms_per_step"]), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (harmonics drawn on the host from the configured spectrum)",
        "config": config,
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": {k: {kk: (_sig(vv) if isinstance(vv, float) else vv) for kk, vv in r.items() if kk != "stats"}
                    for k, r in res.items()},
    }
    if args.workload != "precision_512":
        out["metric"] = f"SRM Gpt·harm/s, workload {args.workload}"
    if "direct_f32" in res and kernel != "direct_f32":
        out["speedup_vs_fp32_direct"] = _sig(res["direct_f32"]["field_ms"] / main_r["field_ms"])
    for k, r in out["kernels"].items():
        r.setdefault("gpt_harm_per_s", _sig(work / (r["field_ms"] * 1e-3) / 1e9))
        if "direct_f32" in res and k != "direct_f32":
            r["speedup_vs_fp32_direct"] = _sig(res["direct_f32"]["field_ms"] / r["field_ms"])
    out.update(extra)
    print(json.dumps(out), flush=True)
    lib.finalize()
    return 0


if __name__ == "__main__":
    sys.exit(main())
