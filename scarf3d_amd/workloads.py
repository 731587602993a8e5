"""Resolved SRM configurations and the benchmark workloads.

A configuration is a plain dict holding exactly what the library's object
keeps after initialize, with the reference defaults filled in
(src/scarflib.f90:132-158 structured, :243-271 point lists): nc/fc from the
grid or the coordinates, dh = ds, mute = taper = -999, rescale = 0, values
rounded to the library's `real`. `init_library` hands one to the C ABI
following the option presence rules of src/c_binding.c:89-103 / :116-130.

Nothing here computes a field; the same dicts drive the library, the bench
and the tests.
"""
from __future__ import annotations

import numpy as np


def structured_cfg(n=(16, 16, 16), ds=50.0, acf=0, cl=(2000.0, 2000.0, 2000.0), sigma=0.05, hurst=0.25,
                   nharm=256, dh=None, alpha=0.0, beta=0.0, gamma=0.0, fs=None, poi=None, mute=-999.0,
                   taper=-999.0, rescale=0, real=np.float32):
    """Structured grid of n points from global index fs (1-based), with the
    reference defaults: nc = (fs-1)*ds, fc = (fe-1)*ds, dh = ds
    (src/scarflib.f90:132-158)."""
    nd = 3 if cl[2] > 0 else 2
    fs = tuple(fs) if fs is not None else (1, 1, 1)
    fe = tuple(fs[i] + n[i] - 1 for i in range(3))
    r = real
    nc = tuple(float(r(r(fs[i] - 1) * r(ds))) for i in range(3))
    fc = tuple(float(r(r(fe[i] - 1) * r(ds))) for i in range(3))
    return dict(nd=nd, structured=True, fs=fs, fe=fe, ds=float(r(ds)), dh=float(r(dh if dh else ds)), acf=acf,
                cl=tuple(float(r(c)) for c in cl), sigma=float(r(sigma)), hurst=float(r(hurst)), nc=nc, fc=fc,
                alpha=float(r(alpha)), beta=float(r(beta)), gamma=float(r(gamma)), nharm=nharm, poi=poi,
                mute=float(r(mute)), taper=float(r(taper)), rescale=rescale)


def unstructured_cfg(npts=2000, box=(10000.0, 10000.0, 5000.0), dh=50.0, acf=1, cl=(3000.0, 3000.0, 1000.0),
                     sigma=0.1, hurst=0.25, nharm=256, seed=7, real=np.float32, nd=3, coord_real=None, **extra):
    """Point list uniform in a box (numpy PCG64 with `seed`), coordinates
    rounded to `coord_real` (default: the library's `real`), nc/fc defaulted
    to the coordinate min/max (src/scarflib.f90:264-271)."""
    cr = coord_real or real
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, box[0], npts).astype(cr).astype(np.float64)
    y = rng.uniform(0, box[1], npts).astype(cr).astype(np.float64)
    z = rng.uniform(0, box[2], npts).astype(cr).astype(np.float64) if nd == 3 else np.zeros(npts)
    if nd == 2:
        cl = (cl[0], cl[1], 0.0)
    r = real
    cfg = dict(nd=nd, structured=False, ds=float(r(dh)), dh=float(r(dh)), acf=acf,
               cl=tuple(float(r(c)) for c in cl), sigma=float(r(sigma)), hurst=float(r(hurst)),
               nc=(x.min(), y.min(), z.min() if nd == 3 else 0.0), fc=(x.max(), y.max(), z.max() if nd == 3 else 0.0),
               alpha=0.0, beta=0.0, gamma=0.0, nharm=nharm, x=x, y=y, z=z, mute=-999.0, taper=-999.0, rescale=0)
    cfg.update(extra)
    return cfg


def init_library(lib, cfg: dict) -> None:
    """Initialize `lib` (scarf3d_amd.Library) so that its object holds
    exactly `cfg`: options are passed only where the presence rules of
    src/c_binding.c:89-103 forward them (nc/fc of a structured grid only when
    they differ from the block's own extents). cfg["method"] selects the
    algorithm (default 1, SRM; 0 is FIM, with cfg["pad"])."""
    nd = cfg["nd"]
    opts = dict(method=cfg.get("method", 1), hurst=cfg.get("hurst", 0.0), alpha=cfg.get("alpha", 0.0),
                beta=cfg.get("beta", 0.0), gamma=cfg.get("gamma", 0.0), rescale=cfg.get("rescale", 0),
                pad=cfg.get("pad", 0))
    if cfg.get("mute", -999.0) > 0:
        opts["mute"] = cfg["mute"]
    if cfg.get("taper", -999.0) > 0:
        opts["taper"] = cfg["taper"]
    if cfg.get("poi") is not None:
        poi = np.asarray(cfg["poi"], dtype=lib.dtype)
        opts["poi"] = poi.reshape(-1, order="F")
        opts["npoi"] = poi.shape[1]
    if cfg.get("structured", True):
        if cfg["dh"] != cfg["ds"]:
            opts["dh"] = cfg["dh"]
        # model extents other than the block's own (a block of a larger grid,
        # as each MPI rank of the reference holds): passed as nc/fc
        r = lib.dtype
        own_nc = [float(r(r(cfg["fs"][i] - 1) * r(cfg["ds"]))) for i in range(nd)]
        own_fc = [float(r(r(cfg["fe"][i] - 1) * r(cfg["ds"]))) for i in range(nd)]
        if list(cfg["nc"][:nd]) != own_nc or list(cfg["fc"][:nd]) != own_fc:
            opts["nc"] = list(cfg["nc"][:nd])
            opts["fc"] = list(cfg["fc"][:nd])
        o = lib.opt(**opts)
        lib.cart_initialize(nd, cfg["fs"][:nd], cfg["fe"][:nd], cfg["ds"], cfg["acf"], cfg["cl"][:nd], cfg["sigma"],
                            o)
    else:
        if cfg["ds"] != cfg["dh"]:
            opts["ds"] = cfg["ds"]
        # model extents other than the points' own min/max (the default of
        # src/scarflib.f90:264-271): passed as nc/fc
        r = lib.dtype
        coords = [cfg["x"], cfg["y"], cfg["z"]][:nd]
        own_nc = [float(np.asarray(c, dtype=r).min()) for c in coords]
        own_fc = [float(np.asarray(c, dtype=r).max()) for c in coords]
        if list(cfg["nc"][:nd]) != own_nc or list(cfg["fc"][:nd]) != own_fc:
            opts["nc"] = list(cfg["nc"][:nd])
            opts["fc"] = list(cfg["fc"][:nd])
        o = lib.opt(**opts)
        lib.nocart_initialize(nd, cfg["x"], cfg["y"], cfg["z"] if nd == 3 else None, cfg["dh"], cfg["acf"],
                              cfg["cl"][:nd], cfg["sigma"], o)
    lib.set_nharm(cfg["nharm"])


# Benchmark workloads, one per measured configuration. flavour: library
# `real` (f = libscarf3df, d = libscarf3d); kernel: the timed field kernel;
# compare: kernels also timed on the same resident table.
WORKLOADS = {
    # 512^3 x 20k, the configuration the headline metric is quoted on (seed 1234)
    "precision_512": dict(n=(512, 512, 512), ds=20.0, acf=0, hurst=0.1, cl=(5000.0,) * 3, sigma=0.05, nharm=20000,
                          seed=1234, flavour="f", kernel="mfma_bf16x3",
                          compare=("direct_f32", "mfma_f16x2", "mfma_bf16x2")),
    # 128^3 x 5k with FP64 host semantics (double library); the reference's
    # CPU path on one rank is timed beside the GPU FP32 direct-sum kernel
    "plain_128": dict(n=(128, 128, 128), ds=50.0, acf=0, hurst=0.25, cl=(2000.0,) * 3, sigma=0.05, nharm=5000,
                      seed=1234, flavour="d", kernel="direct_f32", compare=("mfma_bf16x3", "mfma_f16x2", "direct_f64"),
                      cpu_one_rank=True),
    # 2,000,000 points uniform in 10 x 10 x 5 km (numpy PCG64 seed 7, FP64
    # coordinates through the double library), Gaussian cl = (3000, 3000,
    # 1000) m, dh = 50 m, 10k harmonics, the FP32 direct kernel
    "points_2m": dict(points=2_000_000, box=(10000.0, 10000.0, 5000.0), ds=50.0, acf=1, hurst=0.25,
                      cl=(3000.0, 3000.0, 1000.0), sigma=0.1, nharm=10000, seed=7, flavour="d", kernel="direct_f32",
                      compare=("direct_f64",)),
    # 1024 x 1024 x 512 x 50k, 16 POIs (numpy seed 16), mute 20 m, taper 500 m,
    # rescale to sigma (seed 1234)
    "large_1024": dict(n=(1024, 1024, 512), ds=20.0, acf=0, hurst=0.2, cl=(6000.0, 4000.0, 1500.0), sigma=0.08,
                       nharm=50000, seed=1234, rescale=1, npoi=16, mute=20.0, taper=500.0, flavour="f",
                       kernel="mfma_bf16x3", compare=("mfma_f16x2",)),
}


def workload_cfg(w: dict) -> dict:
    """The resolved configuration of a benchmark workload."""
    real = np.float32 if w["flavour"] == "f" else np.float64
    if "points" in w:
        return unstructured_cfg(npts=w["points"], box=w["box"], dh=w["ds"], acf=w["acf"], cl=w["cl"],
                                sigma=w["sigma"], hurst=w["hurst"], nharm=w["nharm"], seed=w["seed"], real=real)
    kw = {k: w[k] for k in ("n", "ds", "acf", "hurst", "cl", "sigma", "nharm") if k in w}
    extra = {}
    if w.get("npoi"):
        rng = np.random.default_rng(16)
        ext = np.array([(w["n"][i] - 1) * w["ds"] for i in range(3)])
        extra = dict(poi=(rng.uniform(0, 1, (3, w["npoi"])) * ext[:, None]), mute=w["mute"], taper=w["taper"])
    return structured_cfg(rescale=w.get("rescale", 0), real=real, **kw, **extra)


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
