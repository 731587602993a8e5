"""Shared helpers: drive the library and the oracle with the same resolved
configuration and compare."""
from __future__ import annotations

import numpy as np

from oracle_lib import structured_cfg  # noqa: F401  (re-exported)


def init_product(lib, cfg: dict) -> None:
    """Initialize the library so that its object holds exactly `cfg`
    (options follow the presence rules of src/c_binding.c:89-103)."""
    nd = cfg["nd"]
    opts = dict(method=1, hurst=cfg.get("hurst", 0.0), alpha=cfg.get("alpha", 0.0), beta=cfg.get("beta", 0.0),
                gamma=cfg.get("gamma", 0.0), rescale=cfg.get("rescale", 0))
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
        o = lib.opt(**opts)
        lib.cart_initialize(nd, cfg["fs"][:nd], cfg["fe"][:nd], cfg["ds"], cfg["acf"], cfg["cl"][:nd], cfg["sigma"],
                            o)
    else:
        if cfg["ds"] != cfg["dh"]:
            opts["ds"] = cfg["ds"]
        o = lib.opt(**opts)
        lib.nocart_initialize(nd, cfg["x"], cfg["y"], cfg["z"] if nd == 3 else None, cfg["dh"], cfg["acf"],
                              cfg["cl"][:nd], cfg["sigma"], o)
    lib.set_nharm(cfg["nharm"])


def unstructured_cfg(npts=2000, box=(10000.0, 10000.0, 5000.0), dh=50.0, acf=1, cl=(3000.0, 3000.0, 1000.0),
                     sigma=0.1, hurst=0.25, nharm=256, seed=7, real=np.float32, nd=3, **extra):
    """Point list uniform in a box (coordinates rounded to `real`), nc/fc
    defaulted to the coordinate min/max (src/scarflib.f90:264-271)."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, box[0], npts).astype(real).astype(np.float64)
    y = rng.uniform(0, box[1], npts).astype(real).astype(np.float64)
    z = rng.uniform(0, box[2], npts).astype(real).astype(np.float64) if nd == 3 else np.zeros(npts)
    if nd == 2:
        cl = (cl[0], cl[1], 0.0)
    r = real
    cfg = dict(nd=nd, structured=False, ds=float(r(dh)), dh=float(r(dh)), acf=acf,
               cl=tuple(float(r(c)) for c in cl), sigma=float(r(sigma)), hurst=float(r(hurst)),
               nc=(x.min(), y.min(), z.min() if nd == 3 else 0.0), fc=(x.max(), y.max(), z.max() if nd == 3 else 0.0),
               alpha=0.0, beta=0.0, gamma=0.0, nharm=nharm, x=x, y=y, z=z, mute=-999.0, taper=-999.0, rescale=0)
    cfg.update(extra)
    return cfg


def oracle_field(oracle, cfg: dict, seed: int, flavour: str):
    """FP64 field of the oracle on the same draw the library makes."""
    p = oracle.params(cfg)
    s = oracle.setup(p, flavour)
    harm, used, _ = oracle.draw(p, s, seed, flavour)
    if cfg.get("structured", True):
        f = oracle.field_structured(p, s, harm, used)
    else:
        f = oracle.field_unstructured(p, s, harm, used)
    return f, s
