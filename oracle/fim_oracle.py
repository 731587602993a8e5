"""FP64 restatement of the reference's Fourier integral method (FIM) on one
process: the checker of the library's FIM path.

TEST INFRASTRUCTURE ONLY. Nothing in the library imports, links or runs this
module; tests/ and the bench's CPU-baseline leg use it as the checker.

Each function cites the reference lines it follows (src/scarflib_fim.f90 with
DIP off, the default build; src/scarflib_common.f90; src/prng.cpp). What is
exact and what is FP64:
  - the internal grid (npts, off_axis), the spectrum-filter decision and the
    nearest-neighbour indices are discrete: they are computed at the library's
    `real` precision with the reference's order of operations (numpy float32
    or float64 scalars and arrays, no fused operations);
  - the uniforms are the yarn2 stream of the C oracle (liboracle.so,
    orc_uniform_stream) rounded to `real`, as the reference stores them;
  - amplitudes, phases, transforms, statistics, rescale and taper are FP64.
The seed -> uniform map rests on the yarn2 restatement shared with the SRM
oracle: it is as pinned as that one (TRNG4 is not available here).
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
PI = 3.141592653589793


def _nint(x) -> int:
    """Fortran NINT: nearest integer, halfway cases away from zero (exact for
    float32 and for float64 below 2^52)."""
    v = float(x)
    return int(math.copysign(math.floor(abs(v) + 0.5), v))


def _nint_array(x: np.ndarray) -> np.ndarray:
    v = x.astype(np.float64)
    return (np.sign(v) * np.floor(np.abs(v) + 0.5)).astype(np.int64)


def setup(cfg: dict, real=np.float32) -> dict:
    """Internal grid and spectrum constants (src/scarflib_fim.f90:754-849,
    compute_spectrum :2440-2482, scaling :951). cfg holds the object's values
    (nc, fc, ds, dh, cl, acf, sigma, hurst, pad, poi, nd)."""
    r = real
    dh, ds = r(cfg["dh"]), r(cfg["ds"])
    cl = [r(c) for c in cfg["cl"]]
    n = 3 if cl[2] > 0 else 2
    half = dh * r(0.5)
    npts, off = [0, 0, 0], [r(0)] * 3
    for i in range(3):
        lo, hi = r(cfg["nc"][i]), r(cfg["fc"][i])
        np_ = _nint((((hi + half) - lo) + half) / dh) + 1
        offset = _nint(cl[i] / dh) if cfg.get("pad", 0) == 1 else 0
        np_ += 2 * offset
        if np_ % 2:
            np_ += 1
        o = lo - (r(offset) + r(0.5)) * dh
        if n == 2 and i == 2:
            o, np_ = -o, 1
        npts[i], off[i] = np_, o
    info1 = int(any(npts[i] <= _nint(((r(2) * r(PI)) * cl[i]) / dh) for i in range(n)))
    info2 = int(any(dh > cl[i] / r(2) for i in range(n)))
    dk = [(r(2) * r(PI)) / (r(npts[i]) * dh) for i in range(3)]
    hurst, sigma = float(r(cfg["hurst"])), float(r(cfg["sigma"]))
    nu = hurst + n / 2.0
    prod = 1.0
    for i in range(n):
        prod *= float(cl[i])
    if cfg["acf"] == 0:
        cnst = 2.0 ** n * PI ** (n / 2.0) * math.gamma(nu) * sigma ** 2 * prod / math.gamma(hurst)
    else:
        cnst = PI ** (n / 2.0) * sigma ** 2 * prod
    pts = []
    poi = cfg.get("poi")
    if poi is not None:
        poi = np.asarray(poi, dtype=np.float64).reshape(cfg["nd"], -1, order="F")
        for q in range(poi.shape[1]):
            p3 = [r(0)] * 3
            for k in range(min(n, cfg["nd"])):
                p3[k] = r(poi[k, q]) - off[k]
            pts.append(p3)
    return dict(ndim=n, npts=npts, off_axis=off, dh=dh, ds=ds, dk=dk, kc=r(PI) / ds, filter=bool(ds != dh),
                cl=cl, acf=cfg["acf"], nu=nu, cnst=cnst, sigma=sigma, info1=info1, info2=info2,
                scaling=1.0 / math.sqrt(float(npts[0]) * npts[1] * npts[2] * float(dh) ** n), pt=pts, real=real)


def uniforms(seed: int, total: int, real=np.float32) -> np.ndarray:
    """Stream positions 0..total-1 of the seed (prng, src/prng.cpp:66-130 with
    one process: consecutive values in Fortran order), rounded to `real`."""
    lib = C.CDLL(str(_HERE / "liboracle.so"))
    lib.orc_uniform_stream.argtypes = [C.c_long, C.c_int, C.POINTER(C.c_double)]
    out = np.empty(total, dtype=np.float64)
    lib.orc_uniform_stream(seed, total, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out.astype(real)


def wavenumbers(npts: int, dk) -> np.ndarray:
    """kx = [0, dk, ..., (n/2) dk, (n/2 - 1) dk, ..., dk] at `real`
    (src/scarflib_fim.f90:2470-2472)."""
    q = np.arange(npts)
    m = np.where(q <= npts // 2, q, npts - q)
    return m.astype(type(dk)) * dk


def spectrum(s: dict, seed: int) -> np.ndarray:
    """Half spectrum S[k, j, i], i = 0..n1/2 (compute_spectrum,
    src/scarflib_fim.f90:2484-2530): uniform u of stream position
    i + n1 (j + n2 k), amplitude sqrt(const fn(|k cl|)), zero above kc when
    ds != dh, value amp e^(i 2 pi u)."""
    n1, n2, n3 = s["npts"]
    h = n1 // 2 + 1
    r = s["real"]
    u = uniforms(seed, n1 * n2 * n3, r).reshape(n3, n2, n1)[:, :, :h]
    kx = wavenumbers(n1, s["dk"][0])[:h]
    ky = wavenumbers(n2, s["dk"][1])
    kz = wavenumbers(n3, s["dk"][2])
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    kr = np.sqrt((KX * KX + KY * KY) + KZ * KZ)  # at `real`: the filter decision
    keep = ~(s["filter"] & (kr > s["kc"]))
    c = [float(v) for v in s["cl"]]
    x2 = (KX.astype(np.float64) * c[0]) ** 2 + (KY.astype(np.float64) * c[1]) ** 2 + (KZ.astype(np.float64) * c[2]) ** 2
    fn = (1.0 + x2) ** (-s["nu"]) if s["acf"] == 0 else np.exp(-0.25 * x2)
    amp = np.sqrt(s["cnst"] * fn) * keep
    ph = 2.0 * PI * u.astype(np.float64)
    return amp * np.exp(1j * ph)


def enforce_symmetry(S: np.ndarray, n1: int) -> np.ndarray:
    """Hermitian conditions on the planes kx = 0 and kx = n1/2 with one
    process holding the grid (enforce_symmetry, src/scarflib_fim.f90:2660-3008).
    Point conditions first; then every value below is the conjugate of the
    value after the point conditions at the mirrored (j, k): k > n3/2 all j,
    k = n3/2 with j > n2/2, and k = 0 with j > n2/2, which in 3D is conjugated
    in place (the one-process exchange overwrites that line's subarray types
    with those of the j = 0 column, :2866-2960)."""
    S = S.copy()
    n3, n2, _ = S.shape
    jn, kn = n2 // 2, n3 // 2
    for p in (0, n1 // 2):
        P = S[:, :, p]
        if p == 0:
            P[0, 0] = 0
        if p == n1 // 2:
            P[0, 0] = P[0, 0].real
        P[kn, jn] = P[kn, jn].real
        P[kn, 0] = P[kn, 0].real
        P[0, jn] = P[0, jn].real
        B = P.copy()
        for k in range(kn + 1, n3):
            for j in range(n2):
                P[k, j] = np.conj(B[n3 - k, (n2 - j) % n2])
        for j in range(jn + 1, n2):
            P[kn, j] = np.conj(B[kn, n2 - j])
        for j in range(jn + 1, n2):
            P[0, j] = np.conj(B[0, n2 - j]) if n3 == 1 else np.conj(B[0, j])
        S[:, :, p] = P
    return S


def transform(S: np.ndarray, n1: int) -> np.ndarray:
    """Unnormalized backward c2c along z and y, then c2r along x
    (transform_along_z/y/x, src/scarflib_fim.f90:1705-2303, FFTW_BACKWARD =
    e^(+i)); the c2r reads the real parts of the kx = 0 and Nyquist terms
    only. Returns the real internal field R[k, j, i]."""
    T = np.fft.ifft(S, axis=0, norm="forward") if S.shape[0] > 1 else S.copy()
    T = np.fft.ifft(T, axis=1, norm="forward")
    T[:, :, 0] = T[:, :, 0].real
    T[:, :, n1 // 2] = T[:, :, n1 // 2].real
    return np.fft.irfft(T, n=n1, axis=2, norm="forward")


def moments(R: np.ndarray) -> tuple[float, float]:
    """mean and variance (n - 1 divisor) of the internal field (mean_3d,
    variance_3d, src/scarflib_common.f90:694-817; one process: parallel_variance
    returns them unchanged)."""
    v = R.astype(np.float64).ravel()
    mu = float(v.mean())
    d = v - mu
    return mu, float((d @ d - d.sum() ** 2 / v.size) / (v.size - 1))


def taper(R: np.ndarray, s: dict, mute: float, taper_r: float) -> np.ndarray:
    """tapering_structured on the internal grid, positions (i - 1) dh, for
    every point of interest in order (src/scarflib_common.f90:881-941)."""
    n3, n2, n1 = R.shape
    dh = float(s["dh"])
    if not s["pt"]:
        return R
    Z, Y, X = np.meshgrid(np.arange(n3) * dh, np.arange(n2) * dh, np.arange(n1) * dh, indexing="ij")
    out = R.copy()
    for p in s["pt"]:
        d = np.sqrt((X - float(p[0])) ** 2 + (Y - float(p[1])) ** 2 + (Z - float(p[2])) ** 2)
        ramp = (d > mute) & (d <= mute + taper_r)
        t = (d - mute) / taper_r
        f = np.where(d <= mute, 0.0, np.where(ramp, 0.5 - 0.5 * np.cos(t * PI), 1.0))
        out = out * f
    return out


def nearest(R: np.ndarray, s: dict, x: np.ndarray, y: np.ndarray, z: np.ndarray) -> np.ndarray:
    """Values at points given in internal coordinates (`real` arrays relative
    to off_axis): c = (x / dh + 3) - 1 per axis at `real`, inside when
    1 <= c < npts + 1, value of the halo-extended field at NINT(c) - 1 (0 is
    the halo: zero along x and y, the last plane along the periodic z)
    (order_points_*, interpolate, exchange_halo, src/scarflib_fim.f90:1514-1704,
    :3013-3120). Outside points are 0."""
    r = s["real"]
    n1, n2, n3 = s["npts"]
    dh = r(s["dh"])
    cs = [(c.astype(r) / dh + r(3)) - r(1) for c in (x, y, z)]
    inside = np.ones(cs[0].shape, dtype=bool)
    idx = []
    for c, n in zip(cs, (n1, n2, n3)):
        inside &= (c >= r(1)) & (c < r(n + 1))
        idx.append(np.clip(_nint_array(c) - 1, 0, n))
    ix, iy, iz = idx
    iz = np.where(iz == 0, n3, iz)
    ok = inside & (ix > 0) & (iy > 0)
    vals = R[np.clip(iz - 1, 0, n3 - 1), np.clip(iy - 1, 0, n2 - 1), np.clip(ix - 1, 0, n1 - 1)]
    return np.where(ok, vals, 0.0)


def realization(cfg: dict, seed: int, real=np.float32) -> dict:
    """One FIM realization: the field at the caller's points (structured
    block in Fortran order flattened, or the point list), info(1:4) and the
    internal field (src/scarflib_fim.f90:614-1104 / :141-612)."""
    s = setup(cfg, real)
    n1 = s["npts"][0]
    S = enforce_symmetry(spectrum(s, seed), n1)
    R = transform(S, n1) * s["scaling"]
    mu, var = moments(R)
    info3, info4 = math.sqrt(var), mu
    if cfg.get("rescale", 0) == 1:
        R = R * (s["sigma"] / info3) - info4
        info3, info4 = s["sigma"], 0.0
    R = taper(R, s, float(real(cfg.get("mute", -999.0))), float(real(cfg.get("taper", -999.0))))
    r = real
    o = s["off_axis"]
    if cfg.get("structured", True):
        ds = r(cfg["ds"])
        fs, fe = cfg["fs"], cfg["fe"]
        ax = [(np.arange(fs[d] - 1, fe[d]).astype(r) * ds - o[d]) for d in range(3)]
        Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        field = nearest(R, s, X.ravel(), Y.ravel(), Z.ravel())
    else:
        x = np.asarray(cfg["x"], dtype=r) - o[0]
        y = np.asarray(cfg["y"], dtype=r) - o[1]
        z = (np.asarray(cfg["z"], dtype=r) if cfg["nd"] == 3 else np.zeros(len(x), dtype=r)) - o[2]
        field = nearest(R, s, x, y, z)
    return dict(field=field, info=[s["info1"], s["info2"], info3, info4], internal=R, setup=s, mean=mu, var=var)
