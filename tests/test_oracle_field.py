"""Pins the oracle's FP64 field, statistics and taper by properties the
mathematics fixes, so that a dropped term, a wrong sign or a transposed index
in src/scarflib_srm.f90:794-851 / scarflib_common.f90:652-994 as restated
would fail:

- the rotation matrix is orthonormal, alpha = beta = gamma = 0 gives M = I
  and bar = 0, and alpha = 90 degrees maps x to y;
- translating the block (fs) shifts the field (global-index semantics);
- the ensemble variance at a point equals scaling^2 * sum(a^2 + b^2) / 2
  averaged, i.e. sigma^2 * intg(kmin,kmax)/intg() for full harmonics;
- mean/variance equal numpy on random data, and the Chan merge of split
  sets equals the statistics of the whole;
- the taper leaves points beyond mute+taper unchanged, zeroes points
  within mute, and rises from 0 to 1 across the taper ring;
- one harmonic is a closed-form sinusoid: a pure sine for (a, b) = (1, 0)
  and a pure cosine for (0, 1) at zero angles, and at alpha = 90 degrees a
  wave along y whose phase uses the hand-computed bar;
- the sampled rescale + taper used at full size equals the oracle pipeline.
"""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import _DP, structured_cfg


def _setup(oracle, flavour="d", **kw):
    cfg = structured_cfg(real=np.float64 if flavour == "d" else np.float32, **kw)
    p = oracle.params(cfg)
    return cfg, p, oracle.setup(p, flavour)


def test_rotation_identity_and_orthonormal(oracle):
    _, _, s = _setup(oracle)
    assert np.array_equal(np.array(s.matrix[:]), np.eye(3).reshape(-1, order="F"))
    assert np.array_equal(np.array(s.bar[:]), np.zeros(3))
    _, _, s = _setup(oracle, alpha=30.0, beta=-10.0, gamma=5.0)
    m = np.array(s.matrix[:]).reshape(3, 3, order="F")
    assert np.allclose(m @ m.T, np.eye(3), atol=1e-14)
    assert np.linalg.det(m) == pytest.approx(1.0, abs=1e-14)
    # bar = span/2 - M span/2
    span = np.array([15 * 50.0] * 3)
    assert np.allclose(np.array(s.bar[:]), span / 2 - m @ (span / 2), atol=1e-9)


def test_alpha_ninety_maps_x_to_y(oracle):
    _, _, s = _setup(oracle, alpha=90.0)
    m = np.array(s.matrix[:]).reshape(3, 3, order="F")
    assert np.allclose(m @ np.array([1.0, 0, 0]), [0, 1, 0], atol=1e-15)
    assert np.allclose(m @ np.array([0, 1.0, 0]), [-1, 0, 0], atol=1e-15)


def test_block_offset_shifts_the_field(oracle):
    """field(fs = (3, 2, 4))[i, j, k] == field(fs = (1, 1, 1))[i+2, j+1, k+3]
    when nc/fc describe the same model box."""
    cfg1, p1, s1 = _setup(oracle, n=(12, 10, 9), nharm=64, alpha=20.0)
    cfg2 = dict(cfg1, fs=(3, 2, 4), fe=(12, 10, 9))  # sub-block of the same model
    p2 = oracle.params(cfg2)
    s2 = oracle.setup(p2, "d")
    assert s1.scaling == s2.scaling and s1.kmin == s2.kmin
    harm, used, _ = oracle.draw(p1, s1, 3, "d")
    f1 = oracle.field_structured(p1, s1, harm, used)
    f2 = oracle.field_structured(p2, s2, harm, used)
    assert np.allclose(f2, f1[2:, 1:, 3:], rtol=0, atol=1e-15)


def _grid_coords(p):
    """global coordinates (i + fs - 2) * ds of the block, 0-based i (src/scarflib_srm.f90:796-800)"""
    ax = [(np.arange(p.fe[a] - p.fs[a] + 1) + p.fs[a] - 1) * p.ds for a in range(3)]
    return np.meshgrid(*ax, indexing="ij")


@pytest.mark.parametrize("ab,wave", [((1.0, 0.0), np.sin), ((0.0, 1.0), np.cos), ((-0.6, 0.8), None)])
def test_single_harmonic_closed_form(oracle, ab, wave):
    """One harmonic at zero angles (M = I, bar = 0) is the sinusoid
    scaling * (a sin(v.p) + b cos(v.p)) of src/scarflib_srm.f90:802-805: a
    pure sine for (a, b) = (1, 0), a pure cosine for (0, 1). A swapped a/b,
    a sign error in the phase or a transposed axis fails here."""
    cfg, p, s = _setup(oracle, n=(9, 7, 5), fs=(2, 3, 1), nharm=1)
    v = np.array([1.3e-3, -0.7e-3, 2.1e-3])  # rad/m
    harm = np.array([[*v, *ab]])
    f = oracle.field_structured(p, s, harm, np.ones(1, np.uint8))
    x, y, z = _grid_coords(p)
    t = v[0] * x + v[1] * y + v[2] * z
    expect = s.scaling * (wave(t) if wave else ab[0] * np.sin(t) + ab[1] * np.cos(t))
    assert np.allclose(f, expect, rtol=0, atol=1e-15 * s.scaling)
    if wave is np.sin:  # f(0) = 0 only for the sine: at the origin of the model
        cfg0, p0, s0 = _setup(oracle, n=(3, 3, 3), nharm=1)
        assert oracle.field_structured(p0, s0, harm, np.ones(1, np.uint8))[0, 0, 0] == 0.0


def test_single_harmonic_alpha_ninety_by_hand(oracle):
    """alpha = 90 degrees: M p = (-y, x, z) and bar = h - M h = (hx + hy,
    hy - hx, 0) with h the half-span (src/scarflib_srm.f90:677-700), so a
    harmonic v = (v1, 0, 0) has the phase v1 (hx + hy - y): the field does not
    depend on x, and b = 1 gives scaling * cos(v1 (hx + hy - y))."""
    cfg, p, s = _setup(oracle, n=(6, 8, 4), alpha=90.0, nharm=1)
    h = (np.array(cfg["fc"]) - np.array(cfg["nc"])) / 2
    assert np.allclose(np.array(s.bar[:]), [h[0] + h[1], h[1] - h[0], 0.0], rtol=0, atol=1e-12)
    v1 = 2.3e-3
    f = oracle.field_structured(p, s, np.array([[v1, 0.0, 0.0, 0.0, 1.0]]), np.ones(1, np.uint8))
    _, y, _ = _grid_coords(p)
    assert np.allclose(f, s.scaling * np.cos(v1 * (h[0] + h[1] - y)), rtol=0, atol=1e-14 * s.scaling)
    assert np.allclose(f, f[:1, :, :], rtol=0, atol=1e-14 * s.scaling)  # constant along x


def test_ensemble_variance_at_a_point(oracle):
    """Over seeds, f(p) = scaling * sum(a sin + b cos) has variance
    scaling^2 * N_used (a, b iid N(0,1))."""
    # Gaussian spectrum with the band around its peak: high acceptance, fast draws
    cfg, p, s = _setup(oracle, n=(8, 8, 8), acf=1, cl=(100.0, 100.0, 100.0), nharm=64)
    vals, nused = [], []
    for seed in range(1, 601):
        harm, used, _ = oracle.draw(p, s, seed, "d")
        vals.append(oracle.field_at(p, s, harm, used, [1], [2], [3])[0])
        nused.append(int(used.sum()))
    var = np.var(vals, ddof=1)
    expect = s.scaling**2 * np.mean(nused)
    assert var == pytest.approx(expect, rel=0.15)
    # the normalisation: scaling^2 * nharm = sigma^2 * intg(kmin,kmax)/intg()
    assert s.scaling**2 * p.nharm == pytest.approx(cfg["sigma"] ** 2 * s.intg_band / s.intg_all, rel=1e-12)


def test_field_is_sum_over_harmonics(oracle):
    """Linearity: the field of a table is the sum of the fields of its halves."""
    cfg, p, s = _setup(oracle, n=(8, 6, 5), nharm=40, alpha=10.0)
    harm, used, _ = oracle.draw(p, s, 8, "d")
    f = oracle.field_structured(p, s, harm, used)
    u1, u2 = used.copy(), used.copy()
    u1[20:] = 0
    u2[:20] = 0
    assert np.allclose(f, oracle.field_structured(p, s, harm, u1) + oracle.field_structured(p, s, harm, u2),
                       atol=1e-14)
    # without rotation, a single harmonic with v along x is periodic in x
    # with period 2 pi / v1 and constant along y and z
    cfg0, p0, s0 = _setup(oracle, n=(8, 6, 5), nharm=40)
    h1 = np.zeros((40, 5))
    h1[0] = [2 * np.pi / (4 * 50.0), 0.0, 0.0, 0.7, -0.3]
    one = np.zeros(40, dtype=np.uint8)
    one[0] = 1
    g = oracle.field_structured(p0, s0, h1, one)
    assert np.allclose(g[4:, :, :], g[:-4, :, :], atol=1e-14)
    assert np.allclose(g[:, 1:, :], g[:, :-1, :], atol=1e-14)


@pytest.mark.parametrize("flavour,dtype", [("f", np.float32), ("d", np.float64)])
def test_mean_variance_match_numpy(oracle, flavour, dtype):
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(100003) * 0.05 + 0.01).astype(dtype)
    ptr = x.ctypes.data_as(C.c_void_p)
    mean = getattr(oracle.L, f"orc_mean_{flavour}")(ptr, x.size)
    var = getattr(oracle.L, f"orc_variance_{flavour}")(ptr, x.size)
    x64 = x.astype(np.float64)
    rel = 1e-6 if flavour == "f" else 1e-12
    assert mean == pytest.approx(x64.mean(), rel=rel)
    assert var == pytest.approx(x64.var(ddof=1), rel=rel)


def test_chan_merge_equals_whole(oracle):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(10000) * 3 + 1
    parts = np.split(x, [1000, 3500, 3600, 9000])
    var = np.array([q.var(ddof=1) for q in parts])
    avg = np.array([q.mean() for q in parts])
    n = np.array([len(q) for q in parts], dtype=np.int64)
    v, m = C.c_double(), C.c_double()
    oracle.L.orc_parallel_variance(var.ctypes.data_as(_DP), avg.ctypes.data_as(_DP),
                                   n.ctypes.data_as(C.POINTER(C.c_long)), len(parts), C.byref(v), C.byref(m))
    assert v.value == pytest.approx(x.var(ddof=1), rel=1e-12)
    assert m.value == pytest.approx(x.mean(), rel=1e-12)


@pytest.mark.parametrize("flavour,dtype", [("f", np.float32), ("d", np.float64)])
def test_taper_profile(oracle, flavour, dtype):
    nx = ny = nz = 21
    fs = np.array([1, 1, 1], dtype=np.int32)
    fe = np.array([nx, ny, nz], dtype=np.int32)
    ds, mute, taper = 10.0, 30.0, 50.0
    poi = np.array([100.0, 100.0, 100.0])
    f = np.ones(nx * ny * nz, dtype=dtype)
    getattr(oracle.L, f"orc_taper_structured_{flavour}")(
        f.ctypes.data_as(C.c_void_p), fs.ctypes.data_as(C.POINTER(C.c_int)), fe.ctypes.data_as(C.POINTER(C.c_int)),
        ds, poi.ctypes.data_as(_DP), mute, taper)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    d = np.sqrt(((i * ds - 100) ** 2 + (j * ds - 100) ** 2 + (k * ds - 100) ** 2)).reshape(-1, order="F")
    assert np.all(f[d <= mute - 1e-9] == 0)
    assert np.all(f[d > mute + taper + 1e-9] == 1)
    ring = (d > mute + 1e-9) & (d <= mute + taper - 1e-9)
    assert np.all((f[ring] > 0) & (f[ring] < 1))
    # monotone in distance across the ring
    order = np.argsort(d[ring])
    assert np.all(np.diff(f[ring][order]) >= -1e-6)


def test_sampled_epilogue_equals_oracle_pipeline(oracle):
    """epilogue_at (the rescale and the taper of every POI, in FP64 at sampled
    points, used where the whole-field oracle would take hours) equals the
    oracle's own pipeline (orc_srm: statistics, rescale, taper) on a small
    grid, at every point."""
    from oracle_lib import epilogue_at

    poi = np.array([[300.0, 400.0, 200.0], [900.0, 500.0, 600.0], [100.0, 1000.0, 50.0]]).T
    cfg = structured_cfg(n=(24, 22, 14), nharm=128, poi=poi, mute=60.0, taper=300.0, rescale=1, real=np.float64)
    p = oracle.params(cfg)
    s = oracle.setup(p, "d")
    harm, used, _ = oracle.draw(p, s, 5, "d")
    f64 = oracle.field_structured(p, s, harm, used).reshape(-1, order="F")
    mean = f64.mean()
    var = ((f64 - mean) ** 2).sum() / (f64.size - 1)
    flat = np.arange(f64.size)
    mine = epilogue_at(f64, flat, cfg, mean, var)
    ref, _ = oracle.srm(p, 5, "d")
    assert np.count_nonzero(ref == 0.0) > 0  # some points muted
    assert np.array_equal(mine == 0.0, ref == 0.0)
    assert np.abs(mine - ref).max() <= 1e-12 * cfg["sigma"]
