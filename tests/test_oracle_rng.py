"""Pins the oracle's harmonic draw.

yarn2 (TRNG4, not in the reference tree) is restated from its published
design; same-seed bit parity with a TRNG build stays unpinned. What the
design fixes is checked here: the output generator is a primitive root of
m = 2^31 - 1, the characteristic polynomial of the order-2 recurrence is
primitive over GF(m) (full period m^2 - 1), and seeds reduce modulo m as
trng::yarn2::seed does. The sampler (src/scarflib_srm.f90:731-778) is pinned
by its distributions: the accepted radial wavenumbers follow the normalised
srm_fn on [kmin, kmax], directions are uniform on the sphere (disc in 2D),
and a, b are standard normal.
"""
import numpy as np
import pytest
from scipy import integrate, stats

from oracle_lib import structured_cfg

M = 2**31 - 1
A0, A1 = 1498809829, 1160990996
G = 123567893
PRIMES_M_MINUS_1 = [2, 3, 7, 11, 31, 151, 331]  # 2^31 - 2 = 2 * 3^2 * 7 * 11 * 31 * 151 * 331


def test_factorisation_of_m_minus_one():
    n = 1
    for p, e in zip(PRIMES_M_MINUS_1, [1, 2, 1, 1, 1, 1, 1]):
        n *= p**e
    assert n == M - 1


def test_generator_is_primitive_root(oracle):
    for q in PRIMES_M_MINUS_1:
        assert oracle.L.orc_powmod(G, (M - 1) // q, M) != 1
    assert oracle.L.orc_powmod(G, M - 1, M) == 1


def _polymulmod(a, b):
    """(a0 + a1 x)(b0 + b1 x) mod (x^2 - A0 x - A1) over GF(M)."""
    c0 = a[0] * b[0]
    c1 = a[0] * b[1] + a[1] * b[0]
    c2 = a[1] * b[1]
    return ((c0 + c2 * A1) % M, (c1 + c2 * A0) % M)


def _xpow(e):
    r, b = (1, 0), (0, 1)
    while e:
        if e & 1:
            r = _polymulmod(r, b)
        b = _polymulmod(b, b)
        e >>= 1
    return r


def test_recurrence_has_full_period():
    """x has order m^2 - 1 modulo the characteristic polynomial: the MRG with
    these coefficients has period m^2 - 1 (L'Ecuyer's full-period choice)."""
    order = M * M - 1
    assert _xpow(order) == (1, 0)
    for q in sorted(set(PRIMES_M_MINUS_1)):  # primes of (m-1)(m+1); m+1 = 2^31 adds only 2
        assert _xpow(order // q) != (1, 0), q


def test_stream_matches_recurrence_and_output_map(oracle):
    """First values of the stream from the recurrence and g^r mod m written out."""
    seed = 1234
    r0, r1 = seed % M, 1
    expect = []
    for _ in range(50):
        r0, r1 = (A0 * r0 + A1 * r1) % M, r0
        expect.append(0 if r0 == 0 else pow(G, r0, M))
    assert list(oracle.yarn2(seed, 50)) == expect
    u = oracle.uniforms(seed, 50)
    assert np.array_equal(u, np.array(expect, dtype=np.float64) / M)


def test_seed_reduction(oracle):
    assert np.array_equal(oracle.yarn2(5, 20), oracle.yarn2(5 + M, 20))
    # set_seed(int) -> unsigned long -> int64 reduction: -1 behaves as m - 1
    assert np.array_equal(oracle.yarn2(-1, 20), oracle.yarn2(M - 1, 20))


@pytest.mark.parametrize("seed", [1, 1234, 99999])
def test_uniforms_are_uniform(oracle, seed):
    u = oracle.uniforms(seed, 200000)
    assert u.min() > 0.0 and u.max() < 1.0
    assert stats.kstest(u, "uniform").pvalue > 1e-4
    lag = np.corrcoef(u[:-1], u[1:])[0, 1]
    assert abs(lag) < 0.01


def _draw(oracle, flavour, **kw):
    cfg = structured_cfg(**kw)
    p = oracle.params(cfg)
    s = oracle.setup(p, flavour)
    harm, used, n = oracle.draw(p, s, 2024, flavour)
    return cfg, s, harm, used


@pytest.mark.parametrize("flavour", ["f", "d"])
@pytest.mark.parametrize("acf,hurst,cl", [(0, 0.25, 2000.0), (1, 0.0, 300.0)])
def test_radial_wavenumbers_follow_spectrum(oracle, flavour, acf, hurst, cl):
    cfg, s, harm, used = _draw(oracle, flavour, n=(64, 64, 64), acf=acf, hurst=hurst, cl=(cl, cl, cl),
                               nharm=6000, real=np.float32 if flavour == "f" else np.float64)
    u = np.sqrt((harm[used == 1, :3] ** 2).sum(1)) * cl  # |v| * cl = u (isotropic cl)
    nu = hurst + 1.5
    psd = (lambda k: k * k / (1 + k * k) ** nu) if acf == 0 else (lambda k: k * k * np.exp(-k * k / 4))
    total = integrate.quad(psd, s.kmin, s.kmax, limit=200)[0]
    grid = np.linspace(s.kmin, s.kmax, 4001)
    cdf_vals = np.concatenate([[0.0], np.cumsum([integrate.quad(psd, a, b)[0] for a, b in zip(grid[:-1], grid[1:])])])
    cdf = lambda x: np.interp(x, grid, cdf_vals / total)  # noqa: E731
    assert u.min() >= s.kmin * (1 - 1e-6) and u.max() <= s.kmax * (1 + 1e-6)
    assert stats.kstest(u, cdf).pvalue > 1e-4


def test_directions_and_amplitudes(oracle):
    cfg, s, harm, used = _draw(oracle, "d", n=(64, 64, 64), nharm=8000, real=np.float64)
    h = harm[used == 1]
    d = h[:, :3] / np.linalg.norm(h[:, :3], axis=1, keepdims=True)
    assert np.all(np.abs(d.mean(0)) < 0.03)
    assert np.allclose((d**2).mean(0), 1 / 3, atol=0.02)
    # z-component uniform on [-1, 1] (Archimedes) and azimuth uniform
    assert stats.kstest(d[:, 2], "uniform", args=(-1, 2)).pvalue > 1e-4
    phi = np.arctan2(d[:, 0], d[:, 1])  # v1 ~ sin(phi), v2 ~ cos(phi)
    assert stats.kstest((phi + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-4
    for col in (3, 4):
        assert stats.kstest(h[:, col], "norm").pvalue > 1e-4


def test_directions_2d(oracle):
    cfg, s, harm, used = _draw(oracle, "d", n=(128, 128, 1), cl=(2000.0, 2000.0, 0.0), nharm=6000, real=np.float64)
    h = harm[used == 1]
    assert np.all(h[:, 2] == 0.0)
    phi = np.arctan2(h[:, 1], h[:, 0])  # v1 ~ cos(phi), v2 ~ sin(phi)
    assert stats.kstest((phi + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-4


def test_dropped_harmonics_are_those_above_kc(oracle):
    """ds > dh: u in (kc, kmax] are drawn, consume their direction uniforms,
    and are dropped (src/scarflib_srm.f90:748-751)."""
    cfg, s, harm, used = _draw(oracle, "d", n=(20, 20, 20), hurst=0.1, ds=20.0, dh=10.0, nharm=3000,
                               real=np.float64)
    assert s.kc < s.kmax
    assert 0 < used.sum() < len(used)
    u = np.sqrt((harm[used == 1, :3] ** 2).sum(1)) * 2000.0
    assert u.max() <= s.kc * (1 + 1e-12)
    assert np.all(harm[used == 0] == 0.0)
