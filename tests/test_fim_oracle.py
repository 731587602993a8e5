"""FIM oracle (oracle/fim_oracle.py) pinned without a GPU.

- the spectrum kernel's walk of the random stream (fim_stream::walk_item,
  replayed on the host) against the sequential yarn2 stream;
- the internal grid against values worked out by hand from
  src/scarflib_fim.f90:805-835;
- the Hermitian conditions against a line-by-line transcription of the
  one-process exchange of enforce_symmetry (src/scarflib_fim.f90:2730-3006),
  and against the property they exist for (a real plane transform);
- the transform convention against a single-mode closed form;
- the nearest-neighbour rule against the geometry (ties upward, halo rules);
- the spectrum constants against the variance they are built to give.
"""
from __future__ import annotations

import math
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import fim_oracle as fo  # noqa: E402

from oracle_lib import build_oracle  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _oracle_built():
    build_oracle()


def test_kernel_stream_walk_matches_sequential_stream(tmp_path):
    exe = tmp_path / "stream_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I",
                    str(ROOT / "scarf3d_amd" / "csrc"), str(ROOT / "tests" / "fim" / "stream_check.cpp"), "-o",
                    str(exe)], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
    assert "0 mismatches" in res.stdout


def _cfg(**kw):
    c = dict(nd=3, nc=(0.0, 0.0, 0.0), fc=(150.0, 150.0, 150.0), ds=10.0, dh=10.0, cl=(40.0, 40.0, 40.0), acf=1,
             sigma=0.1, hurst=0.3, pad=0, poi=None)
    c.update(kw)
    return c


def test_internal_grid_by_hand():
    # no padding: NINT((150 + 5 - 0 + 5) / 10) + 1 = 17 -> 18 (even); off_axis = 0 - 0.5 * 10
    s = fo.setup(_cfg())
    assert s["npts"] == [18, 18, 18]
    assert [float(v) for v in s["off_axis"]] == [-5.0, -5.0, -5.0]
    # padding: offset NINT(40 / 10) = 4 -> 17 + 8 = 25 -> 26; off_axis = -(4.5) * 10
    s = fo.setup(_cfg(pad=1))
    assert s["npts"] == [26, 26, 26]
    assert [float(v) for v in s["off_axis"]] == [-45.0, -45.0, -45.0]
    # anisotropic extents, 2D: z gets one point and the negated offset
    s = fo.setup(_cfg(nc=(100.0, -40.0, 0.0), fc=(390.0, 200.0, 0.0), cl=(40.0, 25.0, 0.0)))
    assert s["ndim"] == 2
    assert s["npts"] == [31 + 1, 25 + 1, 1]  # NINT(300/10)+1 = 31 -> 32; NINT(250/10)+1 = 26
    assert [float(v) for v in s["off_axis"]] == [95.0, -45.0, 5.0]
    # info(1): npts <= NINT(2 pi cl / dh); info(2): dh > cl / 2
    assert fo.setup(_cfg())["info1"] == 1  # 18 <= NINT(25.13)
    assert fo.setup(_cfg(cl=(20.0, 20.0, 20.0)))["info1"] == 0  # 18 > NINT(12.57)
    assert fo.setup(_cfg(cl=(15.0, 40.0, 40.0)))["info2"] == 1


def _reference_exchange(P0: np.ndarray, is_first: bool, is_nyq: bool) -> np.ndarray:
    """exchange_conjg of src/scarflib_fim.f90:2760-3006 for one process
    (fs = 1, fe = npts, gs = ge = own block), statement by statement, with
    1-based Fortran indices on a padded array. The second mpi_alltoallw sends
    one subarray type per process: when both the beta and the delta blocks
    apply to the same process, the delta block's types are the ones set last."""
    n2, n3 = P0.shape  # P0[j-1, k-1]
    ny2, ny3 = n2 // 2 + 1, n3 // 2 + 1
    spec = np.zeros((n2 + 2, n3 + 2), dtype=complex)
    spec[1:n2 + 1, 1:n3 + 1] = P0
    if is_first:
        spec[1, 1] = 0
    if is_nyq:
        spec[1, 1] = spec[1, 1].real
    spec[ny2, ny3] = spec[ny2, ny3].real
    spec[1, ny3] = spec[1, ny3].real
    spec[ny2, 1] = spec[ny2, 1].real
    buffer = spec.copy()
    sendbuf = np.zeros_like(buffer)

    def region(js, je, ks, ke):
        j0, j1, k0, k1 = max(1, js), min(n2, je), max(1, ks), min(n3, ke)
        if k0 <= k1 and j0 <= j1:
            cy, cz = j1 + j0, k1 + k0
            for k in range(k0, k1 + 1):
                for j in range(j0, j1 + 1):
                    sendbuf[cy - j, cz - k] = buffer[j, k]
            return (j0, j1, k0, k1)
        return None

    # first exchange: the whole region but beta and delta
    r = region(n2 - n2 + 2, n2 - 1 + 2, n3 - n3 + 2, n3 - 1 + 2)
    if r:
        j0, j1, k0, k1 = r
        buffer[j0:j1 + 1, k0:k1 + 1] = sendbuf[j0:j1 + 1, k0:k1 + 1]
    # second exchange: beta (gs(3) = 1), then delta (gs(2) = 1) for the same process
    t = None
    rb = region(n2 - n2 + 2, n2 - 1 + 2, 1, 1)
    if rb:
        t = rb
    rd = region(1, 1, n3 - n3 + 2, n3 - 1 + 2)
    if rd:
        t = rd
    if t:
        j0, j1, k0, k1 = t
        buffer[j0:j1 + 1, k0:k1 + 1] = sendbuf[j0:j1 + 1, k0:k1 + 1]
    for k in range(max(ny3 + 1, 1), n3 + 1):
        for j in range(1, n2 + 1):
            spec[j, k] = np.conj(buffer[j, k])
    for k in range(1, 2):
        for j in range(ny2 + 1, n2 + 1):
            spec[j, k] = np.conj(buffer[j, k])
    for k in range(max(ny3, 1), min(ny3, n3) + 1):
        for j in range(ny2 + 1, n2 + 1):
            spec[j, k] = np.conj(buffer[j, k])
    return spec[1:n2 + 1, 1:n3 + 1]


@pytest.mark.parametrize("shape", [(6, 8, 10), (4, 12, 1), (8, 2, 2), (10, 6, 4), (2, 4, 6)])
def test_symmetry_matches_reference_exchange(shape):
    n1, n2, n3 = shape
    rng = np.random.default_rng(n1 * 100 + n2 * 10 + n3)
    h = n1 // 2 + 1
    S = rng.normal(size=(n3, n2, h)) + 1j * rng.normal(size=(n3, n2, h))
    out = fo.enforce_symmetry(S, n1)
    for p in (0, n1 // 2):
        want = _reference_exchange(S[:, :, p].T, p == 0, p == n1 // 2).T
        np.testing.assert_array_equal(out[:, :, p], want)
    # other planes untouched
    mid = [i for i in range(h) if i not in (0, n1 // 2)]
    np.testing.assert_array_equal(out[:, :, mid], S[:, :, mid])


def test_symmetry_makes_planes_real():
    """2D: both planes exactly Hermitian, so their inverse transforms are real.
    3D: Hermitian except the k = 0, j > n2/2 line (conjugated in place with
    one process); fixing that line alone makes the plane transforms real."""
    rng = np.random.default_rng(5)
    n1, n2 = 8, 12
    S = rng.normal(size=(1, n2, 5)) + 1j * rng.normal(size=(1, n2, 5))
    out = fo.enforce_symmetry(S, n1)
    for p in (0, n1 // 2):
        t = np.fft.ifft(out[0, :, p])
        assert np.abs(t.imag).max() < 1e-14
    n3 = 10
    S = rng.normal(size=(n3, n2, 5)) + 1j * rng.normal(size=(n3, n2, 5))
    out = fo.enforce_symmetry(S, n1)
    for p in (0, n1 // 2):
        P = out[:, :, p].copy()
        t = np.fft.ifft2(P)
        assert np.abs(t.imag).max() > 1e-3  # the in-place line breaks the symmetry
        for j in range(n2 // 2 + 1, n2):
            P[0, j] = np.conj(P[0, n2 - j])
        assert np.abs(np.fft.ifft2(P).imag).max() < 1e-14


def test_transform_single_mode():
    """One coefficient a at (kx, ky, kz) = (1, 2, 0) of an n1 x n2 x n3 grid:
    f(x, y, z) = 2 Re(a e^(+2 pi i (x / n1 + 2 y / n2))) (FFTW_BACKWARD sign,
    unnormalized, the c2r doubling every kx strictly inside (0, n1/2))."""
    n1, n2, n3 = 8, 6, 4
    S = np.zeros((n3, n2, n1 // 2 + 1), dtype=complex)
    a = 0.3 - 0.7j
    S[0, 2, 1] = a
    R = fo.transform(S, n1)
    z, y, x = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    want = 2 * (a * np.exp(2j * np.pi * (x / n1 + 2 * y / n2))).real
    np.testing.assert_allclose(R, want, atol=1e-14)
    # the kx = 0 and Nyquist terms count once, and after the (y, z) transforms
    # only their real parts reach the field
    S[:] = 0
    S[0, 0, n1 // 2] = 0.5 + 9j
    S[1, 0, 0] = 0.25 + 3j
    R = fo.transform(S, n1)
    want = 0.5 * (-1.0) ** x + ((0.25 + 3j) * np.exp(2j * np.pi * z / n3)).real
    np.testing.assert_allclose(R, want, atol=1e-14)


def test_nearest_rule():
    """External points on the dh = ds grid sit halfway between internal nodes
    (internal 1-based point q at off_axis + (q - 1) dh, off_axis = nc - dh/2):
    the reference's NINT takes the upper node; the z halo is periodic."""
    s = fo.setup(_cfg())
    n1, n2, n3 = s["npts"]
    R = np.arange(n1 * n2 * n3, dtype=np.float64).reshape(n3, n2, n1)
    r = np.float32
    xs = np.arange(16).astype(r) * r(10) - s["off_axis"][0]
    v = fo.nearest(R, s, xs, np.full(16, r(5)), np.full(16, r(5)))
    # x = 10 i + 5 -> c = i + 2.5 -> NINT = i + 3 -> internal 0-based i + 1; y = z = 5 -> 0-based 1
    np.testing.assert_array_equal(v, R[1, 1, 1:17])
    # below the grid along x: c < 1 -> outside -> 0; c in [1, 1.5) -> the halo -> 0
    v = fo.nearest(R, s, np.array([-25.0, -8.0], dtype=r), np.full(2, r(5)), np.full(2, r(5)))
    np.testing.assert_array_equal(v, [0.0, 0.0])
    # z in [1, 1.5): NINT = 1 -> halo plane 0 = the last plane (periodic z)
    v = fo.nearest(R, s, np.array([5.0], dtype=r), np.array([5.0], dtype=r), np.array([-8.0], dtype=r))
    np.testing.assert_array_equal(v, [R[n3 - 1, 1, 1]])


def test_field_variance_matches_sigma():
    """Gaussian spectrum with cl far below the box: the FIM field's variance is
    sigma^2 up to the sampling of a few thousand modes (the constant, the
    1/sqrt(N dh^n) scaling and the Hermitian doubling together)."""
    cfg = _cfg(nc=(0.0, 0.0, 0.0), fc=(630.0, 630.0, 630.0), cl=(40.0, 40.0, 40.0), sigma=0.2, structured=True,
               fs=(1, 1, 1), fe=(64, 64, 64))
    out = fo.realization(cfg, seed=11, real=np.float64)
    assert abs(math.sqrt(out["var"]) / 0.2 - 1.0) < 0.05
    # von Karman, 2D
    cfg = _cfg(nc=(0.0, 0.0, 0.0), fc=(2550.0, 2550.0, 0.0), cl=(100.0, 100.0, 0.0), acf=0, hurst=0.5, sigma=0.1,
               ds=10.0, dh=10.0, nd=2, structured=True, fs=(1, 1, 1), fe=(256, 256, 1))
    out = fo.realization(cfg, seed=3, real=np.float64)
    assert abs(math.sqrt(out["var"]) / 0.1 - 1.0) < 0.1


def test_realization_is_seeded_and_rescaled():
    cfg = _cfg(structured=True, fs=(1, 1, 1), fe=(16, 16, 16), rescale=1, poi=[75.0, 75.0, 75.0], mute=10.0,
               taper=30.0)
    a = fo.realization(cfg, seed=9)
    b = fo.realization(cfg, seed=9)
    c = fo.realization(cfg, seed=10)
    np.testing.assert_array_equal(a["field"], b["field"])
    assert np.abs(a["field"] - c["field"]).max() > 0
    assert a["info"][2] == pytest.approx(0.1) and a["info"][3] == 0.0
    # muted around the point of interest: internal nodes within 10 m of it are 0
    s = a["setup"]
    R = a["internal"]
    k = np.argwhere(R == 0.0)
    assert len(k) > 0


def _hermitian_line_rule(S: np.ndarray, n1: int) -> np.ndarray:
    """The library's symmetry pass (fim_symmetry): the reference's rules, with
    the 3D k = 0 line's pairs (j, n2 - j) replaced by their Hermitian part."""
    out = fo.enforce_symmetry(S, n1)
    n3, n2, _ = S.shape
    if n3 > 1:
        for p in (0, n1 // 2):
            P = S[:, :, p].copy()  # values after the point conditions on k = 0, 0 < j < n2 (untouched there)
            for j in range(1, n2 // 2):
                v = (P[0, j] + P[0, n2 - j]) / 2
                out[0, j, p] = v
                out[0, n2 - j, p] = np.conj(v)
    return out


@pytest.mark.parametrize("shape", [(8, 6, 10), (10, 12, 4), (6, 8, 2), (12, 10, 1)])
def test_hermitian_planes_with_one_c2r_equal_reference_sequence(shape):
    """One 3D c2r of the spectrum whose kx = 0 / Nyquist planes carry the
    Hermitian part of the in-place line equals the reference's z, y transforms
    followed by a c2r along x that reads real parts only."""
    n1, n2, n3 = shape
    rng = np.random.default_rng(sum(shape))
    h = n1 // 2 + 1
    S = rng.normal(size=(n3, n2, h)) + 1j * rng.normal(size=(n3, n2, h))
    ref = fo.transform(fo.enforce_symmetry(S, n1), n1)
    H = _hermitian_line_rule(S, n1)
    for p in (0, n1 // 2):  # the planes are exactly Hermitian now
        P = H[:, :, p]
        mirror = np.conj(P[(-np.arange(n3)) % n3][:, (-np.arange(n2)) % n2])
        np.testing.assert_array_equal(P, mirror)
    got = np.fft.irfftn(H, s=(n3, n2, n1), axes=(0, 1, 2), norm="forward")
    np.testing.assert_allclose(got, ref, atol=1e-12)
