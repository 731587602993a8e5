"""GPU parity: the HIP kernels against the FP64 oracle on the same harmonic
draw (same seed, bit-identical tables, see test_host_parity.py).

Tolerances (stated per test) are in units of the `sigma` parameter; the
bar of the path is max|field - FP64 field| <= 1e-5 sigma for the FP32 and
MFMA kernels that meet it, 1e-11 sigma for the FP64 kernel.
"""
import numpy as np
import pytest

from srm_cases import init_product, oracle_field, structured_cfg, unstructured_cfg

pytestmark = pytest.mark.gpu

TOL = {"direct_f32": 1e-5, "mfma_bf16x3": 1e-5, "mfma_f16x2": 1e-5, "mfma_bf16x2": 2e-4, "direct_f64": 1e-11}


def _say(*a):
    print("[case]", *a, flush=True)


def _run(lib, cfg, seed, kernel):
    init_product(lib, cfg)
    assert lib.set_kernel(kernel) == 0
    field, stats = lib.execute(seed)
    err = lib.last_error()
    assert err == "", err
    return field, stats


GRID_CASES = [
    ("vk_iso", dict(n=(40, 36, 20), nharm=384)),
    ("gauss_aniso", dict(n=(33, 30, 17), acf=1, cl=(300.0, 300.0, 100.0), sigma=0.1, nharm=320)),
    ("rotated_offset", dict(n=(24, 40, 12), fs=(5, 3, 9), alpha=30.0, beta=-10.0, gamma=5.0, nharm=256)),
    ("dropped_harmonics", dict(n=(20, 20, 20), hurst=0.1, ds=20.0, dh=10.0, nharm=300)),
    ("grid_2d", dict(n=(64, 48, 1), cl=(2000.0, 2000.0, 0.0), nharm=256)),
    ("x_tile_96", dict(n=(96, 30, 10), nharm=192)),     # 96-point x tiles
    ("x_tile_128", dict(n=(256, 12, 9), nharm=160)),    # two 128-point x tiles, ragged columns
]


@pytest.mark.parametrize("kernel", ["direct_f32", "mfma_bf16x3", "mfma_f16x2", "mfma_bf16x2"])
@pytest.mark.parametrize("name,kw", GRID_CASES, ids=[c[0] for c in GRID_CASES])
def test_grid_kernels_match_fp64_oracle(oracle, lib_f, name, kw, kernel):
    cfg = structured_cfg(real=np.float32, **kw)
    seed = 1234
    _say(name, kernel, kw)
    ref, _ = oracle_field(oracle, cfg, seed, "f")
    field, stats = _run(lib_f, cfg, seed, kernel)
    ref = ref.reshape(field.shape, order="F")  # 2D grids come back as (nx, ny)
    err = np.abs(field.astype(np.float64) - ref).max() / cfg["sigma"]
    print(f"  max|err|/sigma = {err:.3e}", flush=True)
    assert err <= TOL[kernel]


@pytest.mark.parametrize("name,kw", GRID_CASES[:3], ids=[c[0] for c in GRID_CASES[:3]])
def test_grid_fp64_kernel_double_library(oracle, lib_d, name, kw):
    cfg = structured_cfg(real=np.float64, **kw)
    _say(name, "direct_f64 (libscarf3d)")
    ref, _ = oracle_field(oracle, cfg, 77, "d")
    field, _ = _run(lib_d, cfg, 77, "direct_f64")
    err = np.abs(field - ref).max() / cfg["sigma"]
    print(f"  max|err|/sigma = {err:.3e}", flush=True)
    assert err <= TOL["direct_f64"]


@pytest.mark.parametrize("nd", [3, 2])
@pytest.mark.parametrize("kernel", ["direct_f32", "direct_f64"])
def test_point_list_matches_fp64_oracle(oracle, lib_f, nd, kernel):
    cfg = unstructured_cfg(npts=3001, nd=nd, nharm=300)
    _say("points", nd, kernel)
    ref, _ = oracle_field(oracle, cfg, 7, "f")
    field, _ = _run(lib_f, cfg, 7, kernel)
    err = np.abs(field.astype(np.float64) - ref).max() / cfg["sigma"]
    print(f"  max|err|/sigma = {err:.3e}", flush=True)
    # the FP64 kernel writes float: rounding of the output dominates
    assert err <= TOL["direct_f32"] if kernel == "direct_f32" else err <= 1e-6


@pytest.mark.parametrize("flavour", ["f", "d"])
def test_pipeline_stats_rescale_taper(oracle, lib_f, lib_d, flavour):
    """Whole scarf3d_structured_srm: statistics, rescale (Q2) and sequential
    mute/taper of two POIs against the oracle pipeline."""
    lib = lib_f if flavour == "f" else lib_d
    real = np.float32 if flavour == "f" else np.float64
    poi = np.array([[300.0, 400.0, 200.0], [900.0, 500.0, 600.0]]).T  # (nd, npoi)
    cfg = structured_cfg(n=(32, 24, 20), nharm=256, poi=poi, mute=100.0, taper=250.0, rescale=1, real=real)
    from oracle_lib import OrcParams  # noqa: F401

    _say("pipeline", flavour)
    p = oracle.params(cfg)
    ref_field, ref_stats = oracle.srm(p, 4321, flavour)
    # the rescale multiplies the field, and with it the kernel's error, by
    # sigma / sigma_d (sigma_d: the discrete standard deviation before it),
    # so the 1e-5 sigma bar of the field kernel applies times that factor
    _, plain_stats = oracle.srm(oracle.params(dict(cfg, rescale=0, poi=None)), 4321, flavour)
    gain = max(1.0, cfg["sigma"] / float(plain_stats[2]))
    kernel = "direct_f32" if flavour == "f" else "direct_f64"
    field, stats = _run(lib, cfg, 4321, kernel)
    f = field.reshape(-1, order="F").astype(np.float64)
    r = ref_field.astype(np.float64)
    # muted points are exactly zero in both
    assert np.array_equal(f == 0.0, r == 0.0)
    err = np.abs(f - r).max() / cfg["sigma"]
    print(f"  max|err|/sigma = {err:.3e} (rescale gain {gain:.2f})", flush=True)
    assert err <= (1e-5 * gain if flavour == "f" else 1e-9)
    assert stats[0] == ref_stats[0] and stats[1] == ref_stats[1]
    assert stats[2] == pytest.approx(ref_stats[2], rel=1e-12)  # sigma after rescale
    assert stats[3] == 0.0


@pytest.mark.parametrize("flavour", ["f", "d"])
def test_statistics_without_rescale(oracle, lib_f, lib_d, flavour):
    lib = lib_f if flavour == "f" else lib_d
    real = np.float32 if flavour == "f" else np.float64
    cfg = structured_cfg(n=(48, 40, 24), nharm=256, real=real)
    p = oracle.params(cfg)
    _, ref_stats = oracle.srm(p, 99, flavour)
    field, stats = _run(lib, cfg, 99, "direct_f32" if flavour == "f" else "direct_f64")
    f = field.reshape(-1).astype(np.float64)
    # stats(3:4) of the field the library returned, by the reference formulas
    mean = f.mean()
    var = ((f - mean) ** 2).sum() / (f.size - 1)
    assert stats[2] == pytest.approx(np.sqrt(var), rel=2e-6 if flavour == "f" else 1e-12)
    assert stats[3] == pytest.approx(mean, rel=1e-4, abs=1e-7 * cfg["sigma"])
    # and against the oracle's statistics of its FP64 field
    assert stats[2] == pytest.approx(ref_stats[2], rel=1e-5)
    assert stats[0] == ref_stats[0] and stats[1] == ref_stats[1]
    assert stats[4] > 0.0 and stats[5] > 0.0  # host and device seconds


def test_execute_device_and_resident_run(lib_f):
    """scarf_execute_device / scarf_srm_draw + scarf_srm_run write the same
    field as scarf_execute."""
    from scarf3d_amd.hip import Hip

    hip = Hip()
    cfg = structured_cfg(n=(64, 64, 32), nharm=512)
    field, stats = _run(lib_f, cfg, 5, "direct_f32")
    d = hip.empty(field.size, np.float32)
    st2 = lib_f.execute_device(5, d.ptr)
    assert np.array_equal(d.numpy(), field.reshape(-1, order="F"))
    n, host_s = lib_f.srm_draw(5)
    assert n == 512 and host_s >= 0
    d.zero()
    st3 = lib_f.srm_run(d.ptr)
    assert np.array_equal(d.numpy(), field.reshape(-1, order="F"))
    d.free()
    assert st2[2] == stats[2] and st3[2] == stats[2]
    t = lib_f.timings()
    assert t["field"] > 0 and t["total"] >= t["field"]


def test_same_seed_same_field_different_seed_differs(lib_f):
    cfg = structured_cfg(n=(32, 32, 16), nharm=256)
    a, _ = _run(lib_f, cfg, 11, "mfma_bf16x3")
    lib_f.finalize()
    b, _ = _run(lib_f, cfg, 11, "mfma_bf16x3")
    lib_f.finalize()
    c, _ = _run(lib_f, cfg, 12, "mfma_bf16x3")
    assert np.array_equal(a, b)
    assert not np.allclose(a, c)


def test_scaled_fp16_amplitude_headroom(oracle, lib_f, monkeypatch):
    """The scaled-fp16 scheme multiplies the generated operand's high part by
    2^12, which needs |Z Y| < 8; larger amplitudes scale Z down by a power of
    two and the field's scaling up by the same. The limit is lowered here so
    that the rescaled path runs on an ordinary draw."""
    cfg = structured_cfg(n=(96, 40, 12), nharm=256, real=np.float32)
    ref, _ = oracle_field(oracle, cfg, 31, "f")
    monkeypatch.setenv("SCARF_FP16_AMP_LIMIT", "0.125")
    _say("mfma_f16x2 with amplitudes rescaled by 2^-k")
    field, _ = _run(lib_f, cfg, 31, "mfma_f16x2")
    err = np.abs(field.astype(np.float64) - ref.reshape(field.shape, order="F")).max() / cfg["sigma"]
    print(f"  max|err|/sigma = {err:.3e}", flush=True)
    assert err <= TOL["mfma_f16x2"]


def test_tile_tuner_changes_speed_not_field(oracle, lib_f, monkeypatch, tmp_path):
    """Tuned and least-padding tiles sum the same products in the same order
    per output value: the fields agree (to 1e-7 sigma; the same bits where
    the hardware's per-element MFMA order does not depend on the tile). The
    choice persists through SCARF_TUNE_CACHE."""
    cache = tmp_path / "tiles.txt"
    monkeypatch.setenv("SCARF_TUNE_CACHE", str(cache))
    cfg = structured_cfg(n=(200, 70, 9), nharm=256, real=np.float32)
    _say("tuned tile", cfg["fe"])
    tuned, _ = _run(lib_f, cfg, 3, "mfma_bf16x3")
    tile = lib_f.last_tile()
    assert tile is not None and tile[1] * tile[2] == 256 and tile[0] in (64, 96, 128)
    lines = cache.read_text().split("\n")
    assert any(line.split()[2:5] == ["200", "70", "9"] for line in lines if line)
    lib_f.finalize()
    monkeypatch.setenv("SCARF_TUNE", "0")
    plain, _ = _run(lib_f, cfg, 3, "mfma_bf16x3")
    print(f"  tuned tile {tile}, least-padding tile {lib_f.last_tile()}, "
          f"bit-identical: {np.array_equal(tuned, plain)}", flush=True)
    assert np.abs(tuned.astype(np.float64) - plain).max() <= 1e-7 * cfg["sigma"]


@pytest.mark.parametrize("kernel", ["mfma_bf16x3", "mfma_f16x2"])
def test_every_forced_tile_gives_the_same_field(lib_f, monkeypatch, kernel):
    """SCARF_TILE fixes the block tile: every x width (64/96/128) and column
    shape (256x1 .. 32x8) runs, reports itself through scarf_last_tile, and
    sums the same products per output value as the least-padding tile (ragged
    in x, j and k)."""
    cfg = structured_cfg(n=(150, 70, 11), nharm=320, real=np.float32)
    monkeypatch.setenv("SCARF_TUNE", "0")
    _say("least-padding tile", kernel, cfg["fe"])
    ref, _ = _run(lib_f, cfg, 5, kernel)
    for x in (64, 96, 128):
        for k in (1, 2, 4, 8):
            tile = f"{x},{256 // k},{k}"
            monkeypatch.setenv("SCARF_TILE", tile)
            _say("forced tile", kernel, tile)
            lib_f.finalize()
            field, _ = _run(lib_f, cfg, 5, kernel)
            assert lib_f.last_tile() == (x, 256 // k, k)
            assert np.abs(field.astype(np.float64) - ref).max() <= 1e-7 * cfg["sigma"], tile
    monkeypatch.setenv("SCARF_TILE", "100,2,128")  # not a tile: ignored
    lib_f.finalize()
    field, _ = _run(lib_f, cfg, 5, kernel)
    assert np.abs(field.astype(np.float64) - ref).max() <= 1e-7 * cfg["sigma"]


def test_host_copy_equals_device_field(lib_f):
    """scarf_execute returns a 72 MB field (above the 64 MiB slab threshold)
    in z slabs of 11.5 MB, each staged through the two pinned buffers in
    chunks of 4 MiB (two whole chunks and a partial one per slab: the buffer
    reuse, the partial last chunk and the pooled drain all run), the last,
    2.9 MB slab by one copy; the host array equals the device field bit for
    bit."""
    from scarf3d_amd.hip import Hip

    hip = Hip()
    n = (300, 300, 200)
    cfg = structured_cfg(n=n, nharm=64, real=np.float32)
    _say("host copy of a 72 MB field", n)
    field, _ = _run(lib_f, cfg, 11, "mfma_bf16x3")
    d = hip.empty(int(np.prod(n)), np.float32)
    lib_f.execute_device(11, d.ptr)
    assert np.array_equal(field.reshape(-1, order="F"), d.numpy())
    d.free()


def test_device_memory_limit_reports_and_recovers(lib_f, monkeypatch):
    """Allocation failure, forced from the host: with SCARF_DEVICE_MEMORY_LIMIT
    below what the library's device buffers hold or need, execute fails with an
    error (no field, zero statistics) instead of crashing; without the limit
    the next realization runs and equals a clean one."""
    cfg = structured_cfg(n=(64, 48, 16), nharm=512, real=np.float32)
    ref, st_ref = _run(lib_f, cfg, 21, "mfma_bf16x3")
    lib_f.finalize()
    monkeypatch.setenv("SCARF_DEVICE_MEMORY_LIMIT", "65536")
    for kernel in ("mfma_bf16x3", "direct_f32"):
        _say("device memory limit 64 KiB", kernel)
        init_product(lib_f, cfg)
        assert lib_f.set_kernel(kernel) == 0
        field, stats = lib_f.execute(21)
        assert "SCARF_DEVICE_MEMORY_LIMIT" in lib_f.last_error()
        assert not np.any(stats)
        lib_f.finalize()
    monkeypatch.delenv("SCARF_DEVICE_MEMORY_LIMIT")
    field, stats = _run(lib_f, cfg, 21, "mfma_bf16x3")
    assert np.array_equal(field, ref) and np.array_equal(stats[:4], st_ref[:4])


def _block(n, fs, model, ds=50.0, **kw):
    """Block n from global index fs of a model of `model` points (nc/fc of the
    whole model, as each MPI rank of the reference passes them)."""
    cfg = structured_cfg(n=n, fs=fs, ds=ds, real=np.float32, **kw)
    cfg["nc"] = (0.0, 0.0, 0.0)
    cfg["fc"] = tuple(float(np.float32((m - 1) * ds)) for m in model)
    return cfg


EDGE_CASES = [
    ("one_point_block", lambda: _block((1, 1, 1), (7, 5, 3), (40, 30, 20), nharm=200)),
    ("one_x_plane", lambda: _block((1, 33, 9), (4, 1, 1), (12, 33, 9), nharm=176)),
    ("one_column", lambda: _block((130, 1, 1), (1, 6, 2), (130, 10, 6), nharm=208)),
    ("one_harmonic", lambda: structured_cfg(n=(20, 18, 10), nharm=1, real=np.float32)),
    ("odd_harmonic_count", lambda: structured_cfg(n=(20, 18, 10), nharm=23, real=np.float32)),
]


@pytest.mark.parametrize("kernel", ["direct_f32", "mfma_bf16x3", "mfma_f16x2", "direct_f64"])
@pytest.mark.parametrize("name,make", EDGE_CASES, ids=[c[0] for c in EDGE_CASES])
def test_edge_shapes_match_oracle(oracle, lib_f, name, make, kernel):
    """Degenerate blocks (one point, one x plane, one column of x), one
    harmonic and a harmonic count that pads the MFMA stage: field against the
    oracle, statistics against the oracle pipeline. The variance's n - 1
    divisor makes stats(3) of a one-point block 0/0 = NaN, in the reference's
    formula (scarflib_common.f90:684) as in the library."""
    cfg = make()
    seed = 4242
    _say(name, kernel, cfg["fs"], cfg["fe"], "nharm", cfg["nharm"])
    ref, _ = oracle_field(oracle, cfg, seed, "f")
    _, ref_stats = oracle.srm(oracle.params(cfg), seed, "f")
    field, stats = _run(lib_f, cfg, seed, kernel)
    ref = ref.reshape(field.shape, order="F")
    err = np.abs(field.astype(np.float64) - ref).max() / cfg["sigma"]
    print(f"  max|err|/sigma = {err:.3e}, stats {stats[:4]} vs {ref_stats[:4]}", flush=True)
    assert err <= (1e-6 if kernel == "direct_f64" else TOL[kernel])
    assert stats[0] == ref_stats[0] and stats[1] == ref_stats[1]
    if field.size == 1:
        assert np.isnan(ref_stats[2]) and np.isnan(stats[2])
        assert stats[3] == pytest.approx(ref_stats[3], rel=1e-5, abs=1e-6 * cfg["sigma"])
    else:
        assert stats[2] == pytest.approx(ref_stats[2], rel=1e-5, abs=1e-7 * cfg["sigma"])
        assert abs(stats[3] - ref_stats[3]) <= 1e-5 * cfg["sigma"]


@pytest.mark.parametrize("npts", [1, 65])
def test_short_point_lists_match_oracle(oracle, lib_f, npts):
    """Point lists of one point and of one 64-point block plus one, inside a
    model whose nc/fc are given: field and statistics against the oracle."""
    cfg = unstructured_cfg(npts=npts, nharm=300, seed=3, nc=(0.0, 0.0, 0.0), fc=(10000.0, 10000.0, 5000.0))
    _say("points", npts)
    ref, _ = oracle_field(oracle, cfg, 11, "f")
    _, ref_stats = oracle.srm(oracle.params(cfg), 11, "f")
    field, stats = _run(lib_f, cfg, 11, "direct_f32")
    err = np.abs(field.astype(np.float64) - ref).max() / cfg["sigma"]
    print(f"  max|err|/sigma = {err:.3e}, stats {stats[:4]} vs {ref_stats[:4]}", flush=True)
    assert err <= TOL["direct_f32"]
    if npts == 1:
        assert np.isnan(ref_stats[2]) and np.isnan(stats[2])
    else:
        assert stats[2] == pytest.approx(ref_stats[2], rel=1e-5)
