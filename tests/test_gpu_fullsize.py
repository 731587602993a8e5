"""Full-size parity on the benchmark configurations.

512 x 512 x 512, 20,000 harmonics (the configuration the headline throughput
is quoted on): the split-bf16 and scaled-fp16 MFMA fields against the FP64
GPU kernel on the same draw at every point (bar 1e-5 sigma), the FP64 kernel
against the FP64 CPU oracle at 256 sampled points, and the fused statistics
against the field they were computed from.

1024 x 1024 x 512, 50,000 harmonics, 16 points of interest with mute and
taper, rescaled to sigma: the split-bf16 and scaled-fp16 fields against the
FP64 kernel's FP64 field, each after its own statistics, rescale and taper
(bar 1e-5 sigma), and the FP64 CPU oracle at 256 sampled points through the
same rescale and taper.

128 x 128 x 128, 5,000 harmonics with the double library, and 2,000,000
scattered points with 10,000 harmonics (FP64 coordinates through the double
library, and rounded to float through the float library): every kernel
against the FP64 kernel at every point, and against the oracle at 256
sampled points.

The 64-member ensemble: every member equal to its single run bit for bit.

FIM (method 0) on the 1024 x 1024 x 512 grid with padding, rescale and taper:
z-slab blocks equal to the whole field bit for bit, and the rescale and mute
invariants."""
import numpy as np
import pytest

from srm_cases import init_product, structured_cfg, unstructured_cfg

pytestmark = pytest.mark.gpu


def _flat(field):
    return field.reshape(-1, order="F")


def _max_abs_diff(a, b, chunk=1 << 26):
    """max |a - b| in FP64, a slab at a time (the fields hold up to 5e8 points)"""
    a, b = _flat(a), _flat(b)
    m = 0.0
    for i in range(0, a.size, chunk):
        m = max(m, float(np.abs(a[i:i + chunk].astype(np.float64) - b[i:i + chunk]).max()))
    return m


def _run(lib, kernel, seed):
    print(f"[case] {kernel}", flush=True)
    assert lib.set_kernel(kernel) == 0
    field, stats = lib.execute(seed)
    assert lib.last_error() == ""
    return field, stats


def test_headline_configuration_full_size(oracle, lib_f):
    n = (512, 512, 512)
    cfg = structured_cfg(n=n, ds=20.0, acf=0, hurst=0.1, cl=(5000.0, 5000.0, 5000.0), sigma=0.05, nharm=20000,
                         real=np.float32)
    seed = 1234
    print("[case] 512^3 x 20000", flush=True)
    init_product(lib_f, cfg)
    ref, _ = _run(lib_f, "direct_f64", seed)
    for kernel in ("mfma_bf16x3", "mfma_f16x2"):
        field, stats = _run(lib_f, kernel, seed)
        err = _max_abs_diff(field, ref) / cfg["sigma"]
        print(f"  {kernel} vs FP64 kernel, every point: max|err|/sigma = {err:.3e}", flush=True)
        assert err <= 1e-5
        # statistics of the returned field by the reference formulas
        f = _flat(field).astype(np.float64)
        mean = f.mean()
        var = ((f - mean) ** 2).sum() / (f.size - 1)
        assert stats[2] == pytest.approx(np.sqrt(var), rel=2e-6)
        assert stats[3] == pytest.approx(mean, rel=1e-4, abs=1e-7 * cfg["sigma"])
        del f, field

    # the FP64 kernel (its output rounded to float) against the FP64 oracle
    p = oracle.params(cfg)
    s = oracle.setup(p, "f")
    harm, used, _ = oracle.draw(p, s, seed, "f")
    rng = np.random.default_rng(5)
    ii, jj, kk = (rng.integers(0, n[a], 256) for a in range(3))
    vals = oracle.field_at(p, s, harm, used, ii, jj, kk)
    flat = ii + n[0] * (jj + n[1] * kk)
    err64 = np.abs(_flat(ref)[flat].astype(np.float64) - vals).max() / cfg["sigma"]
    print(f"  FP64 kernel vs oracle at 256 points: max|err|/sigma = {err64:.3e}", flush=True)
    assert err64 <= 1e-6


def test_large_grid_with_taper_and_rescale_full_size(oracle, lib_f):
    from oracle_lib import epilogue_at
    from scarf3d_amd.hip import Hip

    n = (1024, 1024, 512)
    ds = 20.0
    rng = np.random.default_rng(16)
    ext = np.array([(n[a] - 1) * ds for a in range(3)])
    poi = rng.uniform(0, 1, (3, 16)) * ext[:, None]
    cfg = structured_cfg(n=n, ds=ds, acf=0, hurst=0.2, cl=(6000.0, 4000.0, 1500.0), sigma=0.08, nharm=50000,
                         rescale=1, poi=poi, mute=20.0, taper=500.0, real=np.float32)
    seed = 1234
    print("[case] 1024 x 1024 x 512 x 50000, 16 POIs, rescale: FP64 kernel (FP64 output)", flush=True)
    init_product(lib_f, cfg)
    # the FP64 kernel's FP64 field after its own FP64 statistics, rescale and taper
    hip = Hip()
    lib_f.srm_draw(seed)
    d_ref = hip.empty(int(np.prod(n)), np.float64)
    st64 = lib_f.srm_run_fp64(d_ref.ptr)  # {count, mean, variance, sigma_d} before the epilogue
    ref = d_ref.numpy()
    d_ref.free()
    for kernel in ("mfma_bf16x3", "mfma_f16x2"):
        field, st = _run(lib_f, kernel, seed)
        err = _max_abs_diff(field, ref) / cfg["sigma"]
        print(f"  {kernel} vs FP64 kernel after rescale and taper: max|err|/sigma = {err:.3e}", flush=True)
        assert err <= 1e-5
        # rescaled: stats(3) reports sigma and stats(4) zero
        assert st[2] == np.float32(cfg["sigma"]) and st[3] == 0.0
        if kernel == "mfma_bf16x3":
            bf = _flat(field)
        del field
    # the FP64 CPU oracle at 256 sampled points, through the reference's
    # rescale and taper with the FP64 field's moments
    p = oracle.params(cfg)
    s = oracle.setup(p, "f")
    harm, used, _ = oracle.draw(p, s, seed, "f")
    rng = np.random.default_rng(5)
    ii, jj, kk = (rng.integers(0, n[a], 256) for a in range(3))
    flat = ii + n[0] * (jj + n[1] * kk)
    vals = epilogue_at(oracle.field_at(p, s, harm, used, ii, jj, kk), flat, cfg, st64[1], st64[2])
    err64 = np.abs(ref[flat] - vals).max() / cfg["sigma"]
    errbf = np.abs(bf[flat].astype(np.float64) - vals).max() / cfg["sigma"]
    print(f"  oracle at 256 points after rescale and taper: FP64 kernel {err64:.3e}, mfma_bf16x3 {errbf:.3e} sigma",
          flush=True)
    assert err64 <= 1e-6 and errbf <= 1e-5
    assert np.count_nonzero(vals == 0.0) == np.count_nonzero(ref[flat] == 0.0)  # muted points


def test_plain_128_double_library_full_size(oracle, lib_d):
    n = (128, 128, 128)
    cfg = structured_cfg(n=n, ds=50.0, acf=0, hurst=0.25, cl=(2000.0, 2000.0, 2000.0), sigma=0.05, nharm=5000,
                         real=np.float64)
    seed = 1234
    print("[case] 128^3 x 5000, double library", flush=True)
    init_product(lib_d, cfg)
    ref, _ = _run(lib_d, "direct_f64", seed)
    for kernel in ("direct_f32", "mfma_bf16x3", "mfma_f16x2"):
        field, _ = _run(lib_d, kernel, seed)
        err = _max_abs_diff(field, ref) / cfg["sigma"]
        print(f"  {kernel} vs FP64 kernel, every point: max|err|/sigma = {err:.3e}", flush=True)
        assert err <= 1e-5
    p = oracle.params(cfg)
    s = oracle.setup(p, "d")
    harm, used, _ = oracle.draw(p, s, seed, "d")
    rng = np.random.default_rng(5)
    ii, jj, kk = (rng.integers(0, n[a], 256) for a in range(3))
    vals = oracle.field_at(p, s, harm, used, ii, jj, kk)
    err64 = np.abs(_flat(ref)[ii + n[0] * (jj + n[1] * kk)] - vals).max() / cfg["sigma"]
    print(f"  FP64 kernel vs oracle at 256 points: max|err|/sigma = {err64:.3e}", flush=True)
    assert err64 <= 1e-11


@pytest.mark.parametrize("flavour", ["f", "d"])
def test_points_configuration_full_size(oracle, lib_f, lib_d, flavour):
    """2,000,000 points with FP64 coordinates through the double library
    (the configuration's coordinates), and the same points rounded to float
    through the float library; the FP32 direct kernel in both."""
    lib = lib_f if flavour == "f" else lib_d
    real = np.float32 if flavour == "f" else np.float64
    cfg = unstructured_cfg(npts=2_000_000, box=(10000.0, 10000.0, 5000.0), dh=50.0, acf=1,
                           cl=(3000.0, 3000.0, 1000.0), sigma=0.1, hurst=0.25, nharm=10000, seed=7, real=real)
    seed = 7
    print(f"[case] 2,000,000 points x 10000, library {flavour}", flush=True)
    init_product(lib, cfg)
    ref, _ = _run(lib, "direct_f64", seed)
    field, _ = _run(lib, "direct_f32", seed)
    err = _max_abs_diff(field, ref) / cfg["sigma"]
    print(f"  direct_f32 vs FP64 kernel, every point: max|err|/sigma = {err:.3e}", flush=True)
    assert err <= 1e-5
    p = oracle.params(cfg)
    s = oracle.setup(p, flavour)
    harm, used, _ = oracle.draw(p, s, seed, flavour)
    flat = np.random.default_rng(5).integers(0, p.npts, 256)
    sub = dict(cfg, x=cfg["x"][flat], y=cfg["y"][flat], z=cfg["z"][flat])  # same parameters, sampled points
    vals = oracle.field_unstructured(oracle.params(sub), s, harm, used)
    err64 = np.abs(_flat(ref)[flat].astype(np.float64) - vals).max() / cfg["sigma"]
    errf = np.abs(_flat(field)[flat].astype(np.float64) - vals).max() / cfg["sigma"]
    print(f"  oracle at 256 points: FP64 kernel {err64:.3e}, direct_f32 {errf:.3e} sigma", flush=True)
    assert err64 <= (1e-6 if flavour == "f" else 1e-11) and errf <= 1e-5


def test_ensemble_64_full_size(lib_f):
    """The 64-member benchmark ensemble (grids 96^3 .. 384 x 384 x 192, 2k-16k
    harmonics; one persistent launch over mixed x tiles, items dealt to the
    least-loaded block): every member's field and statistics equal its single
    run bit for bit."""
    from scarf3d_amd.hip import Hip
    from scarf3d_amd.workloads import ensemble_members

    hip = Hip()
    members = ensemble_members()
    bufs = [hip.empty(int(np.prod(m["fe"])), np.float32) for m in members]
    print("[case] ensemble of 64, mfma_bf16x3", flush=True)
    stats = lib_f.ensemble_execute(members, [b.ptr for b in bufs], "mfma_bf16x3")
    for i, m in enumerate(members):
        cfg = structured_cfg(n=m["fe"], ds=m["ds"], acf=m["acf"], cl=m["cl"], sigma=m["sigma"],
                             hurst=m.get("hurst", 0.0), nharm=m["nharm"], real=np.float32)
        init_product(lib_f, cfg)
        assert lib_f.set_kernel("mfma_bf16x3") == 0
        single, st = lib_f.execute(m["seed"])
        lib_f.finalize()
        assert np.array_equal(bufs[i].numpy(), _flat(single)), i
        assert np.array_equal(stats[i, :4], st[:4]), i
    print(f"  {len(members)} members equal their single runs", flush=True)
    for b in bufs:
        b.free()


def test_fim_large_grid_full_size(lib_f):
    """FIM (method 0) on the 1024 x 1024 x 512 grid with padding, rescale and the
    16 tapered points of interest (internal grid 1626 x 1426 x 664 with the
    padding of cl/dh points per side; 271 and 83 are prime, so rocFFT takes
    Bluestein's algorithm). Size-independent checks, the oracle
    being covered on smaller grids (tests/test_gpu_fim.py):
      - four z-slab blocks computed separately (the model's nc/fc, as MPI
        ranks of the reference pass them) equal the whole field bit for bit;
      - the rescale reports sigma and zero mean, and the field's values
        outside the tapered zones carry the rescaled spread: their standard
        deviation is sigma within sampling (the internal grid is larger than
        the output block);
      - the points of interest are muted: the internal node behind the output
        point nearest each one lies within 2 x 10 sqrt(3) m of it, where the
        taper factor (mute 20 m, taper 500 m) is below 0.0022, so that value is
        below 0.02 sigma.
    """
    from scarf3d_amd.workloads import WORKLOADS, init_library, workload_cfg

    cfg = workload_cfg(WORKLOADS["large_1024"])
    cfg["method"], cfg["pad"] = 0, 1
    n = tuple(cfg["fe"])
    print("[case] FIM 1024 x 1024 x 512, pad, rescale, 16 POIs", flush=True)
    init_library(lib_f, cfg)
    whole, st = lib_f.execute(77)
    assert lib_f.last_error() == "", lib_f.last_error()
    lib_f.finalize()
    assert float(st[2]) == pytest.approx(cfg["sigma"], rel=1e-6) and st[3] == 0.0
    sample = whole[::7, ::5, ::3].astype(np.float64)
    assert abs(sample.std() / cfg["sigma"] - 1.0) < 0.25
    poi = np.asarray(cfg["poi"])
    for q in range(poi.shape[1]):
        i, j, k = (int(round(poi[d, q] / cfg["ds"])) for d in range(3))
        assert abs(float(whole[min(i, n[0] - 1), min(j, n[1] - 1), min(k, n[2] - 1)])) <= 0.02 * cfg["sigma"], q
    for b in range(4):
        z0 = b * (n[2] // 4)
        c = dict(cfg)
        c["fs"] = (1, 1, z0 + 1)
        c["fe"] = (n[0], n[1], z0 + n[2] // 4)
        print(f"[case] FIM block z {z0 + 1}..{z0 + n[2] // 4}", flush=True)
        init_library(lib_f, c)
        part, stb = lib_f.execute(77)
        assert lib_f.last_error() == "", lib_f.last_error()
        lib_f.finalize()
        assert np.array_equal(part, whole[:, :, z0:z0 + n[2] // 4]), b
        assert np.array_equal(stb[:4], st[:4]), b
