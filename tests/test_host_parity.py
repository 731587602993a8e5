"""CPU tests of the product library's host side and C ABI (no GPU compute):

- both library flavours load and export every function include/*.h declares;
- the product's setup (band, QUADPACK normalisation, rotation, bar) and its
  harmonic draw are bit-identical to the oracle's, for both `real` types;
- the object layer applies the reference defaults and the option presence
  rules of src/c_binding.c:89-103 / src/scarflib.f90:132-189, :243-294;
- error paths report instead of crashing, and execute without a GPU fails
  loudly (no CPU fallback).
"""
import re
from pathlib import Path

import numpy as np
import pytest

from srm_cases import init_product, structured_cfg, unstructured_cfg

ROOT = Path(__file__).resolve().parents[1]


def _declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        c_part = text.split("#ifdef __cplusplus\n}")[0] if "#ifdef __cplusplus\n}" in text else text
        for m in re.finditer(r"^\s*(?:void|int|long|const char|double|void \*)\s*\*?\s*(scarf_\w+)\s*\(", c_part, re.M):
            names.add(m.group(1))
    return names


@pytest.mark.parametrize("flavour", ["f", "d"])
def test_library_exports_every_declared_function(flavour):
    import scarf3d_amd

    lib = scarf3d_amd.load(flavour)
    declared = _declared_functions()
    assert len(declared) >= 20
    for name in sorted(declared):
        assert hasattr(lib.lib, name), name
    assert declared == set(scarf3d_amd.api.EXPORTS)


SETUP_CASES = [
    dict(),
    dict(acf=1, cl=(300.0, 300.0, 100.0), sigma=0.1),
    dict(alpha=30.0, beta=-10.0, gamma=5.0),
    dict(cl=(2000.0, 2000.0, 0.0), n=(32, 32, 1)),
    dict(hurst=0.1, ds=20.0, dh=10.0),
    dict(n=(20, 30, 10), fs=(4, 2, 7), hurst=0.5, cl=(1500.0, 900.0, 600.0)),
    dict(acf=1, cl=(150.0, 150.0, 0.0), n=(40, 40, 1), alpha=45.0),
]


@pytest.mark.parametrize("flavour", ["f", "d"])
@pytest.mark.parametrize("case", range(len(SETUP_CASES)))
def test_setup_and_draw_bit_identical_to_oracle(oracle, lib_f, lib_d, flavour, case):
    lib = lib_f if flavour == "f" else lib_d
    cfg = structured_cfg(nharm=400, real=np.float32 if flavour == "f" else np.float64, **SETUP_CASES[case])
    p = oracle.params(cfg)
    s = oracle.setup(p, flavour)
    init_product(lib, cfg)
    sv = lib.setup_values()
    assert sv["ndim"] == s.ndim
    for key, okey in (("kmin", "kmin"), ("kmax", "kmax"), ("kc", "kc"), ("info1", "info1"), ("info2", "info2"),
                      ("scaling", "scaling"), ("band", "intg_band"), ("all", "intg_all")):
        assert sv[key] == getattr(s, okey), key
    assert np.array_equal(sv["matrix"], np.array(s.matrix[:]))
    assert np.array_equal(sv["bar"], np.array(s.bar[:]))
    for seed in (1, 1234, -5):
        h1, u1, n1 = oracle.draw(p, s, seed, flavour)
        h2, u2, n2 = lib.draw_harmonics(seed, 400)
        assert n1 == n2
        assert np.array_equal(u1, u2)
        assert np.array_equal(h1, h2)


@pytest.mark.parametrize("flavour", ["f", "d"])
@pytest.mark.parametrize("threads,block", [(1, 0), (1, 5), (3, 4099), (8, 4096), (16, 0)])
def test_parallel_draw_consumes_the_same_stream(oracle, lib_f, lib_d, flavour, threads, block):
    """The host draw evaluates the stream's trials ahead of the loop, in
    blocks, on several threads that jump into the stream; whatever the cut
    (a block of 5 pairs splits most harmonics across blocks; 4099 pairs over
    3 threads leaves ragged chunks), the harmonics and the number of uniforms
    consumed are those of the oracle's one-pair-at-a-time loop."""
    lib = lib_f if flavour == "f" else lib_d
    real = np.float32 if flavour == "f" else np.float64
    cfg = structured_cfg(n=(64, 64, 64), ds=20.0, hurst=0.1, cl=(5000.0,) * 3, nharm=1500, real=real)
    p = oracle.params(cfg)
    s = oracle.setup(p, flavour)
    init_product(lib, cfg)
    assert lib.set_option("draw_threads", threads) == 0
    assert lib.set_option("draw_block", block) == 0
    for seed in (1234, 98765):
        h1, u1, n1 = oracle.draw(p, s, seed, flavour)
        h2, u2, n2 = lib.draw_harmonics(seed, 1500)
        assert n1 == n2 and n1 > 100 * 1500  # low acceptance: a long stream
        assert np.array_equal(u1, u2)
        assert np.array_equal(h1, h2)


@pytest.mark.parametrize("flavour", ["f", "d"])
@pytest.mark.parametrize("chunk", [1, 7, 128, 640])
def test_resumable_draw_is_cut_invariant(oracle, lib_f, lib_d, flavour, chunk):
    """The streamed execute draws in chunks of kept harmonics, resuming the
    walk where it stopped; cut every `chunk` kept harmonics (and with dh <
    ds, so that harmonics above kc are dropped between and at the cuts), the
    draw equals the oracle's uncut loop."""
    lib = lib_f if flavour == "f" else lib_d
    real = np.float32 if flavour == "f" else np.float64
    cfg = structured_cfg(n=(64, 64, 64), ds=20.0, dh=8.0, hurst=0.3, cl=(3000.0,) * 3, nharm=1300, real=real)
    p = oracle.params(cfg)
    s = oracle.setup(p, flavour)
    init_product(lib, cfg)
    assert lib.set_option("draw_threads", 4) == 0
    assert lib.set_option("draw_chunk", chunk) == 0
    h1, u1, n1 = oracle.draw(p, s, 4321, flavour)
    h2, u2, n2 = lib.draw_harmonics(4321, 1300)
    assert 0 < u1.sum() < 1300  # some harmonics dropped above kc
    assert n1 == n2 and np.array_equal(u1, u2) and np.array_equal(h1, h2)


SQUEEZE_CASES = [
    # the band [kmin, kmax] spans the maximum of srm_fn * const: bins on both
    # sides of it take the bound, the bins around it the exact test
    dict(n=(64, 64, 64), ds=20.0, hurst=0.25, cl=(200.0,) * 3, nharm=900),                 # VK 3D, peak 1.15
    dict(n=(80, 80, 1), ds=20.0, hurst=0.4, cl=(150.0, 150.0, 0.0), nharm=900),            # VK 2D, peak 0.77
    dict(n=(40, 40, 40), ds=50.0, acf=1, cl=(300.0,) * 3, nharm=900),                     # Gaussian 3D, peak 2
    dict(n=(60, 60, 1), ds=25.0, acf=1, cl=(200.0, 200.0, 0.0), nharm=900),               # Gaussian 2D, peak 1.41
]


@pytest.mark.parametrize("flavour", ["f", "d"])
@pytest.mark.parametrize("case", range(len(SQUEEZE_CASES)))
def test_squeezed_trials_decide_as_the_exact_test(oracle, lib_f, lib_d, flavour, case):
    """The draw rejects most trials against a per-bin upper bound of the
    acceptance and evaluates the spectrum only for the rest; with the
    spectrum's maximum inside the sampled band (bins rising, falling, and the
    unbounded bins around the maximum), the draw and the uniforms consumed
    equal the oracle's, which evaluates every trial."""
    lib = lib_f if flavour == "f" else lib_d
    real = np.float32 if flavour == "f" else np.float64
    cfg = structured_cfg(real=real, **SQUEEZE_CASES[case])
    p = oracle.params(cfg)
    s = oracle.setup(p, flavour)
    init_product(lib, cfg)
    for seed in (3, 4444):
        h1, u1, n1 = oracle.draw(p, s, seed, flavour)
        h2, u2, n2 = lib.draw_harmonics(seed, cfg["nharm"])
        assert n1 == n2 and np.array_equal(u1, u2) and np.array_equal(h1, h2)


def test_stream_chunk_option_checked(lib_f):
    """Streamed chunks start at multiples of the field kernels' 128-harmonic
    promotion period: other chunk sizes are refused."""
    init_product(lib_f, structured_cfg(nharm=64))
    assert lib_f.set_option("stream_chunk", 100) != 0
    assert "multiple of 128" in lib_f.last_error()
    assert lib_f.set_option("stream_chunk", 256) == 0
    assert lib_f.set_option("stream_draw", 0) == 0


def test_draw_options_rejected_when_negative(lib_f):
    init_product(lib_f, structured_cfg(nharm=64))
    assert lib_f.set_option("draw_threads", -1) != 0
    assert "draw_threads" in lib_f.last_error()


def test_unstructured_setup_bit_identical(oracle, lib_f):
    cfg = unstructured_cfg(npts=500, nharm=200)
    p = oracle.params(cfg)
    s = oracle.setup(p, "f")
    init_product(lib_f, cfg)
    sv = lib_f.setup_values()
    assert sv["kmin"] == s.kmin and sv["scaling"] == s.scaling
    h1, u1, _ = oracle.draw(p, s, 7, "f")
    h2, u2, _ = lib_f.draw_harmonics(7, 200)
    assert np.array_equal(h1, h2) and np.array_equal(u1, u2)


def test_structured_defaults(lib_f):
    """initialize_structured defaults (src/scarflib.f90:132-158): dh = ds,
    method 0, hurst 0, mute = taper = -999, nc/fc from fs/fe."""
    lib_f.cart_initialize(3, [2, 3, 4], [11, 13, 15], 25.0, 0, [1000.0, 1000.0, 1000.0], 0.1, None)
    c = lib_f.config_values()
    assert c["dh"] == 25.0 and c["method"] == 0 and c["hurst"] == 0
    assert c["mute"] == -999 and c["taper"] == -999 and c["rescale"] == 0
    assert (c["nc1"], c["nc2"], c["nc3"]) == (25.0, 50.0, 75.0)
    assert (c["fc1"], c["fc2"], c["fc3"]) == (250.0, 300.0, 350.0)


def test_option_presence_rules(lib_f):
    """Options count only when method == 1, hurst > 0, dh > 0, mute > 0,
    taper > 0, rescale == 1, nc and fc both set, angles != 0
    (src/c_binding.c:89-103)."""
    o = lib_f.opt(method=2, hurst=-0.3, dh=0.0, mute=0.0, taper=-1.0, rescale=2, pad=3, alpha=0.0)
    lib_f.cart_initialize(3, [1, 1, 1], [8, 8, 8], 10.0, 0, [100.0, 100.0, 100.0], 0.1, o)
    c = lib_f.config_values()
    assert c["method"] == 0 and c["hurst"] == 0 and c["dh"] == 10.0
    assert c["mute"] == -999 and c["taper"] == -999 and c["rescale"] == 0 and c["pad"] == 0
    lib_f.finalize()
    o = lib_f.opt(method=1, hurst=0.3, dh=5.0, mute=20.0, taper=40.0, rescale=1, pad=1, alpha=10.0, beta=-3.0,
                  gamma=2.0, nc=[-5.0, -6.0, -7.0], fc=[500.0, 600.0, 700.0])
    lib_f.cart_initialize(3, [1, 1, 1], [8, 8, 8], 10.0, 0, [100.0, 100.0, 100.0], 0.1, o)
    c = lib_f.config_values()
    assert c["method"] == 1 and c["hurst"] == pytest.approx(0.3) and c["dh"] == 5.0
    assert c["mute"] == 20.0 and c["taper"] == 40.0 and c["rescale"] == 1 and c["pad"] == 1
    assert (c["alpha"], c["beta"], c["gamma"]) == pytest.approx((10.0, -3.0, 2.0))
    assert (c["nc1"], c["fc3"]) == (-5.0, 700.0)
    lib_f.finalize()
    # nc without fc is ignored
    o = lib_f.opt(nc=[-5.0, -6.0, -7.0])
    lib_f.cart_initialize(3, [1, 1, 1], [8, 8, 8], 10.0, 0, [100.0, 100.0, 100.0], 0.1, o)
    assert lib_f.config_values()["nc1"] == 0.0


def test_unstructured_defaults(lib_f):
    """nc/fc default to the coordinate extents, ds = dh; 2D zero z
    (src/scarflib.f90:236-271)."""
    x = np.array([3.0, -2.0, 7.0])
    y = np.array([1.0, 4.0, -6.0])
    z = np.array([0.5, 9.0, 2.0])
    lib_f.nocart_initialize(3, x, y, z, 20.0, 1, [100.0, 100.0, 50.0], 0.1, None)
    c = lib_f.config_values()
    assert (c["nc1"], c["nc2"], c["nc3"]) == (-2.0, -6.0, 0.5)
    assert (c["fc1"], c["fc2"], c["fc3"]) == (7.0, 4.0, 9.0)
    assert c["ds"] == 20.0 and c["structured"] == 0 and c["npts"] == 3
    lib_f.finalize()
    lib_f.nocart_initialize(2, x, y, None, 20.0, 1, [100.0, 100.0], 0.1, lib_f.opt(ds=40.0))
    c = lib_f.config_values()
    assert (c["nc3"], c["fc3"], c["cl3"]) == (0.0, 0.0, 0.0) and c["ds"] == 40.0


def test_initialize_twice_and_execute_before_initialize(lib_f):
    field = np.zeros(8, dtype=np.float32)
    stats = np.full(8, 7.0, dtype=np.float32)
    import ctypes as C

    lib_f.lib.scarf_execute(1, field.ctypes.data_as(C.c_void_p), stats.ctypes.data_as(C.POINTER(C.c_float)))
    assert "before initialize" in lib_f.last_error()
    assert np.all(stats == 0)
    lib_f.cart_initialize(3, [1, 1, 1], [2, 2, 2], 10.0, 0, [100.0, 100.0, 100.0], 0.1, None)
    lib_f.cart_initialize(3, [1, 1, 1], [4, 4, 4], 10.0, 0, [100.0, 100.0, 100.0], 0.1, None)
    assert "twice" in lib_f.last_error()
    assert lib_f.config_values()["fe1"] == 2  # the first object is kept
    lib_f.finalize()
    lib_f.cart_initialize(3, [1, 1, 1], [4, 4, 4], 10.0, 0, [100.0, 100.0, 100.0], 0.1, None)
    assert lib_f.last_error() == "" and lib_f.config_values()["fe1"] == 4


@pytest.mark.parametrize("opts,msg", [
    (dict(), "hurst"),                            # FIM (method 0, the default), von Karman needs hurst > 0
    (dict(method=1), "hurst"),                    # SRM likewise
])
def test_parameter_errors_reported(lib_f, opts, msg):
    lib_f.cart_initialize(3, [1, 1, 1], [16, 16, 16], 50.0, 0, [2000.0] * 3, 0.05, lib_f.opt(**opts))
    field, stats = lib_f.execute(1)
    assert msg in lib_f.last_error()
    assert np.all(field == 0) and np.all(stats == 0)


@pytest.mark.parametrize("method", [0, 1])
def test_user_defined_psd_rejected(lib_f, method):
    lib_f.cart_initialize(3, [1, 1, 1], [16, 16, 16], 50.0, 2, [2000.0] * 3, 0.05,
                          lib_f.opt(method=method, hurst=0.3))
    lib_f.execute(1)
    assert "acf" in lib_f.last_error()


def test_band_without_power_rejected_instead_of_hanging(lib_f):
    # Gaussian spectrum on a box much smaller than cl: the band holds ~1e-68
    # of the spectrum and the reference's rejection loop would never end
    lib_f.cart_initialize(3, [1, 1, 1], [16, 16, 16], 50.0, 1, [3000.0, 3000.0, 1000.0], 0.1, lib_f.opt(method=1))
    lib_f.execute(1)
    assert "no usable power" in lib_f.last_error()


def test_execute_without_gpu_fails_loudly(lib_f):
    from scarf3d_amd.hip import Hip

    if Hip().device_count() > 0:
        pytest.skip("a GPU is present")
    cfg = structured_cfg(n=(8, 8, 8), nharm=16)
    init_product(lib_f, cfg)
    field, stats = lib_f.execute(3)
    assert "no HIP device" in lib_f.last_error()
    assert np.all(field == 0) and np.all(stats == 0)


def test_nharm_setting(lib_f):
    assert lib_f.lib.scarf_get_nharm() == 10000 or lib_f.set_nharm(10000) == 0
    assert lib_f.set_nharm(0) != 0
    assert lib_f.set_nharm(2500) == 0 and lib_f.lib.scarf_get_nharm() == 2500
    lib_f.set_nharm(10000)


def test_io_writers(lib_f, tmp_path):
    """scarf_io_one / scarf_io_slice: raw native-endian real, x fastest
    (single-writer versions of src/scarflib_common.f90:245-646)."""
    n = (5, 4, 3)
    f = np.arange(np.prod(n), dtype=np.float32).reshape(n, order="F")
    lib_f.io_one(n, f, tmp_path / "vol.bin")
    back = np.fromfile(tmp_path / "vol.bin", dtype=np.float32).reshape(n, order="F")
    assert np.array_equal(back, f)
    for axis, plane, ref in (("x", 2, f[1, :, :]), ("y", 4, f[:, 3, :]), ("z", 1, f[:, :, 0])):
        lib_f.io_slice(n, axis, plane, f, tmp_path / f"s{axis}.bin")
        got = np.fromfile(tmp_path / f"s{axis}.bin", dtype=np.float32).reshape(ref.shape, order="F")
        assert np.array_equal(got, ref), axis
    # bad sizes, planes and axes are reported and write nothing
    lib_f.io_one((5, 0, 3), f, tmp_path / "bad.bin")
    assert "npts must be positive" in lib_f.last_error() and not (tmp_path / "bad.bin").exists()
    lib_f.io_slice((5, -4, 3), "z", 1, f, tmp_path / "bad.bin")
    assert "npts must be positive" in lib_f.last_error() and not (tmp_path / "bad.bin").exists()
    lib_f.io_slice(n, "x", 6, f, tmp_path / "bad.bin")
    assert "plane out of range" in lib_f.last_error() and not (tmp_path / "bad.bin").exists()
    lib_f.io_slice(n, "w", 1, f, tmp_path / "bad.bin")
    assert "axis must be x, y or z" in lib_f.last_error() and not (tmp_path / "bad.bin").exists()
