"""The C and C++ drop-in drivers (tests/c/srm_driver.c,
tests/cpp/srm_driver.cpp): callers written against include/scarf3d.h the
way the reference's examples are (examples/c/driver.c:526-612 with
method = 1 and its FIM block :405-430; examples/c++/driver.cpp with
Scarf3D::Initialize<srm> and Scarf3D::Initialize<fim>), built by
scarf3d_amd/build.py for both library flavours. On the GPU every field they
write equals the ctypes path's on the same parameters, bit for bit, and
their printed statistics equal its stats(1:4)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

LIB = Path(__file__).resolve().parents[1] / "scarf3d_amd" / "lib"
N = (40, 36, 20)
DS, CL, SIGMA, SEED, HURST = 50.0, (2000.0,) * 3, 0.05, 1235, 0.25


def _poi(real):
    """the drivers' two points of interest (integer grid positions times ds, exact in both flavours)"""
    n0, n1, n2 = N
    p = [(4 * n0 // 5), n1 // 2, 2 * n2 // 5, 3 * n0 // 10, n1 // 10, 9 * n2 // 10]
    return (np.array(p, dtype=real) * real(DS)).reshape(3, 2, order="F")


def _run_driver(kind, flavour, outdir):
    exe = LIB / f"srm_driver_{kind}_{flavour}"
    assert exe.exists(), f"{exe} is not built (python -m scarf3d_amd.build)"
    print(f"[case] {exe.name} {N}", flush=True)
    res = subprocess.run([str(exe), *map(str, N), str(outdir)], capture_output=True, text=True, timeout=120)
    stats = {}
    for line in res.stdout.splitlines():
        if line.startswith("stats "):
            parts = line.split()
            stats[parts[1]] = np.array([float(v) for v in parts[2:]])
    return res, stats


@pytest.mark.parametrize("kind", ["c", "cpp"])
@pytest.mark.parametrize("flavour", ["f", "d"])
def test_drivers_are_built_and_report_errors(kind, flavour, tmp_path):
    """The drivers link against the library and run; without a GPU every
    execute reports that there is no device (no CPU fallback)."""
    from scarf3d_amd.hip import Hip

    if Hip().device_count() > 0:
        pytest.skip("a GPU is present: the GPU test runs the drivers")
    res, _ = _run_driver(kind, flavour, tmp_path)
    assert res.returncode == 0, res.stderr
    assert "no HIP device available" in res.stderr


def _lib_field(lib, nd, opts, points=None, method=1):
    assert lib.set_kernel("auto") == 0  # as in the drivers' own processes
    o = lib.opt(method=method, **opts)
    if points is None:
        n = N[:nd]
        lib.cart_initialize(nd, [1] * nd, list(n), DS, 0, list(CL[:nd]), SIGMA, o)
    else:
        z = points[2] if nd == 3 else None
        lib.nocart_initialize(nd, points[0], points[1], z, DS, 0, list(CL[:nd]), SIGMA, o)
    lib.set_nharm(10000)
    field, stats = lib.execute(SEED)
    assert lib.last_error() == "", lib.last_error()
    lib.finalize()
    return field.reshape(-1, order="F"), stats


@pytest.mark.gpu
@pytest.mark.parametrize("flavour", ["f", "d"])
def test_drivers_equal_ctypes_path(lib_f, lib_d, flavour, tmp_path):
    lib = lib_f if flavour == "f" else lib_d
    real = np.float32 if flavour == "f" else np.float64
    poi = _poi(real)
    grid = np.meshgrid(*[np.arange(N[a]) * DS for a in range(3)], indexing="ij")
    pts = [g.reshape(-1, order="F").astype(real) for g in grid]
    grid2 = np.meshgrid(*[np.arange(N[a]) * DS for a in range(2)], indexing="ij")
    pts2 = [g.reshape(-1, order="F").astype(real) for g in grid2]
    expect = {
        "c_cart": _lib_field(lib, 3, dict(hurst=HURST, alpha=30.0, beta=-10.0, npoi=2,
                                              poi=poi.reshape(-1, order="F"), taper=2500.0, mute=500.0)),
        "c_nocart": _lib_field(lib, 3, dict(hurst=HURST), points=pts),
        "c_fim_cart": _lib_field(lib, 3, dict(hurst=HURST), method=0),
        "cpp_2d": _lib_field(lib, 2, dict(hurst=HURST)),
        "cpp_3d": _lib_field(lib, 3, dict(hurst=HURST, alpha=30.0, npoi=2, poi=poi.reshape(-1, order="F"),
                                              taper=2500.0, mute=500.0)),
        "cpp_points": _lib_field(lib, 3, dict(hurst=HURST), points=pts),
        "cpp_fim_3d": _lib_field(lib, 3, dict(hurst=HURST, alpha=30.0, npoi=2, poi=poi.reshape(-1, order="F"),
                                                  taper=2500.0, mute=500.0), method=0),
        "cpp_fim_points2d": _lib_field(lib, 2, dict(hurst=HURST), points=pts2, method=0),
    }
    printed = {}
    for kind in ("c", "cpp"):
        res, stats = _run_driver(kind, flavour, tmp_path)
        assert res.returncode == 0, res.stderr
        assert "error" not in res.stderr, res.stderr
        printed.update({f"{kind}_{k}": v for k, v in stats.items()})
    for name, (field, stats) in expect.items():
        written = np.fromfile(tmp_path / f"{name}.bin", dtype=real)
        assert written.size == field.size, name
        assert np.array_equal(written, field), name
        assert np.array_equal(printed[name][:4].astype(real), stats[:4]), name
    # the slices are the volume's planes
    vol = np.fromfile(tmp_path / "c_cart.bin", dtype=real).reshape(N, order="F")
    zs = np.fromfile(tmp_path / "c_cart_zslice.bin", dtype=real)
    assert np.array_equal(zs, vol[:, :, N[2] // 2 - 1].reshape(-1, order="F"))
    vol3 = np.fromfile(tmp_path / "cpp_3d.bin", dtype=real).reshape(N, order="F")
    xs = np.fromfile(tmp_path / "cpp_3d_xslice.bin", dtype=real)
    assert np.array_equal(xs, vol3[N[0] // 2 - 1, :, :].reshape(-1, order="F"))
