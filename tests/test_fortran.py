"""The Fortran module m_scarflib (scarf3d_amd/fortran) over the C ABI,
through the test driver tests/fortran/srm_driver.f90 built by
scarf3d_amd.build with amdflang."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from srm_cases import init_product, structured_cfg

ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "scarf3d_amd" / "lib" / "srm_driver_f"

needs_driver = pytest.mark.skipif(not DRIVER.exists(), reason="Fortran driver not built (amdflang missing)")

KEYS = ["nd", "structured", "fs1", "fs2", "fs3", "fe1", "fe2", "fe3", "ds", "dh", "acf", "cl1", "cl2", "cl3", "sigma",
        "hurst", "nc1", "nc2", "nc3", "fc1", "fc2", "fc3", "alpha", "beta", "gamma", "mute", "taper", "rescale", "pad",
        "method", "npoi", "npts"]


@needs_driver
def test_fortran_optional_arguments_keep_fortran_semantics():
    """mute = 0 is a present argument in Fortran (src/scarflib.f90:166) and is
    kept; nc alone overrides the near corner; fc keeps its default."""
    out = subprocess.run([str(DRIVER), "config"], capture_output=True, text=True, timeout=60, check=True).stdout
    c = dict(zip(KEYS, [float(v) for v in out.split()]))
    assert c["method"] == 1 and c["mute"] == 0.0 and c["taper"] == 500.0 and c["rescale"] == 1
    assert c["hurst"] == pytest.approx(0.25) and c["alpha"] == 30.0
    assert (c["nc1"], c["nc2"], c["nc3"]) == (-10.0, -20.0, -30.0)
    assert (c["fc1"], c["fc2"], c["fc3"]) == (39 * 50.0, 35 * 50.0, 19 * 50.0)
    assert c["dh"] == 50.0 and c["ds"] == 50.0


@needs_driver
@pytest.mark.gpu
def test_fortran_execute_equals_c_api(lib_f, tmp_path):
    out = tmp_path / "field.bin"
    print("[case] fortran driver run", flush=True)
    res = subprocess.run([str(DRIVER), "run", str(out)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    stats_f = np.array([float(v) for v in res.stdout.split()])
    field_f = np.fromfile(out, dtype=np.float32).reshape((40, 36, 20), order="F")
    poi = np.array([[300.0, 400.0, 200.0], [900.0, 500.0, 600.0]]).T
    cfg = structured_cfg(n=(40, 36, 20), nharm=384, poi=poi, mute=100.0, taper=250.0, rescale=1)
    init_product(lib_f, cfg)
    field_c, stats_c = lib_f.execute(1234)
    assert np.array_equal(field_f, field_c)
    assert np.array_equal(stats_f[:4], stats_c[:4].astype(np.float64))


@needs_driver
@pytest.mark.gpu
def test_fortran_fim_equals_c_api(lib_f, tmp_path):
    """FIM through the Fortran module (method omitted: its default 0), with
    pad, rescale and one tapered point of interest, equals the C ABI's field
    and statistics bit for bit."""
    out = tmp_path / "fim.bin"
    print("[case] fortran driver fim", flush=True)
    res = subprocess.run([str(DRIVER), "fim", str(out)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    stats_f = np.array([float(v) for v in res.stdout.split()])
    field_f = np.fromfile(out, dtype=np.float32).reshape((40, 36, 20), order="F")
    poi = np.array([[300.0, 400.0, 200.0]]).T
    cfg = structured_cfg(n=(40, 36, 20), poi=poi, mute=100.0, taper=250.0, rescale=1)
    cfg["method"], cfg["pad"] = 0, 1
    init_product(lib_f, cfg)
    field_c, stats_c = lib_f.execute(1234)
    assert lib_f.last_error() == ""
    assert np.array_equal(field_f, field_c)
    assert np.array_equal(stats_f[:4], stats_c[:4].astype(np.float64))
