"""The standalone program (scarf3d_amd/standalone/scarf3d_main.cpp): the
keyword input file of standalone/main.f90:121-232, its checks and messages
(:253-319), and the field it writes (random_field.bin, :468-472)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from srm_cases import init_product, structured_cfg

LIB = Path(__file__).resolve().parents[1] / "scarf3d_amd" / "lib"
EXE = LIB / "scarf3d_standalone_f"

GOOD = """# dimension
3
# npts
40 36 20
# ds
50.0
# cl
2000 2000 2000
# sigma
0.05
# acf
0
# hurst
0.25
# seed
1234
# method
1
"""


def _run(tmp_path, text, timeout=120):
    inp = tmp_path / "input.txt"
    inp.write_text(text)
    return subprocess.run([str(EXE), str(inp)], capture_output=True, text=True, cwd=tmp_path, timeout=timeout)


def test_standalone_requires_an_input_file(tmp_path):
    r = subprocess.run([str(EXE)], capture_output=True, text=True, cwd=tmp_path, timeout=60)
    assert r.returncode != 0 and "Please specify an input file!" in r.stdout


@pytest.mark.parametrize("drop,replace,msg", [
    ("# seed\n1234\n", "", "seed number (any integer not equal zero) not specified"),
    ("# dimension\n3\n", "", "dimension must be equal to 2 or 3"),
    ("# acf\n0\n", "# acf\n2\n", "accepted ACF are Von Karman (0) and Gaussian (1)"),
    ("# sigma\n0.05\n", "# sigma\n-1\n", "standard deviation (sigma) must be larger than zero"),
    ("# ds\n50.0\n", "# ds\n0\n", "grid-step must be larger than zero"),
    ("# npts\n40 36 20\n", "# npts\n40 0 20\n", "number of points must be larger than zero"),
    ("# method\n1\n", "# method\n1\n# poi\n-5 10 10\n", "POI must be inside the model"),
    ("# method\n1\n", "# method\n1\n# mute\n-1\n", "mute parameter can be only positive"),
    ("# method\n1\n", "# method\n1\n# rescale\n3\n", "rescale parameter can be only off (0) or on (1)"),
])
def test_standalone_input_checks(tmp_path, drop, replace, msg):
    text = GOOD.replace(drop, replace)
    r = _run(tmp_path, text)
    assert r.returncode != 0
    assert msg in r.stdout


def test_standalone_reads_every_descriptor_and_prints_summary(tmp_path):
    """On a host without a GPU the run stops at execute, after the summary."""
    text = GOOD + "# poi\n500 500 500\n# poi\n900 800 700\n# mute\n0\n# taper\n200\n# alpha\n10\n# nwriters\n3\n"
    r = _run(tmp_path, text)
    out = r.stdout
    assert "Summary of input parameters" in out
    assert "Number of POIs" in out and "           2" in out
    assert "Tapering radius" in out and "200.000" in out
    assert "Alpha angle" in out and "10.000" in out


@pytest.mark.gpu
def test_standalone_field_equals_library(tmp_path, lib_f):
    text = GOOD + "# poi\n500 500 500\n# mute\n0\n# taper\n300\n"
    print("[case] standalone 40x36x20", flush=True)
    r = _run(tmp_path, text)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Program completed successfully" in r.stdout
    field = np.fromfile(tmp_path / "random_field.bin", dtype=np.float32)
    cfg = structured_cfg(n=(40, 36, 20), ds=50.0, cl=(2000.0,) * 3, sigma=0.05, hurst=0.25, nharm=10000,
                         poi=np.array([[500.0], [500.0], [500.0]]), mute=0.0, taper=300.0, real=np.float32)
    init_product(lib_f, cfg)
    assert lib_f.set_kernel("auto") == 0  # the standalone's choice (the setting outlives finalize)
    assert lib_f.set_option("mute", 0.0) == 0  # present, as the standalone passes it (a C opt drops mute <= 0)
    ref, _ = lib_f.execute(1234)
    assert np.array_equal(field, ref.reshape(-1, order="F"))


@pytest.mark.gpu
def test_standalone_fim_field_equals_library(tmp_path, lib_f):
    """method 0 (FIM) with a tapered point of interest: the program's field is
    the library's FIM field on the same options. The input's "# pad 1" is read
    and passed to no call, as in standalone/main.f90:204-205, :426-427, so the
    field is the unpadded one."""
    from scarf3d_amd.workloads import init_library

    text = GOOD.replace("# method\n1\n", "# method\n0\n# pad\n1\n") + "# poi\n500 500 500\n# mute\n0\n# taper\n300\n"
    print("[case] standalone FIM 40x36x20", flush=True)
    r = _run(tmp_path, text)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Program completed successfully" in r.stdout
    field = np.fromfile(tmp_path / "random_field.bin", dtype=np.float32)
    cfg = structured_cfg(n=(40, 36, 20), ds=50.0, cl=(2000.0,) * 3, sigma=0.05, hurst=0.25,
                         poi=np.array([[500.0], [500.0], [500.0]]), mute=0.0, taper=300.0, real=np.float32)
    cfg["method"], cfg["pad"] = 0, 0
    init_library(lib_f, cfg)
    assert lib_f.set_option("mute", 0.0) == 0  # present, as the standalone passes it
    ref, _ = lib_f.execute(1234)
    assert lib_f.last_error() == ""
    assert np.array_equal(field, ref.reshape(-1, order="F"))
