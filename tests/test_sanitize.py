"""Host sanitizer build (CPU): the library's host half (setup, QUADPACK, the
resumable parallel draw, the device record: scarf3d_amd/csrc/srm_host.cpp,
quadpack.cpp) compiled with g++ -fsanitize=address,undefined from the
product's own sources, driven by tests/sanitize/host_draw_check.cpp. The
sanitizers abort on any finding; the draws it prints must equal the
library's (scarf_draw_harmonics) on the same configurations."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from srm_cases import init_product, structured_cfg

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "scarf3d_amd" / "csrc"


@pytest.fixture(scope="module")
def sanitized(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not found")
    exe = tmp_path_factory.mktemp("asan") / "host_draw_check"
    flags = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=all", "-ffp-contract=off", "-fno-fast-math", "-fno-builtin-sin",
             "-fno-builtin-cos", "-fno-builtin-sinf", "-fno-builtin-cosf", "-fno-builtin-sincos",
             "-fno-builtin-sincosf", "-pthread", "-I", str(CSRC)]
    srcs = [ROOT / "tests" / "sanitize" / "host_draw_check.cpp", CSRC / "srm_host.cpp", CSRC / "quadpack.cpp"]
    res = subprocess.run([gxx, *flags, *map(str, srcs), "-o", str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300,
                         env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=1", "PATH": "/usr/bin:/bin"})
    assert run.returncode == 0, run.stderr[-4000:]
    assert "runtime error" not in run.stderr and "AddressSanitizer" not in run.stderr, run.stderr[-4000:]
    cases = {}
    cur = None
    for line in run.stdout.splitlines():
        f = line.split()
        if f[0] == "case":
            cur = int(f[1])
            cases[cur] = dict(uniforms=int(f[2]), kept=int(f[3]), harm=[], used=[])
        else:
            cases[cur]["harm"].append([float(v) for v in f[2:7]])
            cases[cur]["used"].append(int(f[7]))
    return cases


CASES = [
    (0, "f", dict(n=(64, 64, 64), ds=20.0, hurst=0.1, cl=(5000.0,) * 3, nharm=700), 1234),
    (1, "d", dict(n=(64, 64, 64), ds=20.0, dh=8.0, hurst=0.3, cl=(3000.0,) * 3, alpha=30.0, nharm=600), 77),
    (2, "f", dict(n=(40, 40, 40), ds=50.0, acf=1, cl=(300.0,) * 3, hurst=0.0, nharm=500), 5),
]


@pytest.mark.parametrize("case", CASES, ids=[str(c[0]) for c in CASES])
def test_sanitized_host_draw_equals_library(sanitized, lib_f, lib_d, case):
    i, flavour, kw, seed = case
    lib = lib_f if flavour == "f" else lib_d
    cfg = structured_cfg(real=np.float32 if flavour == "f" else np.float64, **kw)
    init_product(lib, cfg)
    harm, used, uniforms = lib.draw_harmonics(seed, kw["nharm"])
    got = sanitized[i]
    assert got["uniforms"] == uniforms
    assert got["kept"] == int(used.sum())
    assert np.array_equal(np.array(got["used"], dtype=np.uint8), used)
    assert np.array_equal(np.array(got["harm"]), harm)
