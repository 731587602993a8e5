"""GPU: the reference's domain decomposition on one GPU, the stand-in for
running the path on several GPUs.

Each MPI rank of the reference initializes with its own block fs:fe of the
global grid, redraws the same harmonics from the same seed
(src/scarflib_srm.f90:627-628, :731-778), takes the global extents from an
allreduce (:641-642; passed here as nc/fc) and computes its block; the
blocks' statistics are merged with Chan's formula (:887-895,
src/scarflib_common.f90:823-875). Eight z-slab blocks of the 512^3 x 20k
headline grid, each through its own initialize / execute_device / finalize,
must give the whole-grid field, and the oracle's Chan merge of their
(n, mean, variance) the whole grid's stats(3:4). One block (512 x 512 x 64)
is one GPU's share of the grid at eight GPUs: its time is printed."""
import ctypes as C
import time

import numpy as np
import pytest

from oracle_lib import _DP
from srm_cases import init_product, structured_cfg

pytestmark = pytest.mark.gpu

N = (512, 512, 512)
KW = dict(ds=20.0, acf=0, hurst=0.1, cl=(5000.0,) * 3, sigma=0.05, nharm=20000, real=np.float32)


@pytest.mark.parametrize("kernel", ["mfma_bf16x3", "direct_f32"])
def test_z_slab_blocks_equal_the_whole_grid(oracle, lib_f, kernel):
    from scarf3d_amd.hip import Hip

    hip = Hip()
    seed = 1234
    whole_cfg = structured_cfg(n=N, **KW)
    npts = int(np.prod(N))
    d_whole = hip.empty(npts, np.float32)
    d_blocks = hip.empty(npts, np.float32)
    print(f"[case] whole 512^3 grid, {kernel}", flush=True)
    init_product(lib_f, whole_cfg)
    assert lib_f.set_kernel(kernel) == 0
    st_whole = lib_f.execute_device(seed, d_whole.ptr)
    lib_f.finalize()
    nb, plane = 8, N[0] * N[1]
    depth = N[2] // nb
    var, avg, cnt, block_ms = [], [], [], []
    for b in range(nb):
        cfg = structured_cfg(n=(N[0], N[1], depth), fs=(1, 1, 1 + b * depth), **KW)
        cfg = dict(cfg, nc=whole_cfg["nc"], fc=whole_cfg["fc"])  # the global model box
        print(f"[case] block {b}: z {1 + b * depth}..{(b + 1) * depth}", flush=True)
        init_product(lib_f, cfg)
        assert lib_f.set_kernel(kernel) == 0
        view = d_blocks.ptr + 4 * b * depth * plane
        st = lib_f.execute_device(seed, view)
        if b == 0:  # one GPU's share at eight GPUs, warm
            t0 = time.perf_counter()
            st = lib_f.execute_device(seed, view)
            block_ms.append((time.perf_counter() - t0) * 1e3)
            block_ms.append(float(st[5]) * 1e3)
        assert lib_f.last_error() == "", lib_f.last_error()
        lib_f.finalize()
        var.append(float(st[2]) ** 2)
        avg.append(float(st[3]))
        cnt.append(depth * plane)
    whole = d_whole.numpy()
    blocks = d_blocks.numpy()
    d_whole.free()
    d_blocks.free()
    same = np.array_equal(whole, blocks)
    err = float(np.abs(whole.astype(np.float64) - blocks).max()) / KW["sigma"]
    print(f"  blocks vs whole: bit-identical {same}, max|diff|/sigma = {err:.3e}; one 512x512x64 block "
          f"{block_ms[0]:.1f} ms per execute_device ({block_ms[1]:.1f} ms device)", flush=True)
    assert err <= 1e-7
    # Chan merge of the blocks' (n, mean, variance), as the ranks' allgather does
    v, m = C.c_double(), C.c_double()
    oracle.L.orc_parallel_variance(np.array(var).ctypes.data_as(_DP), np.array(avg).ctypes.data_as(_DP),
                                   np.array(cnt, dtype=np.int64).ctypes.data_as(C.POINTER(C.c_long)), nb,
                                   C.byref(v), C.byref(m))
    print(f"  merged sigma_d {np.sqrt(v.value):.7e} mean {m.value:.4e}; whole {st_whole[2]:.7e} {st_whole[3]:.4e}",
          flush=True)
    assert np.sqrt(v.value) == pytest.approx(float(st_whole[2]), rel=1e-6)
    assert m.value == pytest.approx(float(st_whole[3]), rel=1e-4, abs=1e-7 * KW["sigma"])
