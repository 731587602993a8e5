"""GPU: scarf_srm_run_noalloc, the allocation-free variant of scarf_srm_run
(no counterpart in the reference; for callers that cannot allocate device
memory). It runs the same realization on the resident table with its
operand tables, statistics partials, statistics and points of interest in
the library's static device scratch array, the tables one chunk of
harmonics at a time. Checked through the C ABI against scarf_srm_run on the
same table, bit for bit, on two calls in a row (the second smaller, with a
rescale and a taper), with the device's free memory unchanged by each call;
and grids beyond the scratch are refused with an error."""
import numpy as np
import pytest

from srm_cases import init_product, structured_cfg

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _noalloc(lib, d_ptr):
    return lib.srm_run_noalloc(d_ptr)  # scarf_srm_run_noalloc through the C ABI


POI = np.array([[1500.0, 800.0, 600.0], [4000.0, 3000.0, 1200.0]]).T
CASES = [
    # 9.2 KB of bf16x3 operand tables per harmonic: 16k harmonics take three chunks of the 64 MiB scratch
    ("three chunks", dict(n=(512, 256, 128), nharm=16000), "mfma_bf16x3"),
    # then a smaller input: one chunk, rescale and a taper around two points of interest, scaled fp16
    ("one chunk, rescale + taper", dict(n=(200, 96, 40), nharm=3000, rescale=1, poi=POI, mute=100.0, taper=800.0),
     "mfma_f16x2"),
]


def test_noalloc_equals_srm_run(lib_f):
    from scarf3d_amd.hip import Hip

    hip = Hip()
    for name, kw, kernel in CASES:
        cfg = structured_cfg(real=np.float32, **kw)
        npts = int(np.prod(kw["n"]))
        _say(name, kernel, kw["n"], kw["nharm"])
        init_product(lib_f, cfg)
        assert lib_f.set_kernel(kernel) == 0
        lib_f.srm_draw(7)
        d_ref = hip.empty(npts, np.float32)
        st_ref = lib_f.srm_run(d_ref.ptr)
        d = hip.empty(npts, np.float32)
        hip.synchronize()
        free0, _ = hip.mem_get_info()
        rc, st = _noalloc(lib_f, d.ptr)
        hip.synchronize()
        free1, _ = hip.mem_get_info()
        assert rc == 0, lib_f.last_error()
        ref, got = d_ref.numpy(), d.numpy()
        print(f"  bit-identical {np.array_equal(ref, got)}, free device memory {free0} -> {free1}", flush=True)
        assert np.array_equal(ref, got)
        assert np.array_equal(st[:4], st_ref[:4])
        assert free1 == free0
        d.free()
        d_ref.free()
        lib_f.finalize()


def test_noalloc_refuses_what_does_not_fit(lib_f):
    """A 2D grid 32768 x 32768 needs ~84 MB of operand tables for 128
    harmonics, beyond the 64 MiB scratch: refused before any launch (the
    field buffer is allocated at full size all the same). A direct kernel is
    refused too."""
    from scarf3d_amd.hip import Hip

    hip = Hip()
    n = (32768, 32768, 1)
    cfg = structured_cfg(n=n, ds=50.0, cl=(20000.0, 20000.0, 0.0), nharm=128, real=np.float32)
    _say("grid beyond the scratch", n)
    init_product(lib_f, cfg)
    assert lib_f.set_kernel("mfma_bf16x3") == 0
    lib_f.srm_draw(3)
    d = hip.empty(int(np.prod(n)), np.float32)
    rc, _ = _noalloc(lib_f, d.ptr)
    assert rc != 0 and "no-alloc scratch" in lib_f.last_error()
    lib_f.finalize()
    cfg = structured_cfg(n=(64, 64, 16), nharm=256, real=np.float32)
    init_product(lib_f, cfg)
    assert lib_f.set_kernel("direct_f32") == 0
    lib_f.srm_draw(3)
    rc, _ = _noalloc(lib_f, d.ptr)
    assert rc != 0 and "MFMA" in lib_f.last_error()
    d.free()
