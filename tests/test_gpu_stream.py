"""GPU: streamed realizations. On the MFMA kernels of a structured grid,
scarf_execute and scarf_execute_device draw the harmonics in chunks and
queue each chunk's record upload, operand tables and field pass while the
host draws the next; chunks start at multiples of the kernels' 128-harmonic
promotion period and carry the FP32 running sums from launch to launch, so
the field must equal the whole-table path (stream_draw = 0) bit for bit,
and the whole-table path is checked against the FP64 oracle elsewhere
(test_gpu_parity.py)."""
import numpy as np
import pytest

from srm_cases import init_product, oracle_field, structured_cfg

pytestmark = pytest.mark.gpu

KERNELS = ["mfma_bf16x3", "mfma_f16x2", "mfma_bf16x2"]


def _say(*a):
    print("[case]", *a, flush=True)


def _field(lib, cfg, seed, kernel, stream=True, chunk=0):
    init_product(lib, cfg)
    assert lib.set_kernel(kernel) == 0
    assert lib.set_option("stream_draw", 1 if stream else 0) == 0
    assert lib.set_option("stream_chunk", chunk) == 0
    field, stats = lib.execute(seed)
    assert lib.last_error() == "", lib.last_error()
    lib.finalize()
    return field, stats


STREAM_CASES = [
    # 8000 harmonics: chunks of 1024, 2048 and 4928 kept harmonics
    ("three_chunks", dict(n=(128, 64, 24), nharm=8000)),
    # dh < ds: harmonics above kc are dropped inside and between chunks
    ("dropped", dict(n=(96, 40, 17), ds=20.0, dh=8.0, hurst=0.3, cl=(3000.0,) * 3, nharm=3000)),
    # rotation, a block offset and ragged x/j/k tiles
    ("rotated_offset", dict(n=(150, 70, 11), fs=(5, 3, 9), alpha=30.0, beta=-10.0, gamma=5.0, nharm=2600)),
    ("grid_2d", dict(n=(200, 160, 1), cl=(2000.0, 2000.0, 0.0), nharm=2400)),
]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("name,kw", STREAM_CASES, ids=[c[0] for c in STREAM_CASES])
def test_streamed_equals_whole_table(lib_f, name, kw, kernel):
    cfg = structured_cfg(real=np.float32, **kw)
    _say(name, kernel, kw)
    whole, st_w = _field(lib_f, cfg, 99, kernel, stream=False)
    streamed, st_s = _field(lib_f, cfg, 99, kernel, stream=True)
    assert np.array_equal(whole, streamed)
    assert np.array_equal(st_w[:4], st_s[:4])
    # fixed 128-harmonic chunks: the most launches, every one a promotion period
    small, _ = _field(lib_f, cfg, 99, kernel, stream=True, chunk=128)
    assert np.array_equal(whole, small)


def test_streamed_field_matches_oracle(oracle, lib_f):
    """A streamed bf16x3 realization of 8000 harmonics in three chunks
    against the FP64 oracle on the same draw (1e-5 sigma)."""
    cfg = structured_cfg(n=(64, 32, 16), nharm=8000, real=np.float32)
    _say("streamed bf16x3 vs oracle", cfg["fe"])
    ref, _ = oracle_field(oracle, cfg, 5, "f")
    field, _ = _field(lib_f, cfg, 5, "mfma_bf16x3")
    err = np.abs(field.astype(np.float64) - ref.reshape(field.shape, order="F")).max() / cfg["sigma"]
    print(f"  max|err|/sigma = {err:.3e}", flush=True)
    assert err <= 1e-5


def test_streamed_draw_ending_in_an_empty_chunk(lib_f):
    """When every harmonic after a chunk cut is dropped (u > kc), the last
    chunk brings no harmonic; it then runs one block of zero-amplitude
    harmonics, which add exact zeros. The configuration is found on the host:
    nharm ends right after a cut at a multiple of 128 kept harmonics, on a
    dropped harmonic."""
    base = dict(n=(96, 48, 12), ds=40.0, dh=10.0, hurst=0.3, cl=(3000.0,) * 3)
    probe = structured_cfg(real=np.float32, nharm=3000, **base)
    found = None
    for seed in range(17, 60):
        init_product(lib_f, probe)
        _, used, _ = lib_f.draw_harmonics(seed, 3000)
        lib_f.finalize()
        kept = np.cumsum(used)
        cuts = [p for p in range(1, 2999) if kept[p - 1] % 128 == 0 and used[p - 1] and not used[p]]
        if cuts:
            found = (seed, cuts[0] + 1, int(kept[cuts[0]]))  # the cut's harmonic, then one dropped harmonic
            break
    assert found, "no draw position ends a chunk on a dropped harmonic"
    seed, nharm, nkept = found
    cfg = structured_cfg(real=np.float32, nharm=nharm, **base)
    _say("empty last chunk: seed", seed, "nharm", nharm, "kept", nkept)
    whole, _ = _field(lib_f, cfg, seed, "mfma_bf16x3", stream=False)
    streamed, _ = _field(lib_f, cfg, seed, "mfma_bf16x3", stream=True, chunk=128)
    assert np.array_equal(whole, streamed)


def test_streamed_fp16_restarts_on_a_rescaled_table(lib_f, monkeypatch):
    """The scaled-fp16 scheme needs |Z Y| < 8 in its table; a chunk with a
    larger amplitude (the limit is lowered here so that every chunk has one)
    makes the streamed path finish the draw and run the whole table with its
    power-of-two scale: the field equals the whole-table path's."""
    monkeypatch.setenv("SCARF_FP16_AMP_LIMIT", "0.125")
    cfg = structured_cfg(n=(96, 40, 12), nharm=3000, real=np.float32)
    _say("fp16 restart")
    whole, _ = _field(lib_f, cfg, 31, "mfma_f16x2", stream=False)
    streamed, _ = _field(lib_f, cfg, 31, "mfma_f16x2", stream=True)
    assert np.array_equal(whole, streamed)


def test_streamed_host_copy_in_slabs(lib_f):
    """A 72 MB field (above the 64 MiB slab threshold) of two chunks: the
    final chunk runs in z slabs, each staged to the host array while the
    next is computed; the host array equals the whole-table device field."""
    from scarf3d_amd.hip import Hip

    hip = Hip()
    n = (300, 300, 200)
    cfg = structured_cfg(n=n, nharm=2500, real=np.float32)
    _say("streamed host copy of a 72 MB field", n)
    field, _ = _field(lib_f, cfg, 11, "mfma_bf16x3", stream=True)
    d = hip.empty(int(np.prod(n)), np.float32)
    init_product(lib_f, cfg)
    assert lib_f.set_kernel("mfma_bf16x3") == 0
    assert lib_f.set_option("stream_draw", 0) == 0
    lib_f.execute_device(11, d.ptr)
    lib_f.finalize()
    assert np.array_equal(field.reshape(-1, order="F"), d.numpy())
    d.free()
