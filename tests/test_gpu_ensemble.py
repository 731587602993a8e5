"""Ensembles: one persistent grouped-GEMM launch (replayed from a hipGraph)
must give every member exactly the field and statistics of the same
realization computed alone with the same MFMA kernel, and the FP64 oracle's
field to 1e-5 sigma."""
import numpy as np
import pytest

from srm_cases import init_product, oracle_field, structured_cfg

pytestmark = pytest.mark.gpu

MEMBERS = [
    dict(fe=(96, 80, 40), ds=50.0, acf=0, cl=(2000.0, 2000.0, 2000.0), sigma=0.05, hurst=0.25, nharm=512, seed=1),
    dict(fe=(160, 64, 24), ds=25.0, acf=1, cl=(300.0, 300.0, 100.0), sigma=0.1, nharm=320, seed=2),
    dict(fe=(64, 48, 33), ds=100.0, acf=0, cl=(4000.0, 3000.0, 2000.0), sigma=0.08, hurst=0.1, nharm=768, seed=3,
         alpha=20.0, rescale=1),
    dict(fe=(200, 40, 8), ds=20.0, acf=0, cl=(1500.0, 1500.0, 1500.0), sigma=0.05, hurst=0.4, nharm=256, seed=4),
    dict(fe=(20, 36, 12), ds=50.0, acf=0, cl=(1000.0, 1000.0, 1000.0), sigma=0.05, hurst=0.3, nharm=192, seed=5),
    dict(fe=(180, 24, 8), ds=40.0, acf=1, cl=(600.0, 600.0, 200.0), sigma=0.1, nharm=256, seed=6),
    dict(fe=(270, 16, 8), ds=30.0, acf=0, cl=(2500.0, 2500.0, 2500.0), sigma=0.05, hurst=0.2, nharm=256, seed=7),
    dict(fe=(300, 16, 12), ds=30.0, acf=0, cl=(1200.0, 1200.0, 1200.0), sigma=0.05, hurst=0.15, nharm=224, seed=8),
]
# x tiles (points): 96 | 96 + 64 | 64 | 128 + 96 (ragged) | one 64 for 20 | 128 + 64 (ragged) |
# 128 + 96 + 64 (ragged) | 128 + 128 + 64 (ragged)


def _cfg(m, real=np.float32):
    n = m["fe"]
    return structured_cfg(n=n, ds=m["ds"], acf=m["acf"], cl=m["cl"], sigma=m["sigma"], hurst=m.get("hurst", 0.0),
                          nharm=m["nharm"], alpha=m.get("alpha", 0.0), rescale=m.get("rescale", 0), real=real)


@pytest.mark.parametrize("kernel", ["mfma_bf16x3", "mfma_f16x2"])
def test_ensemble_members_equal_single_runs(oracle, lib_f, kernel):
    from scarf3d_amd.hip import Hip

    hip = Hip()
    bufs = [hip.empty(int(np.prod(m["fe"])), np.float32) for m in MEMBERS]
    print("[case] ensemble", kernel, flush=True)
    stats = lib_f.ensemble_execute(MEMBERS, [b.ptr for b in bufs], kernel)
    fields = [b.numpy() for b in bufs]
    # replay of the captured graph with the same shapes gives the same result
    stats2 = lib_f.ensemble_execute(MEMBERS, [b.ptr for b in bufs], kernel)
    assert np.array_equal(stats[:, :4], stats2[:, :4])  # columns 5-6 are host and device seconds
    for i, m in enumerate(MEMBERS):
        cfg = _cfg(m)
        init_product(lib_f, cfg)
        lib_f.set_kernel(kernel)
        single, st = lib_f.execute(m["seed"])
        lib_f.finalize()
        assert np.array_equal(fields[i], single.reshape(-1, order="F")), i
        assert np.array_equal(stats[i, :4], st[:4]), i
        if not m.get("rescale"):
            ref, _ = oracle_field(oracle, cfg, m["seed"], "f")
            err = np.abs(fields[i].astype(np.float64) - ref.reshape(-1, order="F")).max() / m["sigma"]
            print(f"  member {i}: max|err|/sigma = {err:.2e}", flush=True)
            assert err <= 1e-5
    for b in bufs:
        b.free()
