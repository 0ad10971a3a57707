"""GPU: the learned closure's whole pipeline at test size (padeops_amd/
closure_dns.py): a forced 64^3 FP64 DNS on the device, its snapshots
filtered to 16^3 and 32^3 LES grids, the network trained on the device
(features, loss and BF16 MFMA backward of this library), then

  * a priori on held-out snapshots: the learned stress correlates better
    with the exact subgrid stress, and misses it by less, than the dynamic
    Smagorinsky stress of the standard dynamic procedure on the same field;
  * a posteriori: a 32^3 FP32+refinement LES runs with the learned closure
    (dynamic fallback where its dissipation is negative) and stays finite.

The full-size run (256^3 DNS, 200 snapshots, 64^3 / 128^3 grids, 128^3 LES)
is `python -m padeops_amd.closure_dns`."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_closure_pipeline_small(cuda):
    from padeops_amd import closure_dns as D
    res = D.run(N=64, les_grids=(16, 32), Re=100.0, nsnap=24, every=4, spinup_time=1.0, ntest=8, ntest_keep=4,
                train_steps=400, sub_points=4096, les_grid=32, les_T=1.0, log=lambda m: print(m, flush=True))
    for m in ("16", "32"):
        a = res["apriori"][m]
        print(m, a)
        assert a["mlp_corr"] > a["dyn_corr"]
        assert a["mlp_err"] < a["dyn_err"]
        assert np.isfinite(a["diss_mlp"])
    post = res["aposteriori"]
    assert post["learned+dynamic"]["finite"]
    assert post["dynamic"]["finite"]
    hist = res["train_loss_scaled"]
    assert hist[-1][1] < hist[0][1]
