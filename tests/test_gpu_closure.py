"""GPU: the learned SGS closure against the Smagorinsky model a priori.
Training and test data are Gaussian-filtered fields (filter widths 4 dx and
6 dx) of two independently seeded 128^3 turbulent velocity fields, filtered
by this library's spectral kernels, with cd06 gradients of the filtered
fields (see padeops_amd/sgs_closure.py).  The network is trained on the device by the
BF16 MFMA kernels (Adam, batch 65536) on the first field and judged on the
second: it must beat the best constant-coefficient Smagorinsky stress (the
least-squares coefficient, which no dynamic planar constant can exceed in
homogeneous turbulence) in both the correlation coefficient and the relative
L2 error of the deviatoric SGS stress.

The synthetic fields are random-phase (Gaussian) fields: their velocity
gradients have no skewness, so the Smagorinsky stress, odd in the gradient,
is uncorrelated a priori with the exact stress, which is even in it to
leading order (the gradient model).  The gradient model's scores are
recorded beside the others as the reference point the network should
approach."""
import json
import os

import numpy as np
import pytest

import oracle.mlp as M
import padeops_amd as P

pytestmark = pytest.mark.gpu



def _data(seed, cuda):
    import torch
    from padeops_amd import sgs_closure as C
    from padeops_amd.synthetic import turbulence
    n = 128
    u, v, w = turbulence(n, 20.0, seed, torch.float64, cuda)
    dx = 2 * np.pi / n
    return [C.filtered_dataset(u, v, w, n, f * dx, d0=4 * dx) for f in (4, 6)]


def test_closure_beats_smagorinsky_a_priori(cuda, tmp_path):
    import torch
    from padeops_amd import sgs_closure as C
    train = _data(11, cuda)
    X = torch.cat([p[0] for p in train])
    T = torch.cat([p[1] for p in train])
    net = P.SGSNet(65536)
    net.set_params(P.SGSNet.pack(M.init_params(0)))
    last = C.train(net, X, T, steps=600, lr=2e-3)
    test = _data(12, cuda)
    res = {"train_last_batch_loss_normalised": last, "target_scale": net.tscale}
    for (Xt, Tt, aux), kc in zip(test, (4, 6)):
        pred = C.predict(net, Xt)
        smag, coef = C.smagorinsky_fit(aux, Tt)
        cm, em = C.stress_metrics(pred, Tt)
        cs, es = C.stress_metrics(smag, Tt)
        cg, eg = C.stress_metrics(C.gradient_model(Xt), Tt)
        res["delta_%ddx" % kc] = {"mlp_corr": cm, "mlp_rel_err": em, "smag_corr": cs, "smag_rel_err": es,
                                  "smag_ls_coefficient": coef, "gradient_model_corr": cg,
                                  "gradient_model_rel_err": eg}
        print(kc, res["delta_%ddx" % kc], flush=True)
    with open(os.path.join(str(tmp_path), "closure_apriori.json"), "w") as fh:
        json.dump(res, fh, indent=1)
    for kc in (4, 6):
        r = res["delta_%ddx" % kc]
        assert r["mlp_corr"] > r["smag_corr"] + 0.2
        assert r["mlp_rel_err"] < r["smag_rel_err"]
