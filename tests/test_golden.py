"""The oracle reproduces the committed gp_small fixture (CPU), and the HIP path
matches the same fixture (GPU)."""
import json
import os

import pytest
import torch

from botorch_amd import workloads
from oracle import gp_oracle as O

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gp_small.json")


def _fixture():
    with open(FIX) as f:
        return json.load(f)


def test_oracle_reproduces_fixture():
    fx = _fixture()
    w = workloads.gp_small()
    orc = O.OracleGP(w["X"], w["Y"], O.MATERN52, w["lengthscale"], 1.0, w["noise"])
    Xb, idx = O.prune_inferior_points(orc, w["X"], num_samples=fx["prune_samples"],
                                      seed=fx["prune_seed"])
    assert idx.tolist() == fx["baseline_idx"]
    acq = O.OracleQLogNEI(orc, Xb, num_samples=w["S"], seed=fx["sampler_seed"])
    v, g = acq.value_and_grad(w["Xq"])
    torch.testing.assert_close(v, torch.tensor(fx["qlognei_value"], dtype=torch.float64),
                               rtol=1e-10, atol=1e-12)
    torch.testing.assert_close(g, torch.tensor(fx["qlognei_grad"], dtype=torch.float64),
                               rtol=1e-9, atol=1e-11)
    m, c = orc.posterior(torch.tensor(fx["posterior_X"], dtype=torch.float64))
    torch.testing.assert_close(m, torch.tensor(fx["posterior_mean"], dtype=torch.float64))
    torch.testing.assert_close(c, torch.tensor(fx["posterior_covar"], dtype=torch.float64))


@pytest.mark.gpu
def test_hip_matches_fixture():
    import botorch_amd as B
    print("[case] golden gp_small on HIP", flush=True)
    fx = _fixture()
    w = workloads.gp_small()
    dev = "cuda"
    base = B.MaternKernel(ard_num_dims=6)
    base.lengthscale = w["lengthscale"]
    lik = B.GaussianLikelihood()
    lik.noise = w["noise"]
    covar = B.ScaleKernel(base)
    covar.outputscale = w["outputscale"]
    gp = B.SingleTaskGP(w["X"].to(dev), w["Y"].to(dev), likelihood=lik, covar_module=covar)
    Xb = w["X"][fx["baseline_idx"]].to(dev)
    acq = B.qLogNoisyExpectedImprovement(
        gp, Xb, sampler=B.SobolQMCNormalSampler(torch.Size([w["S"]]), seed=fx["sampler_seed"]),
        prune_baseline=False)
    v, g = acq.value_and_grad(w["Xq"].to(dev))
    vr = torch.tensor(fx["qlognei_value"], dtype=torch.float64)
    gr = torch.tensor(fx["qlognei_grad"], dtype=torch.float64)
    assert ((v.cpu() - vr).abs().max() / vr.abs().max()).item() < 1e-6
    assert ((g.cpu() - gr).abs().max() / gr.abs().max()).item() < 1e-6
    post = gp.posterior(torch.tensor(fx["posterior_X"], dtype=torch.float64, device=dev))
    torch.testing.assert_close(post.mean.squeeze(-1).cpu(),
                               torch.tensor(fx["posterior_mean"], dtype=torch.float64),
                               rtol=1e-8, atol=1e-10)
