"""End-to-end on the GPU: fit_gpytorch_mll gradient parity (with the default
priors) and optimize_acqf over qLogNEI (test/test_end_to_end.py pattern)."""
import math

import pytest
import torch

import botorch_amd as B
from botorch_amd import workloads
from botorch_amd.fit import ExactMarginalLogLikelihood, fit_gpytorch_mll
from botorch_amd.optim import optimize_acqf
from oracle import gp_oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _say(msg):
    print(f"[case] {msg}", flush=True)


def test_mll_loss_and_raw_grads_match_oracle_with_priors():
    _say("mll loss + raw grads with default priors")
    w = workloads.gp_small()
    gp = B.SingleTaskGP(w["X"].to(DEV), w["Y"].to(DEV))  # default RBF + priors
    mll = ExactMarginalLogLikelihood(gp.likelihood, gp)
    gp.covar_module.lengthscale = torch.linspace(0.3, 0.8, 6, dtype=torch.float64)
    gp.likelihood.noise = 0.01
    loss, grads = mll.loss_and_grads()
    d = 6
    ls = gp.covar_module.lengthscale.detach().cpu().clone().requires_grad_(True)
    nz = torch.tensor(0.01, dtype=torch.float64, requires_grad=True)
    c = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    y = gp.train_targets.cpu()
    priors = {"lengthscale": (math.sqrt(2) + 0.5 * math.log(d), math.sqrt(3)),
              "noise": (-4.0, 1.0)}
    ref = -O.exact_mll(O.RBF, w["X"], y, ls, torch.tensor(1.0, dtype=torch.float64), nz, c,
                       priors=priors)
    ref.backward()
    assert loss == pytest.approx(ref.item(), rel=1e-9)
    # raw lengthscale / noise are untransformed (GreaterThan(transform=None))
    torch.testing.assert_close(grads[0], ls.grad, rtol=1e-7, atol=1e-10)
    assert grads[1].item() == pytest.approx(nz.grad.item(), rel=1e-7)
    assert grads[2].item() == pytest.approx(c.grad.item(), rel=1e-7, abs=1e-10)


def test_fit_gpytorch_mll_decreases_loss():
    _say("fit_gpytorch_mll")
    w = workloads.gp_small()
    gp = B.SingleTaskGP(w["X"].to(DEV), w["Y"].to(DEV),
                        covar_module=B.ScaleKernel(B.MaternKernel(ard_num_dims=6)))
    mll = ExactMarginalLogLikelihood(gp.likelihood, gp)
    l0, _ = mll.loss_and_grads()
    fit_gpytorch_mll(mll, optimizer_kwargs={"options": {"maxiter": 50}})
    l1, _ = mll.loss_and_grads()
    assert l1 < l0


def test_optimize_acqf_qlognei_end_to_end():
    _say("optimize_acqf qLogNEI q=2")
    w = workloads.gp_small()
    base = B.MaternKernel(ard_num_dims=6)
    base.lengthscale = w["lengthscale"]
    lik = B.GaussianLikelihood()
    lik.noise = w["noise"]
    covar = B.ScaleKernel(base)
    covar.outputscale = w["outputscale"]
    gp = B.SingleTaskGP(w["X"].to(DEV), w["Y"].to(DEV), likelihood=lik, covar_module=covar)
    acq = B.qLogNoisyExpectedImprovement(gp, w["X"].to(DEV),
                                         sampler=B.SobolQMCNormalSampler(torch.Size([128]),
                                                                         seed=0))
    bounds = torch.stack([torch.zeros(6), torch.ones(6)]).to(torch.float64)
    cand, val = optimize_acqf(acq, bounds=bounds, q=2, num_restarts=8, raw_samples=64,
                              options={"seed": 0, "maxiter": 50})
    assert cand.shape == (2, 6)
    assert (cand >= 0).all() and (cand <= 1).all()
    from botorch_amd.sampling import draw_sobol_samples
    raw = draw_sobol_samples(bounds.to(DEV), 64, 2, seed=0)
    with torch.no_grad():
        assert val.item() >= acq(raw).max().item() - 1e-9
