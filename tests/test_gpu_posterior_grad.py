"""SingleTaskGP.posterior is differentiable in X, as the reference's is
(models/gpytorch.py:544-610): the posterior adjoint (bta_gp_posterior_backward),
the Cholesky adjoint of the posterior's scale_tril (bta_chol_small_backward) and
the z L^T product's adjoint against torch autograd through the FP64 oracle, and
the posterior mean's gradient against central finite differences."""
import pytest
import torch

import botorch_amd as B
from botorch_amd import workloads
from oracle import gp_oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _rel(a, r):
    return ((a.detach().cpu() - r).abs().max() / r.abs().max()).item()


def _pair(kind=O.MATERN52):
    w = workloads.gp_small()
    d = w["X"].shape[-1]
    base = B.MaternKernel(ard_num_dims=d) if kind == O.MATERN52 else B.RBFKernel(ard_num_dims=d)
    base.lengthscale = w["lengthscale"]
    covar = B.ScaleKernel(base)
    covar.outputscale = 1.3
    lik = B.GaussianLikelihood()
    lik.noise = 1e-3
    gp = B.SingleTaskGP(w["X"].to(DEV), w["Y"].to(DEV), likelihood=lik, covar_module=covar)
    orc = O.OracleGP(w["X"], w["Y"], kind, w["lengthscale"], 1.3, 1e-3)
    return gp, orc


@pytest.mark.parametrize("kind,batch,mq", [(O.MATERN52, (3,), 5), (O.RBF, (2, 2), 4),
                                           (O.MATERN52, (1,), 70)])
def test_posterior_moments_gradient_matches_oracle_autograd(kind, batch, mq):
    _say(f"posterior adjoint kind={kind} batch={batch} mq={mq}")
    gp, orc = _pair(kind)
    g = torch.Generator().manual_seed(3 + mq)
    X = torch.rand(*batch, mq, 6, generator=g, dtype=torch.float64)
    wm = torch.randn(*batch, mq, generator=g, dtype=torch.float64)
    wc = torch.randn(*batch, mq, mq, generator=g, dtype=torch.float64)
    Xd = X.to(DEV).requires_grad_(True)
    post = gp.posterior(Xd)
    loss = (post.mean.squeeze(-1) * wm.to(DEV)).sum() + (post.covariance_matrix * wc.to(DEV)).sum()
    loss.backward()
    Xr = X.clone().requires_grad_(True)
    m_r, c_r = orc.posterior(Xr)
    ((m_r * wm).sum() + (c_r * wc).sum()).backward()
    err = _rel(Xd.grad, Xr.grad)
    print(f"dX rel err {err:.2e}", flush=True)
    assert err < 1e-9


def test_posterior_samples_gradient_matches_oracle_autograd():
    _say("rsample_from_base_samples gradient through Cholesky and z L^T")
    gp, orc = _pair()
    g = torch.Generator().manual_seed(11)
    X = torch.rand(4, 3, 6, generator=g, dtype=torch.float64)
    Z = torch.randn(32, 4, 3, generator=g, dtype=torch.float64)
    W = torch.randn(32, 4, 3, generator=g, dtype=torch.float64)
    Xd = X.to(DEV).requires_grad_(True)
    s = gp.posterior(Xd).rsample_from_base_samples(torch.Size([32]), Z.to(DEV))
    (s.squeeze(-1) * W.to(DEV)).sum().backward()
    Xr = X.clone().requires_grad_(True)
    m_r, c_r = orc.posterior(Xr)
    L_r = torch.linalg.cholesky(c_r)
    s_r = m_r.unsqueeze(0) + torch.einsum("bij,sbj->sbi", L_r, Z)
    (s_r * W).sum().backward()
    err = _rel(Xd.grad, Xr.grad)
    print(f"dX rel err {err:.2e}", flush=True)
    assert err < 1e-8


def test_posterior_mean_gradient_matches_central_fd():
    _say("posterior mean gradient vs central FD")
    gp, _ = _pair()
    X = torch.rand(2, 3, 6, generator=torch.Generator().manual_seed(5), dtype=torch.float64)
    Xd = X.to(DEV).requires_grad_(True)
    gp.posterior(Xd).mean.sum().backward()
    h = 1e-6
    fd = torch.zeros_like(X)
    with torch.no_grad():
        for idx in range(X.numel()):
            e = torch.zeros(X.numel(), dtype=torch.float64)
            e[idx] = h
            e = e.view_as(X).to(DEV)
            fd.view(-1)[idx] = ((gp.posterior(X.to(DEV) + e).mean.sum()
                                 - gp.posterior(X.to(DEV) - e).mean.sum()) / (2 * h)).item()
    assert _rel(Xd.grad, fd) < 1e-6
