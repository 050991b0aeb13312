"""The large and the batched configurations at their full sizes.

Large single GP (n = 16384, d = 20, 10% of the points clustered, noise 1e-4,
q = 8, 1024 restarts, 256 samples): the FP64 HIP path against the FP64 oracle
on a slice of the restarts (value and gradient 1e-6 relative), and the mixed
path against the FP64 HIP path on all restarts (the mixed-mode tolerance is
stated in the test).  Batched models (n log-uniform in [128, 4096], d in
{4..20}, q = 2, 256 samples): a spread of the 64 models, each against the
oracle on a slice of its restarts.
"""
import pytest
import torch

import botorch_amd as B
from botorch_amd import workloads
from oracle import gp_oracle as O

import mixed_contract

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _rel(a, r):
    return ((a - r).abs().max() / r.abs().max()).item()


def _build(w, precision, dev=DEV):
    d = w["X"].shape[-1]
    base = B.MaternKernel(ard_num_dims=d)
    base.lengthscale = w["lengthscale"]
    covar = B.ScaleKernel(base)
    covar.outputscale = w["outputscale"]
    lik = B.GaussianLikelihood()
    lik.noise = w["noise"]
    gp = B.SingleTaskGP(w["X"].to(dev), w["Y"].to(dev), likelihood=lik, covar_module=covar)
    with torch.no_grad():
        Xb = B.prune_inferior_points(
            gp, w["X"].to(dev), sampler=B.SobolQMCNormalSampler(torch.Size([2048]), seed=123))
    acqs = {p: B.qLogNoisyExpectedImprovement(
        gp, Xb, sampler=B.SobolQMCNormalSampler(torch.Size([w["S"]]), seed=0),
        prune_baseline=False, precision=p) for p in precision}
    return acqs, Xb


def test_large_clustered_gp_n16384():
    _say("large GP n=16384 d=20 clustered, q=8 b=1024 S=256")
    w = workloads.gp_large()
    acqs, Xb = _build(w, ["fp64", "mixed"])
    X = w["Xq"].to(DEV)
    v64, g64 = acqs["fp64"].value_and_grad(X)
    vm, gm = acqs["mixed"].value_and_grad(X)
    orc = O.OracleGP(w["X"].to(DEV), w["Y"].to(DEV), O.MATERN52, w["lengthscale"].to(DEV),
                     w["outputscale"], w["noise"])
    ref = O.OracleQLogNEI(orc, Xb, num_samples=w["S"], seed=0)
    v_ref, g_ref = ref.value_and_grad(X[:16])
    e_v, e_g = _rel(v64[:16], v_ref), _rel(g64[:16], g_ref)
    print(f"n_b={Xb.shape[0]} fp64 vs oracle: value {e_v:.2e} grad {e_g:.2e}", flush=True)
    assert e_v < 1e-6 and e_g < 1e-6
    # the mixed mode against the FP64 path (itself checked above) on all 1024
    # restarts: the clustered points leave posterior variances far below the prior
    # one, measured 2.2e-7 value / 4.4e-4 gradient (all restarts improving)
    mixed_contract.check(vm, gm, v64, g64, ref, X, 1e-4, 1e-3, "mixed vs fp64")


@pytest.mark.parametrize("idx", [0, 7, 21, 38, 50, 63])
def test_batched_uneven_models(idx):
    # the full set of 64 models is generated; a spread of them is checked
    ws = workloads.gp_batch()
    w = ws[idx]
    n, d = w["X"].shape
    _say(f"batched model {idx}: n={n} d={d} q=2 S=256")
    acqs, Xb = _build(w, ["fp64", "mixed"])
    X = w["Xq"][:32].to(DEV)
    orc = O.OracleGP(w["X"], w["Y"], O.MATERN52, w["lengthscale"], w["outputscale"],
                     w["noise"])
    ref = O.OracleQLogNEI(orc, Xb.cpu(), num_samples=w["S"], seed=0)
    v_ref, g_ref = ref.value_and_grad(X.cpu())
    v, gr = acqs["fp64"].value_and_grad(X)
    ev, eg = _rel(v.cpu(), v_ref), _rel(gr.cpu(), g_ref)
    print(f"fp64: value {ev:.2e} grad {eg:.2e}", flush=True)
    assert ev < 1e-6 and eg < 1e-6
    v, gr = acqs["mixed"].value_and_grad(X)
    # measured up to 2.1e-4 gradient (model 63: n = 224, d = 4)
    mixed_contract.check(v, gr, v_ref, g_ref, ref, X.cpu(), 1e-4, 3e-4, "mixed")


def test_large_gp_cholesky_jitter_and_refinement():
    """The large configuration's stated numerics: the factor of K + 1e-6 I and one
    refinement step of the solves against K itself.  Against the FP64 oracle's
    exact solve (no jitter): the jitter alone moves alpha = K^-1 (y - c) and the
    posterior mean visibly; one refinement step takes most of that back."""
    _say("large GP n=16384 d=20: cholesky_jitter 1e-6, refine_solves 0 / 1")
    w = workloads.gp_large()
    d = w["X"].shape[-1]
    base = B.MaternKernel(ard_num_dims=d)
    base.lengthscale = w["lengthscale"]
    covar = B.ScaleKernel(base)
    covar.outputscale = w["outputscale"]
    lik = B.GaussianLikelihood()
    lik.noise = w["noise"]
    gp = B.SingleTaskGP(w["X"].to(DEV), w["Y"].to(DEV), likelihood=lik, covar_module=covar)
    orc = O.OracleGP(w["X"].to(DEV), w["Y"].to(DEV), O.MATERN52, w["lengthscale"].to(DEV),
                     w["outputscale"], w["noise"])
    n = w["X"].shape[0]
    Xt = torch.rand(64, d, generator=torch.Generator().manual_seed(9), dtype=torch.float64)
    with torch.no_grad():
        m_ref, _ = orc.posterior(Xt.to(DEV))
    errs = {}
    gp.cholesky_jitter = 1e-6
    for steps in (0, 1):
        gp.refine_solves = steps
        alpha = gp.prediction_cache()["alpha"][:n]
        with torch.no_grad():
            mean = gp.posterior(Xt.to(DEV)).mean.squeeze(-1)
        errs[steps] = (_rel(alpha, orc.mean_cache), _rel(mean, m_ref))
    print(f"jitter 1e-6: alpha / mean rel err, no refinement {errs[0][0]:.2e} / {errs[0][1]:.2e}, "
          f"one step {errs[1][0]:.2e} / {errs[1][1]:.2e}", flush=True)
    assert errs[1][0] < errs[0][0] / 20 and errs[1][1] < errs[0][1] / 20
    assert errs[1][0] < 1e-3 and errs[1][1] < 1e-4
