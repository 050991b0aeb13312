"""The reference's fallback from cached-root to joint sampling
(acquisition/cached_cholesky.py:143-158; test/acquisition/test_logei.py:622-626):
when the low-rank update of the cached baseline root fails (NaN or a q x q block
that stays not PSD after the jitter tries), the batch is sampled again from the
Cholesky of the joint (X_baseline, X) posterior, with one BotorchWarning.

* joint mode on its own against the oracle's joint path (cache_root=False),
  values and gradients, FP64 and mixed;
* a zeroed cached root (the reference test's way to make the update fail):
  exactly one BotorchWarning, and values and gradients equal the joint path's.
"""
import warnings

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


def _pair(w, precision, q):
    d = w["X"].shape[-1]
    base = B.MaternKernel(ard_num_dims=d)
    base.lengthscale = w["lengthscale"]
    covar = B.ScaleKernel(base)
    covar.outputscale = w["outputscale"]
    lik = B.GaussianLikelihood()
    lik.noise = w["noise"]
    gp = B.SingleTaskGP(w["X"].to(DEV), w["Y"].to(DEV), likelihood=lik, covar_module=covar)
    orc = O.OracleGP(w["X"], w["Y"], O.MATERN52, w["lengthscale"], w["outputscale"], w["noise"])
    # a few baseline points so that the joint factor has a real baseline block
    Xb = w["X"][[3, 17, 40, 51]]
    acq = B.qLogNoisyExpectedImprovement(
        gp, Xb.to(DEV), sampler=B.SobolQMCNormalSampler(torch.Size([w["S"]]), seed=0),
        prune_baseline=False, precision=precision)
    ref = O.OracleQLogNEI(orc, Xb, num_samples=w["S"], seed=0, cache_root=False)
    g = torch.Generator().manual_seed(5 + q)
    X = torch.rand(16, q, d, generator=g, dtype=torch.float64)
    return acq, ref, X


@pytest.mark.parametrize("precision,q", [("fp64", 1), ("fp64", 3), ("mixed", 2)])
def test_joint_mode_matches_oracle_joint_path(precision, q):
    _say(f"joint sampling mode {precision} q={q} (n=64 d=6, n_b=4)")
    acq, ref, X = _pair(workloads.gp_small(), precision, q)
    v, g, st = acq._launch(X.to(DEV).contiguous(), True, joint=True)
    v_ref, g_ref = ref.value_and_grad(X)
    ev, eg = _rel(v, v_ref), _rel(g, g_ref)
    print(f"value {ev:.2e} grad {eg:.2e}", flush=True)
    assert int(st.abs().max()) == 0
    tol = 1e-9 if precision == "fp64" else 1e-4
    assert ev < tol and eg < tol


def test_zeroed_cached_root_falls_back_with_one_warning():
    _say("zeroed cached root -> joint sampling fallback (fp64, q=2)")
    acq, ref, X = _pair(workloads.gp_small(), "fp64", 2)
    v_joint, g_joint, _ = acq._launch(X.to(DEV).contiguous(), True, joint=True)
    acq._baseline_L.zero_()  # the cached root the kernels read
    Xr = X.to(DEV).requires_grad_(True)
    with warnings.catch_warnings(record=True) as ws:
        warnings.simplefilter("always")
        val = acq(Xr)
        val.sum().backward()
    assert sum(issubclass(x.category, B.BotorchWarning) for x in ws) == 1
    assert torch.equal(val.detach(), v_joint)
    assert torch.equal(Xr.grad, g_joint)
    v_ref, g_ref = ref.value_and_grad(X)
    assert _rel(val, v_ref) < 1e-9 and _rel(Xr.grad, g_ref) < 1e-9
