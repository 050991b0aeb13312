"""Parity of the product's index work, posterior sampling and value-only step
against the FP64 oracle, on the GPU.

* prune_inferior_points (acquisition/utils.py:347-437): the kept points are
  index work, so the product must return exactly the oracle's points for the
  same Sobol seed (bit-identical), including the max_frac truncation.
* GPyTorchPosterior.rsample_from_base_samples (posteriors/gpytorch.py:86-127):
  mean + L z against the oracle's posterior and torch's Cholesky, 1e-10.
* The value-only step (no gradient: the forward that gen_batch_initial_conditions
  runs under no_grad, optim/initializers.py:447-460) against the oracle, and equal
  bit for bit to the value of the value+grad step.
"""
import os

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
    return ((a - r).abs().max() / r.abs().max()).item()


def _gp(w, dev=DEV):
    d = w["X"].shape[-1]
    base = B.MaternKernel(ard_num_dims=d)
    base.lengthscale = w["lengthscale"]
    covar = B.ScaleKernel(base)
    covar.outputscale = w["outputscale"]
    lik = B.GaussianLikelihood()
    lik.noise = w["noise"]
    gp = B.SingleTaskGP(w["X"].to(dev), w["Y"].to(dev), likelihood=lik, covar_module=covar)
    return gp


def _oracle(w, dev="cpu"):
    return O.OracleGP(w["X"].to(dev), w["Y"].to(dev), O.MATERN52, w["lengthscale"].to(dev),
                      w["outputscale"], w["noise"])


# ------------------------------------------------------------------ pruning
@pytest.mark.parametrize("case", ["small", "small_max_frac", "headline", "large"])
def test_prune_inferior_points_matches_oracle(case):
    seed, max_frac, oracle_dev = 123, 1.0, "cpu"
    if case.startswith("small"):
        w = workloads.gp_small()
        if case == "small_max_frac":
            # a noisier, shorter-lengthscale model spreads the argmax over 12
            # points, so ceil(0.05 * 64) = 4 takes the truncation branch (which
            # keeps positions into the sorted unique indices, as the reference does)
            w = dict(w, noise=0.1, lengthscale=torch.full((6,), 0.2, dtype=torch.float64))
            max_frac = 0.05
    elif case == "headline":
        w = workloads.gp_headline()
    else:
        w = workloads.gp_large()
        oracle_dev = DEV  # the FP64 oracle through PyTorch-ROCm ops: n = 16384
    n = w["X"].shape[0]
    _say(f"prune {case}: n={n} seed={seed} max_frac={max_frac}")
    gp = _gp(w)
    Xd = w["X"].to(DEV)
    with torch.no_grad():
        Xp = B.prune_inferior_points(
            gp, Xd, max_frac=max_frac,
            sampler=B.SobolQMCNormalSampler(torch.Size([2048]), seed=seed))
    orc = _oracle(w, oracle_dev)
    Xr, idx = O.prune_inferior_points(orc, w["X"].to(oracle_dev), num_samples=2048, seed=seed,
                                      max_frac=max_frac)
    print(f"kept {Xp.shape[0]} (oracle {Xr.shape[0]}): {idx.tolist()[:16]}", flush=True)
    assert Xp.shape == Xr.shape
    assert torch.equal(Xp.cpu(), Xr.cpu())
    if case == "small_max_frac":
        assert Xp.shape[0] == 4  # ceil(0.05 * 64), the truncation branch ran


# ------------------------------------------------------------------ rsample
@pytest.mark.parametrize("batch,m", [((3,), 7), ((2,), 90)])
def test_rsample_from_base_samples_matches_oracle(batch, m):
    _say(f"rsample_from_base_samples batch={batch} m={m}")
    w = workloads.gp_small()
    gp = _gp(w)
    orc = _oracle(w)
    g = torch.Generator().manual_seed(17 + m)
    X = torch.rand(*batch, m, 6, generator=g, dtype=torch.float64)
    post = gp.posterior(X.to(DEV))
    mean_r, cov_r = orc.posterior(X)
    L_r = torch.linalg.cholesky(cov_r)
    Z = torch.randn(16, *batch, m, generator=g, dtype=torch.float64)
    s = post.rsample_from_base_samples(torch.Size([16]), Z.to(DEV))
    assert s.shape == (16, *batch, m, 1)
    ref = mean_r.unsqueeze(0) + torch.einsum("...ij,s...j->s...i", L_r, Z)
    err = _rel(s.squeeze(-1).cpu(), ref)
    print(f"max rel err {err:.2e}", flush=True)
    assert err < 1e-10


# ------------------------------------------------------------------ value-only step
@pytest.mark.parametrize("case,precision", [("small", "fp64"), ("headline", "fp64"),
                                            ("headline", "mixed")])
def test_value_only_step_matches_oracle(case, precision):
    w = workloads.gp_small() if case == "small" else workloads.gp_headline(b=128)
    _say(f"value-only step {case} {precision} b={w['b']} q={w['q']} S={w['S']}")
    gp = _gp(w)
    orc = _oracle(w)
    Xb, _ = O.prune_inferior_points(orc, w["X"], num_samples=2048, seed=123)
    acq = B.qLogNoisyExpectedImprovement(
        gp, Xb.to(DEV), sampler=B.SobolQMCNormalSampler(torch.Size([w["S"]]), seed=0),
        prune_baseline=False, precision=precision)
    ref = O.OracleQLogNEI(orc, Xb, num_samples=w["S"], seed=0)
    X = w["Xq"]
    with torch.no_grad():
        v_ref = ref(X)
    Xd = X.to(DEV).contiguous()
    v_only, g_none, st = acq._launch(Xd, False)
    assert g_none is None
    v_full, _, _ = acq._launch(Xd, True)
    torch.cuda.synchronize()
    assert int(st.abs().max()) == 0
    err = _rel(v_only.cpu(), v_ref)
    print(f"value-only rel err {err:.2e}", flush=True)
    assert err < (1e-6 if precision == "fp64" else 1e-4)
    assert torch.equal(v_only, v_full)
    with torch.no_grad():  # forward() without grad takes the same path
        assert torch.equal(acq(Xd), v_only)


# ------------------------------------------------------------------ restart chunks
def test_step_in_restart_chunks_equals_whole_batch():
    """With max_workspace_bytes below the whole batch's workspace, the step runs
    in restart chunks on one workspace (the K_x, V, U rows recomputed chunk by
    chunk instead of held for all restarts): values, gradients and status words
    equal the one-call step's bit for bit."""
    w = workloads.gp_headline(b=512)
    _say("headline step, 512 restarts, in chunks vs whole (mixed)")
    gp = _gp(w)
    orc = _oracle(w)
    Xb, _ = O.prune_inferior_points(orc, w["X"], num_samples=2048, seed=123)
    kw = dict(sampler=B.SobolQMCNormalSampler(torch.Size([w["S"]]), seed=0),
              prune_baseline=False, precision="mixed")
    whole = B.qLogNoisyExpectedImprovement(gp, Xb.to(DEV), **kw)
    X = w["Xq"].to(DEV).contiguous()
    v1, g1, s1 = whole._launch(X, True)
    budget = whole._workspace(512, 4).numel() // 3
    chunked = B.qLogNoisyExpectedImprovement(gp, Xb.to(DEV), max_workspace_bytes=budget, **kw)
    assert chunked._chunk(512, 4) < 512 // 2
    v2, g2, s2 = chunked._launch(X, True)
    torch.cuda.synchronize()
    assert torch.equal(v1, v2) and torch.equal(g1, g2) and torch.equal(s1, s2)
