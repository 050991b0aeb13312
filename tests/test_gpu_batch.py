"""The batched step over many models (bta_qlognei_batch_value_grad) on the
batched configuration at full size: 64 independent Matern-5/2 GPs, n
log-uniform in [128, 4096], d in {4..20}, each scoring 256 restarts x q = 2
with 256 Sobol samples, value and gradient, mixed precision.

* every model's values, gradients and status words equal those of the
  one-model step (bta_qlognei_value_grad) bit for bit;
* a slice of restarts of a spread of models against the FP64 oracle;
* the step captured in a HIP graph and replayed with refilled inputs equals the
  eager step bit for bit on every replay;
* a value-only batch equals the values of the value+grad batch.
"""
import pytest
import torch

import botorch_amd as B
from botorch_amd import workloads
from oracle import gp_oracle as O

import mixed_contract

pytestmark = pytest.mark.gpu
DEV = "cuda"

_CACHE = {}


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _models():
    if "acqs" not in _CACHE:
        ws = workloads.gp_batch()
        acqs, Xbs = [], []
        for w in ws:
            d = w["X"].shape[-1]
            base = B.MaternKernel(ard_num_dims=d)
            base.lengthscale = w["lengthscale"]
            covar = B.ScaleKernel(base)
            covar.outputscale = w["outputscale"]
            lik = B.GaussianLikelihood()
            lik.noise = w["noise"]
            gp = B.SingleTaskGP(w["X"].to(DEV), w["Y"].to(DEV), likelihood=lik,
                                covar_module=covar)
            with torch.no_grad():
                Xb = B.prune_inferior_points(
                    gp, w["X"].to(DEV),
                    sampler=B.SobolQMCNormalSampler(torch.Size([2048]), seed=123))
            acqs.append(B.qLogNoisyExpectedImprovement(
                gp, Xb, sampler=B.SobolQMCNormalSampler(torch.Size([w["S"]]), seed=0),
                prune_baseline=False, precision="mixed"))
            Xbs.append(Xb)
        _CACHE.update(ws=ws, acqs=acqs, Xbs=Xbs)
    return _CACHE["ws"], _CACHE["acqs"], _CACHE["Xbs"]


def test_batched_step_equals_one_model_steps():
    """Every model's restarts run the one-model step's bodies with the same MC
    workgroup choice (two restarts per workgroup by default, one with
    BTA_MC_PAIR=0, fixed when the batch is prepared): values, gradients and
    status words equal the one-model steps' bit for bit in both modes."""
    import os
    ws, acqs, _ = _models()
    ns = [w["X"].shape[0] for w in ws]
    old = os.environ.get("BTA_MC_PAIR")
    try:
        for pair in ("1", "0"):
            os.environ["BTA_MC_PAIR"] = pair
            _say(f"batched step: {len(ws)} models, n {min(ns)}..{max(ns)}, b=256 q=2 S=256 "
                 f"mixed, BTA_MC_PAIR={pair}")
            batch = B.BatchedQLogNEI(acqs, 256, q=2)
            vals, grads = batch.value_and_grad([w["Xq"].to(DEV) for w in ws])
            for i, (w, acq) in enumerate(zip(ws, acqs)):
                v1, g1, st1 = acq._launch(w["Xq"].to(DEV).contiguous(), True)
                torch.cuda.synchronize()
                assert torch.equal(batch.status[i], st1), (pair, i)
                assert torch.equal(vals[i], v1), (pair, i)
                assert torch.equal(grads[i], g1), (pair, i)
    finally:
        if old is None:
            del os.environ["BTA_MC_PAIR"]
        else:
            os.environ["BTA_MC_PAIR"] = old


@pytest.mark.parametrize("idx", [0, 21, 50, 63])
def test_batched_step_matches_oracle(idx):
    ws, acqs, Xbs = _models()
    w = ws[idx]
    n, d = w["X"].shape
    _say(f"batched step vs oracle, model {idx}: n={n} d={d}, 32 restarts")
    batch = B.BatchedQLogNEI(acqs, 256, q=2)
    vals, grads = batch.value_and_grad([x["Xq"].to(DEV) for x in ws])
    orc = O.OracleGP(w["X"], w["Y"], O.MATERN52, w["lengthscale"], w["outputscale"],
                     w["noise"])
    ref = O.OracleQLogNEI(orc, Xbs[idx].cpu(), num_samples=w["S"], seed=0)
    X = w["Xq"][:32]
    v_ref, g_ref = ref.value_and_grad(X)
    # the mixed-mode tolerance of the one-model batched-configuration test
    mixed_contract.check(vals[idx][:32], grads[idx][:32], v_ref, g_ref, ref, X, 1e-4, 3e-4,
                         f"model {idx}")


def test_batched_step_graph_replay():
    ws, acqs, _ = _models()
    sub = list(range(0, 64, 4))  # 16 of the models
    _say(f"batched step graph replay: {len(sub)} models")
    batch = B.BatchedQLogNEI([acqs[i] for i in sub], 256, q=2)
    g = torch.Generator().manual_seed(3)
    batch.value_and_grad([ws[i]["Xq"].to(DEV) for i in sub])  # warm-up outside the capture
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        batch.launch()
    for _ in range(3):
        Xn = [torch.rand(ws[i]["Xq"].shape, generator=g, dtype=torch.float64).to(DEV)
              for i in sub]
        for buf, x in zip(batch.X, Xn):
            buf.copy_(x)
        graph.replay()
        torch.cuda.synchronize()
        v_g = [v.clone() for v in batch.value]
        g_g = [t.clone() for t in batch.grad]
        v_e, g_e = batch.value_and_grad(Xn)
        torch.cuda.synchronize()
        for a, b in zip(v_g, v_e):
            assert torch.equal(a, b)
        for a, b in zip(g_g, g_e):
            assert torch.equal(a, b)


def test_batched_value_only():
    ws, acqs, _ = _models()
    sub = [3, 17, 40]
    _say(f"batched value-only step: models {sub}")
    Xs = [ws[i]["Xq"].to(DEV) for i in sub]
    full = B.BatchedQLogNEI([acqs[i] for i in sub], 256, q=2)
    v_full, _ = full.value_and_grad(Xs)
    vonly = B.BatchedQLogNEI([acqs[i] for i in sub], 256, q=2, need_grad=False)
    v_only, g_none = vonly.value_and_grad(Xs)
    assert g_none is None
    for a, b in zip(v_only, v_full):
        assert torch.equal(a, b)
