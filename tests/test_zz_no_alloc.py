"""bta_qlognei_value_grad_noalloc, the qLogNEI step working in the library's
static device array instead of a caller workspace, against
bta_qlognei_value_grad with a workspace from torch's allocator: identical
values, gradients and status words on two calls in a row, the second with
fewer restarts (and a value-only call), in the mixed and the FP64 modes; a call
that needs more scratch than the array holds is refused before anything runs.
"""
import ctypes

import pytest
import torch

import botorch_amd as B
from botorch_amd import _lib, workloads

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _acq(precision):
    w = workloads.gp_headline()
    base = B.MaternKernel(ard_num_dims=12)
    base.lengthscale = w["lengthscale"]
    covar = B.ScaleKernel(base)
    covar.outputscale = w["outputscale"]
    lik = B.GaussianLikelihood()
    lik.noise = w["noise"]
    gp = B.SingleTaskGP(w["X"].to(DEV), w["Y"].to(DEV), likelihood=lik, covar_module=covar)
    with torch.no_grad():
        Xb = B.prune_inferior_points(
            gp, w["X"].to(DEV), sampler=B.SobolQMCNormalSampler(torch.Size([2048]), seed=123))
    acq = B.qLogNoisyExpectedImprovement(
        gp, Xb, sampler=B.SobolQMCNormalSampler(torch.Size([w["S"]]), seed=0),
        prune_baseline=False, precision=precision)
    return acq, w


def _noalloc(acq, X, need_grad=True):
    """One step through the allocation-free entry point, outputs pre-filled with
    NaN (so a result that was not written shows)."""
    b, q, d = X.shape
    st = acq._state
    st.q = q
    st.Zt = _lib.ptr(acq.base_samples_t(q))
    value = torch.full((b,), float("nan"), dtype=torch.float64, device=DEV)
    grad = (torch.full((b, q, d), float("nan"), dtype=torch.float64, device=DEV)
            if need_grad else None)
    status = torch.full((b,), -1, dtype=torch.int32, device=DEV)
    rc = _lib.lib().bta_qlognei_value_grad_noalloc(
        ctypes.byref(acq._keep["cache"]["model"]), ctypes.byref(st), _lib.ptr(X), b,
        _lib.ptr(value), _lib.ptr(grad), _lib.ptr(status), _lib.stream())
    return rc, value, grad, status


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_step_equals_workspace_step(precision):
    acq, w = _acq(precision)
    g = torch.Generator().manual_seed(99)
    for b in (512, 96):  # the full headline batch (59 / 76 MB of scratch), then a smaller one
        _say(f"noalloc step {precision} n=2048 d=12 q=4 S=512 b={b}")
        X = torch.rand(b, 4, 12, generator=g, dtype=torch.float64).to(DEV).contiguous()
        rc, v_n, g_n, st_n = _noalloc(acq, X)
        _lib.check(rc, "qlognei_value_grad_noalloc")
        v_w, g_w, st_w = acq._launch(X, True)
        torch.cuda.synchronize()
        assert torch.equal(st_n, st_w) and int(st_n.abs().max()) == 0
        assert torch.equal(v_n, v_w)
        assert torch.equal(g_n, g_w)
    # value only (no gradient buffers written or read)
    _say(f"noalloc value-only {precision} b=64")
    X = torch.rand(64, 4, 12, generator=g, dtype=torch.float64).to(DEV).contiguous()
    rc, v_n, g_n, st_n = _noalloc(acq, X, need_grad=False)
    _lib.check(rc, "qlognei_value_grad_noalloc")
    v_w, _, st_w = acq._launch(X, False)
    torch.cuda.synchronize()
    assert g_n is None and torch.equal(v_n, v_w) and torch.equal(st_n, st_w)


def test_noalloc_refuses_oversize_call():
    acq, _ = _acq("mixed")
    b = 1024  # ~117 MB of scratch at n = 2048, q = 4
    acq._state.q = 4
    need = _lib.lib().bta_qlognei_workspace_size(ctypes.byref(acq._keep["cache"]["model"]),
                                                 ctypes.byref(acq._state), b)
    assert need > (96 << 20)
    _say(f"noalloc oversize b={b} needs {need} bytes")
    X = torch.rand(b, 4, 12, dtype=torch.float64).to(DEV).contiguous()
    rc, v_n, g_n, st_n = _noalloc(acq, X)
    torch.cuda.synchronize()
    assert rc == _lib.BTA_ERR_WORKSPACE
    assert b"static array" in _lib.lib().bta_last_error()
    assert torch.isnan(v_n).all() and torch.isnan(g_n).all() and (st_n == -1).all()
