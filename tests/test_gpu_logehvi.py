"""The fused qLogEHVI log-areas op (the reference's csrc/logei_fused.cpp, bound at
acquisition/multi_objective/logei.py:107-170) on HIP against the oracle's
restatement of the pure-Python path (logei.py:437-479): values and gradients
(autograd through the oracle), FP64 and FP32, shared and batched cell bounds,
unbounded cells, subset sizes from 1 to 12 and up to 4 objectives."""
import pytest
import torch

from botorch_amd.multi_objective import fused_log_areas
from oracle import gp_oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _case(B, n_sub, i, m, nc, batched, dtype, seed):
    g = torch.Generator().manual_seed(seed)
    obj = torch.randn(B, n_sub, i, m, generator=g, dtype=torch.float64)
    cl = torch.randn(B if batched else 1, nc, m, generator=g, dtype=torch.float64) - 1.0
    cu = cl + 0.2 + torch.rand(cl.shape, generator=g, dtype=torch.float64)
    cu[..., -1, :] = float("inf")  # the top cell of a box decomposition is unbounded
    if not batched:
        cl, cu = cl[0], cu[0]
    return obj.to(dtype), cl.to(dtype), cu.to(dtype)


@pytest.mark.parametrize("B,n_sub,i,m,nc,batched,dtype", [
    (64, 6, 2, 3, 20, False, torch.float64),
    (32, 4, 1, 2, 9, False, torch.float64),
    (16, 10, 3, 2, 30, True, torch.float64),
    (8, 3, 12, 4, 7, False, torch.float64),
    (64, 6, 2, 3, 20, False, torch.float32),
    (16, 5, 8, 4, 11, True, torch.float32),
])
def test_log_areas_match_oracle(B, n_sub, i, m, nc, batched, dtype):
    _say(f"log areas B={B} n_sub={n_sub} i={i} m={m} nc={nc} batched={batched} {dtype}")
    obj, cl, cu = _case(B, n_sub, i, m, nc, batched, dtype, seed=B + i + m)
    x = obj.to(DEV).requires_grad_(True)
    out = fused_log_areas(x, cl.to(DEV), cu.to(DEV), 1e-6, 1e-2)
    assert out.shape == (B, nc, n_sub)
    w = torch.randn(out.shape, generator=torch.Generator().manual_seed(1), dtype=dtype)
    (out * w.to(DEV)).sum().backward()
    xr = obj.double().requires_grad_(True)
    ref = O.log_areas(xr, cl.double(), cu.double(), 1e-6, 1e-2)
    (ref * w.double()).sum().backward()
    ev = ((out.detach().double().cpu() - ref.detach()).abs().max() / ref.abs().max()).item()
    eg = ((x.grad.double().cpu() - xr.grad).abs().max() / xr.grad.abs().max()).item()
    print(f"value {ev:.2e} grad {eg:.2e}", flush=True)
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    assert ev < tol and eg < (1e-10 if dtype == torch.float64 else 2e-4)
