"""Writes tests/golden/gp_small.json from the FP64 oracle on the gp_small
workload (n=64, d=6, Matern-5/2, lengthscale 0.5, outputscale 1, noise 1e-4;
qLogNEI q=1, 512 Sobol samples, 32 restarts) and a small MLL case.

The reference itself cannot be imported (gpytorch / linear_operator absent),
so these vectors come from the oracle, whose functions are pinned separately
in tests/test_oracle.py; the fixture freezes them so later changes to the
oracle or the HIP path are checked against the same numbers.

Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from botorch_amd import workloads  # noqa: E402
from oracle import gp_oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gp_small.json")


def main():
    w = workloads.gp_small()
    orc = O.OracleGP(w["X"], w["Y"], O.MATERN52, w["lengthscale"], 1.0, w["noise"])
    Xb, idx = O.prune_inferior_points(orc, w["X"], num_samples=2048, seed=123)
    acq = O.OracleQLogNEI(orc, Xb, num_samples=w["S"], seed=0)
    v, g = acq.value_and_grad(w["Xq"])
    gen = torch.Generator().manual_seed(99)
    Xt = torch.rand(4, 6, generator=gen, dtype=torch.float64)
    mean, cov = orc.posterior(Xt)
    # MLL on the same data, standardized targets
    y = orc.train_targets
    ls = w["lengthscale"].clone().requires_grad_(True)
    os_ = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    nz = torch.tensor(w["noise"], dtype=torch.float64, requires_grad=True)
    c = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    ll = O.exact_mll(O.MATERN52, w["X"], y, ls, os_, nz, c) * w["X"].shape[0]
    ll.backward()
    data = {
        "workload": "gp_small (botorch_amd.workloads.gp_small, seed 0)",
        "prune_seed": 123, "prune_samples": 2048, "sampler_seed": 0,
        "baseline_idx": idx.tolist(),
        "best_f_first8": acq.baseline_best_f[:8].tolist(),
        "qlognei_value": v.tolist(), "qlognei_grad": g.tolist(),
        "posterior_X": Xt.tolist(), "posterior_mean": mean.tolist(),
        "posterior_covar": cov.tolist(),
        "mll_loglik": ll.item(),
        "mll_grad": {"lengthscale": ls.grad.tolist(), "outputscale": os_.grad.item(),
                     "noise": nz.grad.item(), "constant": c.grad.item()},
    }
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
