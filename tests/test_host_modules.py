"""Host-side module defaults (no GPU): the starting values of the
hyperparameters, which decide where fit_gpytorch_mll begins.

gpytorch initialises every softplus-constrained raw parameter to zero, so the
value starts at lower_bound + softplus(0) = lower_bound + ln 2 (GaussianLikelihood:
GreaterThan(1e-4) -> 1e-4 + ln 2; ScaleKernel and the kernels' lengthscale:
Positive() -> ln 2).  BoTorch's default modules start at their priors' modes
(models/utils/gpytorch_modules.py:74-133; the modes are the closed form
exp(loc - scale^2) of a LogNormal, checked against torch.distributions here).
"""
import math

import pytest
import torch

import botorch_amd as B
from botorch_amd.gp import (get_covar_module_with_dim_scaled_prior,
                            get_gaussian_likelihood_with_lognormal_prior)

LN2 = math.log(2.0)


def test_softplus_defaults_start_at_raw_zero():
    lik = B.GaussianLikelihood()
    assert lik.raw_noise.item() == pytest.approx(0.0, abs=1e-12)
    assert lik.noise.item() == pytest.approx(1e-4 + LN2, rel=1e-15)
    sk = B.ScaleKernel(B.MaternKernel(ard_num_dims=3))
    assert sk.raw_outputscale.item() == pytest.approx(0.0, abs=1e-12)
    assert sk.outputscale.item() == pytest.approx(LN2, rel=1e-15)
    k = B.RBFKernel(ard_num_dims=4)
    assert torch.allclose(k.raw_lengthscale, torch.zeros(4, dtype=torch.float64), atol=1e-12)
    assert torch.allclose(k.lengthscale, torch.full((4,), LN2, dtype=torch.float64))
    # a softplus constraint with a lower bound keeps raw = 0 as well
    k2 = B.MaternKernel(ard_num_dims=2, lengthscale_constraint=B.GreaterThan(0.05))
    assert torch.allclose(k2.lengthscale, torch.full((2,), 0.05 + LN2, dtype=torch.float64))


def test_botorch_default_modules_start_at_prior_modes():
    lik = get_gaussian_likelihood_with_lognormal_prior()
    mode = torch.distributions.LogNormal(torch.tensor(-4.0, dtype=torch.float64),
                                         torch.tensor(1.0, dtype=torch.float64))
    # LogNormal mode exp(loc - scale^2): the maximiser of the density
    xs = torch.linspace(1e-4, 0.05, 200001, dtype=torch.float64)
    x_star = xs[torch.argmax(mode.log_prob(xs))].item()
    assert lik.noise.item() == pytest.approx(math.exp(-5.0), rel=1e-12)
    assert lik.noise.item() == pytest.approx(x_star, abs=5e-7)
    d = 7
    k = get_covar_module_with_dim_scaled_prior(d)
    loc = math.sqrt(2) + 0.5 * math.log(d)
    assert torch.allclose(k.lengthscale, torch.full((d,), math.exp(loc - 3.0),
                                                    dtype=torch.float64), rtol=1e-14)
    assert isinstance(k, B.RBFKernel)
    assert isinstance(get_covar_module_with_dim_scaled_prior(d, use_rbf_kernel=False),
                      B.MaternKernel)


def test_out_of_bounds_assignment_refused():
    lik = B.GaussianLikelihood()
    with pytest.raises(RuntimeError, match="out of bounds"):
        lik.noise = 1e-6
