"""Host-side optimiser logic (no GPU): the lock-step batched L-BFGS-B reproduces
SciPy's own L-BFGS-B per problem (optim/batched_lbfgs_b.py semantics, tested in
the reference by test/optim/test_batched_lbfgs_b.py on toy quadratics), and the
Boltzmann initial-condition pick keeps the best raw point."""
import numpy as np
import pytest
import scipy.optimize
import torch

from botorch_amd.optim import fmin_l_bfgs_b_batched, initialize_q_batch


def _rosen_batch(x, idx):
    f = np.array([scipy.optimize.rosen(xi) for xi in x])
    g = np.stack([scipy.optimize.rosen_der(xi) for xi in x])
    return f, g


def test_batched_lbfgsb_matches_scipy_per_problem():
    rng = np.random.default_rng(0)
    x0 = rng.uniform(-1.5, 1.5, size=(5, 4))
    lb, ub = np.full(4, -2.0), np.full(4, 2.0)
    x, f, info = fmin_l_bfgs_b_batched(_rosen_batch, x0, (lb, ub), maxiter=200)
    for i in range(5):
        r = scipy.optimize.minimize(scipy.optimize.rosen, x0[i], jac=scipy.optimize.rosen_der,
                                    method="L-BFGS-B", bounds=list(zip(lb, ub)),
                                    options={"maxiter": 200, "maxcor": 10, "gtol": 1e-5,
                                             "ftol": 1e7 * np.finfo(float).eps})
        np.testing.assert_allclose(x[i], r.x, rtol=1e-10, atol=1e-12)
        assert info[i]["nit"] == r.nit


def test_batched_lbfgsb_respects_bounds_and_maxiter():
    x0 = np.array([[3.0, -3.0], [0.5, 0.5]])

    def quad(x, idx):
        c = np.array([5.0, -5.0])
        return ((x - c) ** 2).sum(1), 2 * (x - c)

    x, f, info = fmin_l_bfgs_b_batched(quad, np.clip(x0, -1, 1), (np.full(2, -1.0),
                                                                   np.full(2, 1.0)))
    np.testing.assert_allclose(x, [[1.0, -1.0], [1.0, -1.0]])
    x2, _, info2 = fmin_l_bfgs_b_batched(_rosen_batch, np.zeros((2, 3)),
                                         (np.full(3, -2.0), np.full(3, 2.0)), maxiter=2)
    assert all(i["nit"] == 2 and i["warnflag"] == 1 for i in info2)


def test_initialize_q_batch_keeps_best():
    g = torch.Generator().manual_seed(0)
    X = torch.rand(100, 2, 3, generator=g)
    y = torch.randn(100, generator=g)
    Xs, ys = initialize_q_batch(X, y, 10, generator=torch.Generator().manual_seed(1))
    assert Xs.shape == (10, 2, 3)
    assert y.max() in ys


def test_constraint_violating_hyperparameter_raises():
    """As gpytorch's Module.initialize: a value below a GreaterThan bound is
    refused instead of producing a NaN raw parameter."""
    import botorch_amd as B
    lik = B.GaussianLikelihood()
    with pytest.raises(RuntimeError, match="out of bounds"):
        lik.noise = 1e-6
    lik.noise = 1e-4  # the bound itself is allowed
    assert abs(float(lik.noise.detach()) - 1e-4) < 1e-12


class _WrongGradAcqf:
    """Host stand-in for an acquisition function whose gradient has the wrong
    sign: every line search fails, so each restart ends abnormally and
    gen_candidates_scipy raises an OptimizationWarning."""

    def __init__(self, sign=-1.0):
        self.X_baseline = torch.zeros(1, 2, dtype=torch.float64)
        self.sign = sign
        self.calls = 0

    def __call__(self, X):
        return -((X - 0.3) ** 2).sum(dim=(-1, -2))

    def value_and_grad(self, X):
        self.calls += 1
        return self(X), self.sign * (-2.0) * (X - 0.3)


def test_optimize_acqf_retries_on_optimization_warning():
    """optim/optimize.py:476-526: an OptimizationWarning from the restarts
    draws new initial conditions and retries once; a second failure warns
    again; retry_on_optimization_warning=False keeps the first result."""
    import warnings

    from botorch_amd.optim import optimize_acqf
    bounds = torch.tensor([[0.0, 0.0], [1.0, 1.0]], dtype=torch.float64)
    acq = _WrongGradAcqf()
    with warnings.catch_warnings(record=True) as ws:
        warnings.simplefilter("always")
        optimize_acqf(acq, bounds, q=1, num_restarts=3, raw_samples=16)
    msgs = [str(w.message) for w in ws if issubclass(w.category, RuntimeWarning)]
    assert any("Trying again with a new set of initial conditions" in m for m in msgs)
    assert any("Optimization failed on the second try" in m for m in msgs)
    first = acq.calls
    acq2 = _WrongGradAcqf()
    with warnings.catch_warnings(record=True) as ws:
        warnings.simplefilter("always")
        optimize_acqf(acq2, bounds, q=1, num_restarts=3, raw_samples=16,
                      retry_on_optimization_warning=False)
    assert not [w for w in ws if issubclass(w.category, RuntimeWarning)]
    assert acq2.calls * 2 == first
    # provided initial conditions covering num_restarts: no new draw, one warning
    acq3 = _WrongGradAcqf()
    ics = torch.rand(3, 1, 2, dtype=torch.float64)
    with warnings.catch_warnings(record=True) as ws:
        warnings.simplefilter("always")
        optimize_acqf(acq3, bounds, q=1, num_restarts=3, batch_initial_conditions=ics)
    msgs = [str(w.message) for w in ws if issubclass(w.category, RuntimeWarning)]
    assert len(msgs) == 1 and "will not be retried" in msgs[0]
    # a correct gradient converges without warnings or retries
    ok = _WrongGradAcqf(sign=1.0)
    with warnings.catch_warnings(record=True) as ws:
        warnings.simplefilter("always")
        cand, val = optimize_acqf(ok, bounds, q=1, num_restarts=3, raw_samples=16)
    assert not [w for w in ws if issubclass(w.category, RuntimeWarning)]
    assert torch.allclose(cand, torch.full((1, 2), 0.3, dtype=torch.float64), atol=1e-5)
