"""Device-resident batched L-BFGS-B (bta_lbfgsb_*) against SciPy's setulb, the
per-restart routine of the reference's lock-step loop
(optim/batched_lbfgs_b.py:440-575), and gen_candidates_scipy's device path
against its host SciPy path on qLogNEI.

The first tests feed both optimisers f and g from the same host function (the
device rounds copy x to the host each round), so equal points give equal f, g
and the comparison isolates the optimiser's arithmetic: every evaluation point
of a restart agrees with SciPy's, except after a bound knife edge.  When a line
search ends at its step bound, the variable that set the bound lands within an
ulp of it - exactly on it or one ulp inside, depending on the last bit of the
step, which the two codes' different summation orders decide.  The variable is
then active in one run and free in the other, and the paths part.  A
divergence is accepted only when an earlier common point shows exactly that (a
coordinate within 1e-12 in both runs, on the bound in one only), and most
restarts must agree to the end."""
import numpy as np
import pytest
import torch

from botorch_amd.exceptions import UnsupportedError
from botorch_amd.optim import fmin_l_bfgs_b_device

from test_lbfgsb_host import scipy_trace

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _say(msg):
    print(f"[lbfgsb] {msg}", flush=True)


def _sinquad(n, seed):
    r = np.random.default_rng(seed)
    a, b = r.uniform(1, 4, n), r.uniform(-1, 1, n)
    A = r.standard_normal((n, n)) / np.sqrt(n)

    def fun(x):
        z = A @ x
        return float(np.sum(np.sin(a * x + b)) + 0.3 * z @ z), a * np.cos(a * x + b) + 0.6 * A.T @ z
    return fun


def _quad(n, seed):
    """A convex quadratic with its minimiser inside the box: several iterations."""
    r = np.random.default_rng(seed)
    A = r.standard_normal((n, n))
    H = A @ A.T / n + 0.5 * np.eye(n)
    c = r.uniform(0.2, 0.8, n)
    return lambda x: (float(0.5 * (x - c) @ H @ (x - c)), H @ (x - c))


def _host_evaluate(fun, b, n):
    def evaluate(x_eval):
        X = x_eval.cpu().numpy()
        fg = [fun(X[i]) for i in range(b)]
        f = torch.tensor([v[0] for v in fg], dtype=torch.float64, device=DEV)
        g = torch.tensor(np.stack([v[1] for v in fg]), dtype=torch.float64, device=DEV)
        return f, g
    return evaluate


def _knife_edge_before(ps, pm, k, lb, ub):
    """Whether one of the first k common points has a coordinate within 1e-12
    in both runs that lies on a bound in one run only."""
    for j in range(min(k, len(ps), len(pm))):
        a, c = ps[j], pm[j]
        on_a = (a == lb) | (a == ub)
        on_c = (c == lb) | (c == ub)
        if ((np.abs(a - c) <= 1e-12) & (on_a != on_c)).any():
            return True
    return False


def _device_vs_scipy(fun, x0, lb, ub, tol, evaluate=None, min_exact=0.8, **kw):
    """Per restart: SciPy's evaluation points (fun per point on the host) against
    the device rounds (evaluate, by default fun on the host for every row)."""
    b, n = x0.shape
    trace = []
    x, f, info = fmin_l_bfgs_b_device(evaluate or _host_evaluate(fun, b, n),
                                      torch.tensor(x0, dtype=torch.float64, device=DEV),
                                      (lb, ub), trace=trace, **kw)
    rounds = np.stack([t.cpu().numpy() for t in trace])  # (rounds+1) x b x n
    exact = 0
    for i in range(b):
        ps, xs, nit, task = scipy_trace(fun, x0[i].copy(), lb, ub, m=kw.get("maxcor", 10),
                                        maxls=kw.get("maxls", 20),
                                        maxiter=kw.get("maxiter", 200))
        pm = rounds[:info[i]["nfev"], i]
        k = min(len(ps), len(pm))
        scale = np.maximum(1.0, np.abs(ps[:k]).max(axis=1))
        err = np.abs(ps[:k] - pm[:k]).max(axis=1) / scale
        first = int(np.argmax(err > tol)) if (err > tol).any() else k
        same = first == k and len(ps) == len(pm)
        if same:
            exact += 1
            assert info[i]["nit"] == nit, (i, info[i], nit)
            assert {4: info[i]["task"] in (1, 2), 5: info[i]["task"] == 4}.get(
                task, info[i]["task"] == 3), (i, task, info[i])
            np.testing.assert_allclose(x[i].cpu().numpy(), xs, rtol=0,
                                       atol=tol * max(1, abs(xs).max()))
        else:
            assert _knife_edge_before(ps, pm, first, lb, ub), \
                (i, first, len(ps), len(pm), float(err[first]) if first < k else None)
    _say(f"{exact} of {b} restarts identical to setulb at every evaluation")
    assert exact >= min_exact * b, (exact, b)
    return info


@pytest.mark.parametrize("n", [5, 48, 160])
def test_device_lbfgsb_matches_setulb_box(n):
    _say(f"device vs setulb, box [0,1]^{n}, 16 restarts")
    rng = np.random.default_rng(n)
    x0 = rng.uniform(0, 1, (16, n))
    info = _device_vs_scipy(_sinquad(n, 100 + n), x0, np.zeros(n), np.ones(n), 1e-8,
                            maxiter=60)
    assert all(i["warnflag"] in (0, 1) for i in info)


def test_device_lbfgsb_matches_setulb_bound_types_and_settings():
    _say("device vs setulb: mixed bound types, x0 outside, m=3, maxls=2")
    n = 24
    rng = np.random.default_rng(7)
    lb = np.where(np.arange(n) % 4 < 2, 0.0, -np.inf)
    ub = np.where(np.arange(n) % 4 >= 1, 1.0, np.inf)
    lb[5] = ub[5] = 0.5  # fixed
    x0 = rng.uniform(-0.5, 1.5, (12, n))
    _device_vs_scipy(_sinquad(n, 3), x0, lb, ub, 1e-8, maxcor=3, maxls=2, maxiter=40)


def test_device_lbfgsb_abnormal_and_limits():
    _say("device vs setulb: wrong-sign gradient (abnormal end), maxiter stop")
    n, b = 8, 6
    rng = np.random.default_rng(9)
    c = rng.uniform(-1, 2, n)
    wrong = lambda x: (float(np.sum((x - c) ** 2)), -2 * (x - c))  # noqa: E731
    info = _device_vs_scipy(wrong, rng.uniform(0, 1, (b, n)), np.zeros(n), np.ones(n), 1e-10)
    assert all(i["task"] == 3 and i["warnflag"] == 2 for i in info)
    info = _device_vs_scipy(_quad(n, 4), rng.uniform(0, 1, (b, n)), np.zeros(n), np.ones(n),
                            1e-10, maxiter=3)
    assert all(i["nit"] == 3 and i["task"] == 4 and i["warnflag"] == 1 for i in info)
    # maxfun: the host driver's rule (stop before the evaluation that would exceed it)
    x, f, info = fmin_l_bfgs_b_device(_host_evaluate(_quad(n, 4), b, n),
                                      torch.rand(b, n, dtype=torch.float64, device=DEV),
                                      (np.zeros(n), np.ones(n)), maxfun=4)
    assert all(i["nfev"] == 4 and i["task"] == 5 and i["warnflag"] == 1 for i in info)


def test_device_lbfgsb_nan_gradient_ends_that_restart():
    _say("device L-BFGS-B: NaN gradient of one restart")
    n, b = 6, 4
    fun = _quad(n, 11)
    calls = [0]

    def evaluate(x_eval):
        f, g = _host_evaluate(fun, b, n)(x_eval)
        calls[0] += 1
        if calls[0] == 3:
            g[2, 1] = float("nan")
        return f, g
    x, f, info = fmin_l_bfgs_b_device(evaluate, torch.rand(b, n, dtype=torch.float64, device=DEV),
                                      (np.zeros(n), np.ones(n)), rounds_per_check=1)
    assert info[2]["task"] == 6 and info[2]["nfev"] == 3
    assert all(info[i]["task"] in (1, 2) for i in (0, 1, 3))


def test_device_lbfgsb_graph_replay_equals_eager():
    _say("device L-BFGS-B: HIP graph replay vs eager rounds")
    n, b = 12, 32
    r = np.random.default_rng(5)
    a = torch.tensor(r.uniform(1, 4, n), device=DEV)
    c = torch.tensor(r.uniform(-1, 1, n), device=DEV)

    def evaluate(x_eval):  # device-only objective: capturable
        f = torch.sin(a * x_eval + c).sum(1) + 0.5 * (x_eval * x_eval).sum(1)
        g = a * torch.cos(a * x_eval + c) + x_eval
        return f, g
    x0 = torch.rand(b, n, dtype=torch.float64, device=DEV)
    bounds = (np.zeros(n), np.ones(n))
    xe, fe, ie = fmin_l_bfgs_b_device(evaluate, x0, bounds, rounds_per_check=4)
    xg, fg, ig = fmin_l_bfgs_b_device(evaluate, x0, bounds, rounds_per_check=4, use_graph=True)
    assert torch.equal(xe, xg) and torch.equal(fe, fg)
    assert [(i["nit"], i["nfev"], i["task"]) for i in ie] == \
        [(i["nit"], i["nfev"], i["task"]) for i in ig]


def test_device_lbfgsb_rejects_oversize_problem():
    with pytest.raises(UnsupportedError):
        fmin_l_bfgs_b_device(lambda x: None, torch.zeros(2, 300, dtype=torch.float64, device=DEV),
                             (np.zeros(300), np.ones(300)))


def test_device_lbfgsb_on_qlognei_matches_setulb():
    """qLogNEI (FP64) on a small GP, q = 2: the device rounds evaluate all
    restarts with the batched HIP step; SciPy's setulb per restart evaluates
    the same acquisition at its own points, in a batch of the same shape (every
    row set to the point), so each restart's value and gradient come from the
    same arithmetic in both runs and the paths agree evaluation by evaluation
    (up to bound knife edges), with identical evaluation and iteration counts.
    Bound 1e-5 relative on the points: the two optimisers' sums differ in the
    last bit and the acquisition's flat approach to its maxima amplifies that
    (measured: 3e-8 after 14 evaluations, 1.1e-6 after 22)."""
    import botorch_amd as B
    from botorch_amd import workloads
    _say("device L-BFGS-B vs setulb on qLogNEI q=2 (8 restarts)")
    w = workloads.gp_small()
    base = B.MaternKernel(ard_num_dims=6)
    base.lengthscale = w["lengthscale"]
    lik = B.GaussianLikelihood()
    lik.noise = w["noise"]
    covar = B.ScaleKernel(base)
    covar.outputscale = w["outputscale"]
    gp = B.SingleTaskGP(w["X"].to(DEV), w["Y"].to(DEV), likelihood=lik, covar_module=covar)
    acq = B.qLogNoisyExpectedImprovement(gp, w["X"].to(DEV), prune_baseline=False,
                                         sampler=B.SobolQMCNormalSampler(torch.Size([128]),
                                                                         seed=0))
    b, q, d = 8, 2, 6

    def fun(x):  # one restart, in a batch of b copies
        X = torch.tensor(x, device=DEV).view(1, q, d).expand(b, q, d).contiguous()
        v, g, st = acq._launch(X, True)
        return -float(v[0]), -g[0].reshape(-1).cpu().numpy()

    def evaluate(x_eval):  # all restarts in one step
        v, g, st = acq._launch(x_eval.view(b, q, d), True)
        return -v, -g
    x0 = np.random.default_rng(0).uniform(0, 1, (b, q * d))
    _device_vs_scipy(fun, x0, np.zeros(q * d), np.ones(q * d), 1e-5, evaluate=evaluate,
                     maxiter=25, min_exact=0.75)


def test_gen_candidates_device_path_runs_the_device_optimiser():
    """gen_candidates_scipy on the GPU takes the device path by default and
    returns candidates in the box with their acquisition values."""
    import botorch_amd as B
    from botorch_amd import optim, workloads
    _say("gen_candidates_scipy device path")
    w = workloads.gp_small()
    gp = B.SingleTaskGP(w["X"].to(DEV), w["Y"].to(DEV))
    acq = B.qLogNoisyExpectedImprovement(gp, w["X"].to(DEV), prune_baseline=False,
                                         sampler=B.SobolQMCNormalSampler(torch.Size([64]),
                                                                         seed=0))
    used = []
    orig = optim.fmin_l_bfgs_b_device

    def spy(*a, **k):
        used.append(1)
        return orig(*a, **k)
    optim.fmin_l_bfgs_b_device = spy
    try:
        X0 = torch.rand(4, 2, 6, dtype=torch.float64, device=DEV)
        c, v = optim.gen_candidates_scipy(X0, acq, 0.0, 1.0, options={"maxiter": 10})
    finally:
        optim.fmin_l_bfgs_b_device = orig
    assert used and c.shape == (4, 2, 6) and v.shape == (4,)
    assert (c >= 0).all() and (c <= 1).all()
    with torch.no_grad():
        assert torch.allclose(v, acq(c), rtol=1e-12, atol=0)
