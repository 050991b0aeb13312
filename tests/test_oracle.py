"""Pins the FP64 oracle (oracle/gp_oracle.py) with checks that do not restate
its own formulas: closed forms at chosen points, brute force with explicit
inverses, independent identities (LOO, joint-vs-cached Cholesky sampling),
the reference tests' invariants, finite differences, and library routines."""
import math

import pytest
import torch

from oracle import gp_oracle as O

D = torch.float64


def _data(n=12, d=3, seed=0):
    g = torch.Generator().manual_seed(seed)
    X = torch.rand(n, d, generator=g, dtype=D)
    Y = (torch.sin(3 * X).sum(-1) + 0.1 * torch.randn(n, generator=g, dtype=D)).unsqueeze(-1)
    ls = 0.3 + 0.5 * torch.rand(d, generator=g, dtype=D)
    return X, Y, ls


# ---------------------------------------------------------------- kernels
def test_matern_closed_form_points():
    # two points at scaled distance r: k = (1 + sqrt5 r + 5/3 r^2) exp(-sqrt5 r)
    for r in (0.0, 0.3, 1.0, 2.5):
        x1 = torch.tensor([[0.1, 0.2]], dtype=D)
        x2 = torch.tensor([[0.1 + r * 0.7, 0.2]], dtype=D)
        k = O.covar(O.MATERN52, x1, x2, torch.tensor([0.7, 1.3], dtype=D), 2.0).item()
        s5 = math.sqrt(5.0)
        expect = 2.0 * (1 + s5 * r + 5.0 / 3.0 * r * r) * math.exp(-s5 * r)
        assert k == pytest.approx(expect, rel=1e-12, abs=1e-14)


def test_rbf_closed_form_points():
    x1 = torch.tensor([[0.0, 0.0]], dtype=D)
    x2 = torch.tensor([[0.6, 0.8]], dtype=D)  # scaled distance 1 with unit lengthscales
    k = O.covar(O.RBF, x1, x2, torch.ones(2, dtype=D), 1.5).item()
    assert k == pytest.approx(1.5 * math.exp(-0.5), rel=1e-13)


@pytest.mark.parametrize("kind", [O.MATERN52, O.RBF])
def test_gemm_expansion_matches_direct_differences(kind):
    X, _, ls = _data(30, 4)
    g = torch.Generator().manual_seed(3)
    Z = torch.rand(17, 4, generator=g, dtype=D)
    torch.testing.assert_close(O.covar(kind, Z, X, ls, 1.3), O.covar_direct(kind, Z, X, ls, 1.3),
                               rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------- cholesky
def test_psd_safe_cholesky_on_pd_matches_library():
    X, _, ls = _data(20, 3)
    A = O.covar(O.MATERN52, X, X, ls) + 1e-3 * torch.eye(20, dtype=D)
    torch.testing.assert_close(O.psd_safe_cholesky(A), torch.linalg.cholesky(A))


def test_psd_safe_cholesky_jitter_only_failing_batch_entries():
    good = torch.eye(3, dtype=D) * 2
    bad = torch.ones(3, 3, dtype=D)  # rank one: eigenvalues 3, 0, 0
    L = O.psd_safe_cholesky(torch.stack([good, bad]), warn=False)
    torch.testing.assert_close(L[0], torch.linalg.cholesky(good))
    # the bad entry got the first jitter of the schedule (1e-8), nothing else
    torch.testing.assert_close(L[1] @ L[1].T, bad + 1e-8 * torch.eye(3, dtype=D), rtol=0,
                               atol=1e-12)


def test_psd_safe_cholesky_errors():
    with pytest.raises(O.NotPSDError):
        O.psd_safe_cholesky(-torch.eye(2, dtype=D), warn=False)
    A = torch.eye(2, dtype=D)
    A[0, 0] = float("nan")
    with pytest.raises(O.NanError):
        O.psd_safe_cholesky(A, warn=False)


# ---------------------------------------------------------------- Standardize
def test_standardize():
    Y = torch.tensor([[1.0], [2.0], [4.0]], dtype=D)
    Yt, m, s = O.standardize(Y)
    assert Yt.mean().item() == pytest.approx(0.0, abs=1e-15)
    assert Yt.std().item() == pytest.approx(1.0)
    assert m.item() == pytest.approx(7 / 3)
    Yt1, _, s1 = O.standardize(Y[:1])
    assert s1.item() == 1.0 and Yt1.item() == 0.0


# ---------------------------------------------------------------- posterior
def test_posterior_matches_explicit_inverse_brute_force():
    X, Y, ls = _data(12, 3)
    gp = O.OracleGP(X, Y, O.MATERN52, ls, 1.4, 1e-2, constant=0.2)
    g = torch.Generator().manual_seed(1)
    Xt = torch.rand(5, 3, generator=g, dtype=D)
    mean, cov = gp.posterior(Xt)
    y = gp.train_targets
    Khat = O.covar_direct(O.MATERN52, X, X, ls, 1.4) + 1e-2 * torch.eye(12, dtype=D)
    Kinv = torch.linalg.inv(Khat)
    Ks = O.covar_direct(O.MATERN52, Xt, X, ls, 1.4)
    Kss = O.covar_direct(O.MATERN52, Xt, Xt, ls, 1.4)
    m_bf = 0.2 + Ks @ Kinv @ (y - 0.2)
    c_bf = Kss - Ks @ Kinv @ Ks.T
    sy, my = gp.y_std, gp.y_mean
    torch.testing.assert_close(mean, my + sy * m_bf, rtol=1e-9, atol=1e-10)
    torch.testing.assert_close(cov, sy**2 * c_bf, rtol=1e-8, atol=1e-10)


def test_posterior_leave_one_out_identity():
    """cross_validation.py:644-679: mu_loo_i = y_i - [K^-1 r]_i / [K^-1]_ii and
    sigma2_loo_i = 1/[K^-1]_ii equal the predictive of the GP fit without point i
    (latent variance + noise), checked against a refit on n-1 points."""
    X, Y, ls = _data(10, 2, seed=4)
    noise = 1e-2
    gp = O.OracleGP(X, Y, O.MATERN52, ls, 1.0, noise, standardize_Y=False)
    Kinv = torch.cholesky_inverse(gp.L)
    r = gp.train_targets
    i = 3
    mu_loo = r[i] - (Kinv @ r)[i] / Kinv[i, i]
    var_loo = 1.0 / Kinv[i, i]
    keep = [j for j in range(10) if j != i]
    gp_i = O.OracleGP(X[keep], Y[keep], O.MATERN52, ls, 1.0, noise, standardize_Y=False)
    m, c = gp_i.posterior(X[i : i + 1])
    assert m.item() == pytest.approx(mu_loo.item(), rel=1e-9)
    assert (c.item() + noise) == pytest.approx(var_loo.item(), rel=1e-9)


# ---------------------------------------------------------------- base samples
def test_sobol_normal_samples_reproducible_and_normal():
    a = O.draw_sobol_normal_samples(5, 1024, seed=7)
    b = O.draw_sobol_normal_samples(5, 1024, seed=7)
    c = O.draw_sobol_normal_samples(5, 1024, seed=8)
    assert torch.equal(a, b) and not torch.equal(a, c)
    assert a.mean(0).abs().max() < 0.01 and (a.std(0) - 1).abs().max() < 0.02
    assert torch.isfinite(a).all()


def _joint_baseline_columns(acq, Xq, Z):
    """Samples at the baseline columns of the joint (X_baseline, X) posterior,
    drawn with the full joint Cholesky (no cached root) from base samples Z."""
    b = Xq.shape[0]
    X_full = torch.cat([acq.X_baseline.expand(b, -1, -1), Xq], dim=-2)
    mean, cov = acq.gp.posterior(X_full)
    L = torch.linalg.cholesky(cov)
    s = torch.einsum("bij,sj->sbi", L, Z) + mean.unsqueeze(0)
    return s[..., : acq.X_baseline.shape[-2]]


def test_cached_root_base_samples_keep_baseline_samples():
    """test/utils/test_low_rank.py:160-175: adding the q new points and their
    base samples does not change the posterior samples at the baseline points,
    i.e. the joint samples at the baseline columns equal the baseline samples
    drawn when the acquisition function was built (logei.py:431-447).  This fixes
    where the baseline columns of the joint base samples come from; a joint
    Sobol draw over all n_b + q columns (the mistake planted below) fails it.
    Where the q new columns come from (the sampler's own Sobol(n_b + q) draw,
    sampling/normal.py:128-131,182-213) is a convention no fixture pins."""
    X, Y, ls = _data(14, 3, seed=13)
    gp = O.OracleGP(X, Y, O.MATERN52, ls, 1.0, 1e-3)
    acq = O.OracleQLogNEI(gp, X[[1, 4, 9]], num_samples=64, seed=5)
    g = torch.Generator().manual_seed(4)
    Xq = torch.rand(5, 2, 3, generator=g, dtype=D)
    Z = acq.base_samples(2)
    got = _joint_baseline_columns(acq, Xq, Z)
    want = acq.baseline_samples.unsqueeze(1).expand_as(got)
    torch.testing.assert_close(got, want, rtol=1e-9, atol=1e-10)
    planted = O.draw_sobol_normal_samples(5, 64, seed=5)  # one joint draw: wrong source
    bad = _joint_baseline_columns(acq, Xq, planted)
    assert (bad - want).abs().max() > 1e-2


# ---------------------------------------------------------------- cached sampling
def test_cached_cholesky_equals_joint_sampling():
    """test/utils/test_low_rank.py: sample_cached_cholesky == bottom rows of the
    joint Cholesky applied to the same base samples."""
    X, Y, ls = _data(15, 3, seed=2)
    gp = O.OracleGP(X, Y, O.MATERN52, ls, 1.0, 1e-3)
    g = torch.Generator().manual_seed(9)
    Xb = torch.rand(4, 3, generator=g, dtype=D)
    Xq = torch.rand(6, 2, 3, generator=g, dtype=D)
    Xfull = torch.cat([Xb.expand(6, -1, -1), Xq], dim=-2)
    mean, cov = gp.posterior(Xfull)
    Lb = torch.linalg.cholesky(cov[0, :4, :4])
    Z = O.cached_root_base_samples(32, 4, 2, seed=1)
    s_cached = O.sample_cached_cholesky(mean[..., -2:], cov[..., -2:, :], Lb, Z)
    Ljoint = torch.linalg.cholesky(cov)
    s_joint = torch.einsum("bij,sj->sbi", Ljoint, Z)[..., -2:] + mean[..., -2:].unsqueeze(0)
    torch.testing.assert_close(s_cached, s_joint, rtol=1e-9, atol=1e-10)


# ---------------------------------------------------------------- safe math
def test_pareto_properties():
    # _pareto: p(0) = 1, |p'(0)| = 1, p''(0) = 1, decays like x^-alpha
    x = torch.tensor([0.0], dtype=D, requires_grad=True)
    p = O.pareto(x)
    (g,) = torch.autograd.grad(p.sum(), x, create_graph=True)
    (h,) = torch.autograd.grad(g.sum(), x)
    assert p.item() == 1.0 and g.item() == pytest.approx(-1.0) and h.item() == pytest.approx(1.0)
    big = torch.tensor([1e4], dtype=D)
    assert (O.pareto(big) * big**2).item() == pytest.approx(2.0, rel=1e-3)


def test_fatplus_value_at_zero_and_closed_form_derivative():
    """fatplus(x, tau) = tau (softplus(x / tau) + ALPHA cauchy(x / tau)) with
    ALPHA = 0.1 (utils/safe_math.py:322, csrc/logei_fused.cpp:28): at x = 0 the
    value is tau (ln 2 + ALPHA), which fixes the weight.  The derivative of
    log_fatplus is checked against the reference's closed form
    (csrc/logei_fused.cpp:97-111): (sigmoid(y) - 2 ALPHA y c(y)^2) / (tau f),
    y = x / tau, c(y) = 1 / (1 + y^2), f = softplus(y) + ALPHA c(y)."""
    for tau in (1e-6, 1e-2, 0.7):
        v = O.fatplus(torch.tensor([0.0], dtype=D), tau).item()
        assert v == pytest.approx(tau * (math.log(2.0) + 0.1), rel=1e-15)
    tau = 0.05
    x = torch.linspace(-3, 3, 241, dtype=D).requires_grad_(True)
    (g,) = torch.autograd.grad(O.log_fatplus(x, tau).sum(), x)
    y = x.detach() / tau
    c = 1.0 / (1.0 + y * y)
    f = torch.nn.functional.softplus(y) + 0.1 * c
    closed = (torch.sigmoid(y) - 2 * 0.1 * y * c * c) / (tau * f)
    # above y = 20 torch's softplus (which the reference calls) is linear, so its
    # derivative there is 1 instead of sigmoid(y) = 1 - O(e^-20): 2e-9 relative
    inner = y.abs() <= 20
    torch.testing.assert_close(g[inner], closed[inner], rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(g, closed, rtol=3e-9, atol=1e-12)
    # a weight of 0.2 instead of 0.1 would move the derivative by ~1e-1 near 0
    f2 = torch.nn.functional.softplus(y) + 0.2 * c
    wrong = (torch.sigmoid(y) - 2 * 0.2 * y * c * c) / (tau * f2)
    assert (g - wrong).abs().max() > 1e-2 * g.abs().max()


def test_fatplus_bounds_and_asymptotics():
    x = torch.linspace(-5, 5, 101, dtype=D)
    tau = 0.5
    f = O.fatplus(x, tau)
    assert (f >= torch.relu(x)).all()
    # softplus part is within tau*log2 of relu; the Cauchy term adds at most 0.1 tau
    assert (f - torch.relu(x)).max() <= tau * (math.log(2) + 0.1) + 1e-12
    lf = O.log_fatplus(torch.tensor([-1e3], dtype=D), 1e-6)
    assert torch.isfinite(lf).all()


def test_fatmax_within_tau_of_max():
    g = torch.Generator().manual_seed(0)
    x = torch.randn(50, 4, generator=g, dtype=D)
    tau = 1e-2
    fm = O.fatmax(x, dim=-1, tau=tau)
    mx = x.amax(-1)
    assert (fm >= mx - 1e-12).all()
    assert (fm - mx).max() <= tau * math.log(4) + 1e-12


def test_logmeanexp_constant_and_shift():
    x = torch.full((7,), 3.0, dtype=D)
    assert O.logmeanexp(x, 0).item() == pytest.approx(3.0)
    y = torch.tensor([0.0, math.log(3.0)], dtype=D)
    assert O.logmeanexp(y, 0).item() == pytest.approx(math.log(2.0))


# ---------------------------------------------------------------- qLogNEI reductions
def _reducer(best):
    class R(O.OracleQLogNEI):
        def __init__(self):
            self.tau_relu, self.tau_max = O.TAU_RELU, O.TAU_MAX
            self.baseline_best_f = best
    return R()


def test_qlognei_mock_posterior_invariant_one():
    """test/acquisition/test_logei.py:313-339: baseline sample 0, candidate sample 1
    in every MC sample -> exp(qLogNEI) = 1."""
    r = _reducer(torch.zeros(2, dtype=D))
    Y = torch.ones(2, 1, 1, dtype=D)  # S x b x q
    assert r.reduce(Y).exp().item() == pytest.approx(1.0, rel=1e-5)


def test_qlognei_mock_posterior_batch_invariant():
    """test_logei.py:383-405: batch results [1, 0] within tau_relu, second > 0."""
    r = _reducer(torch.zeros(2, dtype=D))
    Y = torch.zeros(2, 2, 2, dtype=D)
    Y[:, 0, 1] = 1.0
    res = r.reduce(Y).exp()
    assert abs(res[0].item() - 1.0) <= O.TAU_RELU
    assert 0.0 < res[1].item() < O.TAU_RELU


def test_qlognei_autograd_matches_central_fd():
    X, Y, ls = _data(16, 2, seed=6)
    gp = O.OracleGP(X, Y, O.MATERN52, ls, 1.0, 1e-3)
    acq = O.OracleQLogNEI(gp, X[:3], num_samples=64, seed=3)
    g = torch.Generator().manual_seed(2)
    Xq = torch.rand(3, 2, 2, generator=g, dtype=D)
    v, grad = acq.value_and_grad(Xq)
    h = 1e-6
    fd = torch.zeros_like(Xq)
    for idx in range(Xq.numel()):
        e = torch.zeros(Xq.numel(), dtype=D)
        e[idx] = h
        e = e.view_as(Xq)
        fd.view(-1)[idx] = ((acq(Xq + e) - acq(Xq - e)).sum() / (2 * h))
    # elements far below the largest gradient are FD-noise dominated: compare
    # on the scale of the largest entry
    assert ((grad - fd).abs().max() / grad.abs().max()).item() < 1e-7


# ---------------------------------------------------------------- MLL
def test_mll_closed_form_with_slogdet():
    X, Y, ls = _data(14, 3, seed=8)
    y = Y.squeeze(-1)
    ll = O.exact_mll(O.MATERN52, X, y, ls, torch.tensor(1.3, dtype=D), torch.tensor(0.02, dtype=D),
                     torch.tensor(0.1, dtype=D))
    K = O.covar_direct(O.MATERN52, X, X, ls, 1.3) + 0.02 * torch.eye(14, dtype=D)
    r = y - 0.1
    expect = -0.5 * (r @ torch.linalg.solve(K, r) + torch.linalg.slogdet(K)[1]
                     + 14 * math.log(2 * math.pi))
    assert ll.item() * 14 == pytest.approx(expect.item(), rel=1e-11)


@pytest.mark.parametrize("kind", [O.MATERN52, O.RBF])
def test_mll_adjoint_formula_matches_autograd(kind):
    X, Y, ls = _data(20, 3, seed=9)
    y = Y.squeeze(-1)
    lsr = ls.clone().requires_grad_(True)
    osr = torch.tensor(1.3, dtype=D, requires_grad=True)
    nr = torch.tensor(0.02, dtype=D, requires_grad=True)
    cr = torch.tensor(0.1, dtype=D, requires_grad=True)
    (O.exact_mll(kind, X, y, lsr, osr, nr, cr) * 20).backward()
    g_ls, g_os, g_n, g_c = O.exact_mll_adjoint(kind, X, y, ls, 1.3, 0.02, 0.1)
    torch.testing.assert_close(g_ls, lsr.grad, rtol=1e-8, atol=1e-10)
    assert g_os.item() == pytest.approx(osr.grad.item(), rel=1e-8)
    assert g_n.item() == pytest.approx(nr.grad.item(), rel=1e-8)
    assert g_c.item() == pytest.approx(cr.grad.item(), rel=1e-8, abs=1e-12)


def test_lognormal_prior_matches_torch_distribution():
    x = torch.tensor([0.3, 1.7], dtype=D)
    ref = torch.distributions.LogNormal(torch.tensor(-4.0, dtype=D),
                                        torch.tensor(1.0, dtype=D)).log_prob(x)
    torch.testing.assert_close(O.lognormal_log_prob(x, -4.0, 1.0), ref)


# ---------------------------------------------------------------- pruning
def test_prune_keeps_argmax_points():
    X, Y, ls = _data(10, 2, seed=11)
    gp = O.OracleGP(X, Y, O.MATERN52, ls, 1.0, 1e-4)
    Xp, idx = O.prune_inferior_points(gp, X, num_samples=256, seed=1)
    best = torch.argmax(gp.train_targets)
    assert best.item() in idx.tolist()
    Xp2, _ = O.prune_inferior_points(gp, X[:1], num_samples=16, seed=1)
    assert torch.equal(Xp2, X[:1])
    Xp3, idx3 = O.prune_inferior_points(gp, X, num_samples=256, seed=1, max_frac=0.1)
    assert len(idx3) <= 1


def test_joint_path_equals_cached_root_path():
    """acquisition/cached_cholesky.py:143-171 with test/acquisition/test_logei.py:
    513-626: sampling the q new points from the joint posterior's Cholesky
    (cache_root=False) gives the cached-root values and gradients."""
    X, Y, ls = _data(20, 3, seed=21)
    gp = O.OracleGP(X, Y, O.MATERN52, ls, 1.0, 1e-3)
    Xb = X[[2, 7, 11]]
    a = O.OracleQLogNEI(gp, Xb, num_samples=64, seed=2)
    b = O.OracleQLogNEI(gp, Xb, num_samples=64, seed=2, cache_root=False)
    Xq = torch.rand(4, 2, 3, generator=torch.Generator().manual_seed(8), dtype=D)
    va, ga = a.value_and_grad(Xq)
    vb, gb = b.value_and_grad(Xq)
    torch.testing.assert_close(va, vb, rtol=1e-9, atol=1e-12)
    torch.testing.assert_close(ga, gb, rtol=1e-7, atol=1e-10)


def test_log_areas_limit_is_the_log_box_volume():
    """acquisition/multi_objective/logei.py:437-479: as tau_relu, tau_max -> 0 the
    log-area of a subset in a cell tends to sum_k log(min(min_j (obj_jk - l_k),
    u_k - l_k)) wherever every improvement is positive; a sign slip, a max for a
    min or a dropped cell length fails this."""
    g = torch.Generator().manual_seed(0)
    obj = torch.rand(3, 4, 2, 3, generator=g, dtype=D)
    cl = torch.rand(5, 3, generator=g, dtype=D) * 0.5
    cu = cl + torch.rand(5, 3, generator=g, dtype=D)
    a = O.log_areas(obj, cl, cu, tau_relu=1e-9, tau_max=1e-9)
    Z = obj.unsqueeze(1) - cl[None, :, None, None, :]
    exact = torch.minimum(Z.min(dim=-2).values, (cu - cl)[None, :, None, :]).log().sum(-1)
    ok = Z.min(dim=-2).values.min(-1).values > 1e-3
    assert ok.sum() > 10
    torch.testing.assert_close(a[ok], exact[ok], rtol=0, atol=1e-12)
    # an unbounded cell side is clamped at 1e10 (FP64)
    cu_inf = cu.clone()
    cu_inf[0, 0] = float("inf")
    b = O.log_areas(obj, cl, cu_inf)
    c = O.log_areas(obj, cl, torch.where(torch.isinf(cu_inf), torch.full_like(cu, 1e10), cu))
    torch.testing.assert_close(b, c)
    # batched cells (one set per batch element) equal the shared cells repeated
    torch.testing.assert_close(O.log_areas(obj, cl.expand(3, -1, -1), cu.expand(3, -1, -1)),
                               O.log_areas(obj, cl, cu))
