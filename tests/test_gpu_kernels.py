"""GPU parity of the individual C-ABI kernels against the FP64 oracle / torch.

Each test prints its case before launching, so a failing run names it.
"""
import ctypes

import pytest
import torch

from botorch_amd import _lib
from oracle import gp_oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _scaled(X, center, ls, n_pad):
    xs = torch.zeros(n_pad, X.shape[1], dtype=torch.float64, device=DEV)
    xs[: X.shape[0]] = (X - center) / ls
    return xs


@pytest.mark.parametrize("kind", [O.MATERN52, O.RBF])
@pytest.mark.parametrize("out_f32", [0, 1])
def test_covar_matches_oracle(kind, out_f32):
    _say(f"covar kind={kind} f32={out_f32}")
    L = _lib.lib()
    g = torch.Generator().manual_seed(1)
    X1 = torch.rand(37, 6, generator=g, dtype=torch.float64)
    X2 = torch.rand(50, 6, generator=g, dtype=torch.float64)
    ls = 0.3 + torch.rand(6, generator=g, dtype=torch.float64)
    c = X1.mean(0)
    x1s = _scaled(X1.to(DEV), c.to(DEV), ls.to(DEV), 64)
    x2s = _scaled(X2.to(DEV), c.to(DEV), ls.to(DEV), 64)
    dt = torch.float32 if out_f32 else torch.float64
    K = torch.full((64, 128), 7.0, dtype=dt, device=DEV)
    _lib.check(L.bta_covar(kind, _lib.ptr(x1s), 37, _lib.ptr(x2s), 50, 6, 1.7, 0.0, out_f32,
                           _lib.ptr(K), 128, 64, 128, _lib.stream()))
    ref = O.covar_direct(kind, X1, X2, ls, 1.7)
    got = K[:37, :50].double().cpu()
    tol = 1e-6 if out_f32 else 1e-13
    torch.testing.assert_close(got, ref, rtol=tol, atol=tol)
    assert (K[37:, :].abs().max() == 0) and (K[:, 50:].abs().max() == 0)


@pytest.mark.parametrize("f32", [False, True])
@pytest.mark.parametrize("flags,shape", [
    (0, (128, 192, 256)),
    (_lib.GEMM_B_LOWER, (192, 256, 256)),
    (_lib.GEMM_B_KN | _lib.GEMM_B_UPPER, (128, 256, 384)),
    (_lib.GEMM_A_LOWER | _lib.GEMM_B_KN, (256, 128, 256)),
    (_lib.GEMM_A_KM | _lib.GEMM_B_KN, (128, 128, 192)),
    (0, (1024, 1280, 512)),
    (_lib.GEMM_B_LOWER, (2048, 2176, 2048)),
])
def test_gemm_matches_torch(flags, shape, f32):
    _say(f"gemm flags={flags} shape={shape} f32={f32}")
    M, N, K = shape
    dt = torch.float32 if f32 else torch.float64
    g = torch.Generator().manual_seed(M + N + K + flags)
    A = torch.randn(M, K, generator=g, dtype=torch.float64)
    Bnk = torch.randn(N, K, generator=g, dtype=torch.float64)
    if flags & _lib.GEMM_A_LOWER:
        A = torch.tril(A)
    opB = Bnk  # op(B) as N x K
    if flags & _lib.GEMM_B_LOWER:
        opB = torch.tril(opB)
    if flags & _lib.GEMM_B_UPPER:
        opB = torch.triu(opB)
    Bstore = opB.t().contiguous() if flags & _lib.GEMM_B_KN else opB.contiguous()
    Astore = A.t().contiguous() if flags & _lib.GEMM_A_KM else A.contiguous()
    C0 = torch.randn(M, N, generator=g, dtype=torch.float64)
    ref = 0.5 * A @ opB.t() - 2.0 * C0
    Ad, Bd, C0d = (t.to(DEV, dt) for t in (Astore, Bstore, C0))
    C = torch.empty(M, N, dtype=dt, device=DEV)
    lda = M if flags & _lib.GEMM_A_KM else K
    ldb = N if flags & _lib.GEMM_B_KN else K
    fl = flags | (_lib.GEMM_F32 if f32 else 0)
    _lib.check(_lib.lib().bta_gemm(fl, M, N, K, 0.5, _lib.ptr(Ad), lda, _lib.ptr(Bd), ldb, -2.0,
                                   _lib.ptr(C0d), N, _lib.ptr(C), N, 1, 0, 0, 0, 0,
                                   _lib.stream()))
    got = C.double().cpu()
    if f32:
        ref32 = (0.5 * Ad.float() @ (Bd.float().t() if not flags & _lib.GEMM_B_KN else Bd.float())
                 if not flags & _lib.GEMM_A_KM else None)
        torch.testing.assert_close(got, ref, rtol=1e-4, atol=1e-3 * K**0.5)
    else:
        torch.testing.assert_close(got, ref, rtol=1e-12, atol=1e-11 * K**0.5)


@pytest.mark.parametrize("flags", [_lib.GEMM_B_KN | _lib.GEMM_B_UPPER,
                                   _lib.GEMM_A_LOWER | _lib.GEMM_B_KN])
@pytest.mark.parametrize("n,batch", [(1024, 2), (2048, 1)])
def test_gemm_triangular_split_k(flags, n, batch):
    """Triangular FP64 products without C0 (the shapes of the triangular
    inversion levels), where the longest tiles are cut in two and added with
    atomics: against torch, and bit-identical over repeated calls."""
    _say(f"gemm split-K flags={flags} n={n} batch={batch}")
    g = torch.Generator().manual_seed(n + flags)
    A = torch.randn(batch, n, n, generator=g, dtype=torch.float64)
    Bnk = torch.randn(batch, n, n, generator=g, dtype=torch.float64)
    if flags & _lib.GEMM_A_LOWER:
        A = torch.tril(A)
    opB = torch.triu(Bnk) if flags & _lib.GEMM_B_UPPER else Bnk
    ref = -1.0 * A @ opB.transpose(-1, -2)
    Ad = A.to(DEV).contiguous()
    Bd = opB.transpose(-1, -2).contiguous().to(DEV)  # K x N storage (B_KN)
    outs = []
    for _ in range(3):
        C = torch.full((batch, n, n), 3.0, dtype=torch.float64, device=DEV)
        _lib.check(_lib.lib().bta_gemm(flags, n, n, n, -1.0, _lib.ptr(Ad), n, _lib.ptr(Bd), n,
                                       0.0, None, 0, _lib.ptr(C), n, batch, n * n, n * n, 0,
                                       n * n, _lib.stream()))
        outs.append(C.cpu())
    torch.testing.assert_close(outs[0], ref, rtol=1e-12, atol=1e-11 * n**0.5)
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])


def test_gemm_batched_lower_output():
    _say("gemm batched C_LOWER")
    L = _lib.lib()
    g = torch.Generator().manual_seed(3)
    nb, M, K = 3, 192, 128
    A = torch.randn(nb, M, K, generator=g, dtype=torch.float64)
    C0 = torch.randn(nb, M, M, generator=g, dtype=torch.float64)
    ref = C0 - A @ A.transpose(-1, -2)
    Ad, C = A.to(DEV), C0.to(DEV).clone()
    _lib.check(L.bta_gemm(_lib.GEMM_C_LOWER, M, M, K, -1.0, _lib.ptr(Ad), K, _lib.ptr(Ad), K, 1.0,
                          _lib.ptr(C), M, _lib.ptr(C), M, nb, M * K, M * K, M * M, M * M,
                          _lib.stream()))
    got = C.cpu()
    torch.testing.assert_close(torch.tril(got), torch.tril(ref), rtol=1e-12, atol=1e-11)


def _spd(n, seed, noise=1e-4):
    g = torch.Generator().manual_seed(seed)
    X = torch.rand(n, 5, generator=g, dtype=torch.float64)
    return O.covar_direct(O.MATERN52, X, X, torch.full((5,), 0.4, dtype=torch.float64)) + \
        noise * torch.eye(n, dtype=torch.float64)


@pytest.mark.parametrize("n", [64, 200, 640])
def test_potrf_and_trtri(n):
    _say(f"potrf/trtri n={n}")
    L = _lib.lib()
    A = _spd(n, n)
    n_pad = _lib.round_up(n, 128)
    Ad = torch.eye(n_pad, dtype=torch.float64, device=DEV)
    Ad[:n, :n] = A.to(DEV)
    Linv = torch.full((n_pad, n_pad), 3.0, dtype=torch.float64, device=DEV)
    ws = _lib.workspace(max(L.bta_potrf_workspace_size(n_pad), n_pad * n_pad * 8), DEV)
    jit = ctypes.c_double(-1)
    _lib.check(L.bta_potrf(_lib.ptr(Ad), n_pad, n_pad, _lib.ptr(Linv), n_pad, 6, 1e-8,
                           ctypes.byref(jit), _lib.ptr(ws), ws.numel(), _lib.stream()))
    Lref = torch.linalg.cholesky(A)
    torch.testing.assert_close(Ad[:n, :n].cpu(), Lref, rtol=1e-10, atol=1e-12)
    assert jit.value == 0.0
    assert torch.triu(Ad, 1).abs().max().item() == 0.0
    _lib.check(L.bta_trtri_lower(_lib.ptr(Ad), n_pad, _lib.ptr(Linv), n_pad, n_pad, _lib.ptr(ws),
                                 ws.numel(), _lib.stream()))
    Iref = torch.linalg.solve_triangular(Lref, torch.eye(n, dtype=torch.float64), upper=False)
    torch.testing.assert_close(Linv[:n, :n].cpu(), Iref, rtol=1e-8, atol=1e-9 * Iref.abs().max())
    assert torch.triu(Linv, 1).abs().max().item() == 0.0


@pytest.mark.parametrize("n,n_pad", [(64, 64), (100, 128), (200, 256), (300, 320),
                                       (1000, 1088), (1500, 1536)])
def test_potrf_inverse(n, n_pad):
    """Cholesky with L^-1 built inside the factorisation's launches (odd block
    counts included: 5 and 17 blocks of 64) against torch's FP64 Cholesky and
    triangular solve, and against the separate inverse pass."""
    _say(f"potrf_inverse n={n} n_pad={n_pad}")
    L = _lib.lib()
    A = _spd(n, n)
    Ad = torch.eye(n_pad, dtype=torch.float64, device=DEV)
    Ad[:n, :n] = A.to(DEV)
    A2 = Ad.clone()
    Linv = torch.full((n_pad, n_pad), 3.0, dtype=torch.float64, device=DEV)
    ws = _lib.workspace(max(L.bta_potrf_workspace_size(n_pad), n_pad * n_pad * 8), DEV)
    jit = ctypes.c_double(-1)
    _lib.check(L.bta_potrf_inverse(_lib.ptr(Ad), n_pad, n_pad, _lib.ptr(Linv), n_pad, 6, 1e-8,
                                   ctypes.byref(jit), _lib.ptr(ws), ws.numel(), _lib.stream()))
    Lref = torch.linalg.cholesky(A)
    torch.testing.assert_close(Ad[:n, :n].cpu(), Lref, rtol=1e-10, atol=1e-12)
    assert jit.value == 0.0
    assert torch.triu(Ad, 1).abs().max().item() == 0.0
    Iref = torch.linalg.solve_triangular(Lref, torch.eye(n, dtype=torch.float64), upper=False)
    torch.testing.assert_close(Linv[:n, :n].cpu(), Iref, rtol=1e-8, atol=1e-9 * Iref.abs().max())
    assert torch.triu(Linv, 1).abs().max().item() == 0.0
    # the padding block of the identity inverts to the identity
    torch.testing.assert_close(Linv[n:, n:].cpu(), torch.eye(n_pad - n, dtype=torch.float64))
    # the two-pass route (factorisation, then bta_trtri_lower) gives the same inverse
    Linv2 = torch.zeros_like(Linv)
    _lib.check(L.bta_potrf(_lib.ptr(A2), n_pad, n_pad, _lib.ptr(Linv2), n_pad, 6, 1e-8,
                           ctypes.byref(jit), _lib.ptr(ws), ws.numel(), _lib.stream()))
    _lib.check(L.bta_trtri_lower(_lib.ptr(A2), n_pad, _lib.ptr(Linv2), n_pad, n_pad,
                                 _lib.ptr(ws), ws.numel(), _lib.stream()))
    err = ((Linv - Linv2).abs().max() / Linv2.abs().max()).item()
    assert err < 1e-11, err


def test_potrf_inverse_large_and_jitter():
    """n = 4096 (64 steps with side tasks in every one) twice on one workspace,
    then a rank-deficient matrix that needs the jitter retry: the inverse must
    be that of the jittered factor."""
    _say("potrf_inverse n=4096 x2")
    L = _lib.lib()
    n = 4096
    g = torch.Generator(device=DEV).manual_seed(11)
    B = torch.randn(n, n, dtype=torch.float64, device=DEV, generator=g)
    A0 = B @ B.T / n + 0.05 * torch.eye(n, dtype=torch.float64, device=DEV)
    Lref = torch.linalg.cholesky(A0)
    Iref = torch.linalg.solve_triangular(Lref, torch.eye(n, dtype=torch.float64, device=DEV),
                                         upper=False)
    ws = _lib.workspace(L.bta_potrf_workspace_size(n), DEV)
    Linv = torch.empty_like(A0)
    for _ in range(2):
        A = A0.clone()
        Linv.fill_(float("nan"))
        jit = ctypes.c_double(-1)
        _lib.check(L.bta_potrf_inverse(_lib.ptr(A), n, n, _lib.ptr(Linv), n, 6, 1e-8,
                                       ctypes.byref(jit), _lib.ptr(ws), ws.numel(),
                                       _lib.stream()))
        err = ((Linv - Iref).abs().max() / Iref.abs().max()).item()
        assert err < 1e-10, err
    _say("potrf_inverse jitter retry n=100 (n_pad 192)")
    gc = torch.Generator().manual_seed(5)
    V = torch.randn(100, 30, generator=gc, dtype=torch.float64)
    A = V @ V.t()
    n_pad = 192
    Ad = torch.eye(n_pad, dtype=torch.float64, device=DEV)
    Ad[:100, :100] = A.to(DEV)
    Linv = torch.full((n_pad, n_pad), float("nan"), dtype=torch.float64, device=DEV)
    ws = _lib.workspace(L.bta_potrf_workspace_size(n_pad), DEV)
    jit = ctypes.c_double(-1)
    _lib.check(L.bta_potrf_inverse(_lib.ptr(Ad), n_pad, n_pad, _lib.ptr(Linv), n_pad, 6, 1e-8,
                                   ctypes.byref(jit), _lib.ptr(ws), ws.numel(), _lib.stream()))
    assert jit.value > 0.0
    Lg = Ad.cpu()
    Ig = torch.linalg.solve_triangular(Lg, torch.eye(n_pad, dtype=torch.float64), upper=False)
    err = ((Linv.cpu() - Ig).abs().max() / Ig.abs().max()).item()
    assert err < 1e-8, err


def test_potrf_large_repeated():
    """n = 4096: the factorisation's step launches then hold more workgroups than
    CUs (strip and trailing-update workgroups share CUs); three runs must all
    agree with torch's Cholesky."""
    _say("potrf n=4096 x3")
    L = _lib.lib()
    n = 4096
    g = torch.Generator(device=DEV).manual_seed(7)
    B = torch.randn(n, n, dtype=torch.float64, device=DEV, generator=g)
    A0 = B @ B.T / n + torch.eye(n, dtype=torch.float64, device=DEV)
    Lref = torch.linalg.cholesky(A0)
    ws = _lib.workspace(L.bta_potrf_workspace_size(n), DEV)
    Linv = torch.zeros_like(A0)
    for _ in range(3):
        A = A0.clone()
        jit = ctypes.c_double(-1)
        _lib.check(L.bta_potrf(_lib.ptr(A), n, n, _lib.ptr(Linv), n, 6, 1e-8, ctypes.byref(jit),
                               _lib.ptr(ws), ws.numel(), _lib.stream()))
        err = ((A - Lref).abs().max() / Lref.abs().max()).item()
        assert err < 1e-12, err


def test_potrf_jitter_schedule():
    _say("potrf jitter on a singular matrix")
    L = _lib.lib()
    g = torch.Generator().manual_seed(5)
    V = torch.randn(100, 30, generator=g, dtype=torch.float64)
    A = V @ V.t()  # rank 30: needs jitter
    ref = O.psd_safe_cholesky(A, warn=False)
    n_pad = 128
    Ad = torch.eye(n_pad, dtype=torch.float64, device=DEV)
    Ad[:100, :100] = A.to(DEV)
    ws = _lib.workspace(L.bta_potrf_workspace_size(n_pad), DEV)
    jit = ctypes.c_double(-1)
    _lib.check(L.bta_potrf(_lib.ptr(Ad), n_pad, n_pad, None, 0, 6, 1e-8, ctypes.byref(jit),
                           _lib.ptr(ws), ws.numel(), _lib.stream()))
    # the oracle's schedule: the first jitter 1e-8 * 10^i at which cholesky_ex succeeds
    ref_jit = next(1e-8 * 10**i for i in range(6)
                   if torch.linalg.cholesky_ex(A + 1e-8 * 10**i * torch.eye(100,
                                               dtype=torch.float64))[1] == 0)
    assert jit.value == pytest.approx(ref_jit)
    Lg = Ad[:100, :100].cpu()
    recon = Lg @ Lg.t() - (A + jit.value * torch.eye(100, dtype=torch.float64))
    assert recon.abs().max().item() < 1e-10 * A.abs().max().item()
    assert ref.shape == Lg.shape


def test_potrf_not_psd_raises():
    _say("potrf negative definite -> NotPSDError")
    from botorch_amd.exceptions import NotPSDError
    L = _lib.lib()
    Ad = -torch.eye(128, dtype=torch.float64, device=DEV)
    ws = _lib.workspace(L.bta_potrf_workspace_size(128), DEV)
    jit = ctypes.c_double(0)
    rc = L.bta_potrf(_lib.ptr(Ad), 128, 128, None, 0, 6, 1e-8, ctypes.byref(jit), _lib.ptr(ws),
                     ws.numel(), _lib.stream())
    with pytest.raises(NotPSDError):
        _lib.check(rc)


def test_chol_small_batched_matches_oracle():
    _say("chol_small_batched")
    L = _lib.lib()
    g = torch.Generator().manual_seed(7)
    V = torch.randn(9, 5, 5, generator=g, dtype=torch.float64)
    A = V @ V.transpose(-1, -2) + 1e-3 * torch.eye(5, dtype=torch.float64)
    A[4] = torch.ones(5, 5, dtype=torch.float64)  # singular: jitter path
    ref = O.psd_safe_cholesky(A, warn=False)
    Ad = A.to(DEV).contiguous()
    status = torch.zeros(9, dtype=torch.int32, device=DEV)
    _lib.check(L.bta_chol_small_batched(_lib.ptr(Ad), 5, 5, 25, 9, 6, 1e-8, _lib.ptr(status),
                                        _lib.stream()))
    torch.testing.assert_close(Ad.cpu(), ref, rtol=1e-6, atol=1e-6)
    assert status.cpu()[4].item() & _lib.STATUS_JITTER


def test_gemm_f32_acc64_segments():
    _say("gemm f32 ACC64 (FP64 segment sums)")
    M, N, K = 256, 256, 2048
    g = torch.Generator().manual_seed(77)
    A = torch.randn(M, K, generator=g, dtype=torch.float64).float().double()
    B = torch.tril(torch.randn(N, K, generator=g, dtype=torch.float64)).float().double()
    ref = A @ B.t()
    Ad, Bd = A.float().to(DEV), B.float().to(DEV)
    C = torch.empty(M, N, dtype=torch.float32, device=DEV)
    for flags in (_lib.GEMM_F32 | _lib.GEMM_B_LOWER | _lib.GEMM_ACC64,):
        _lib.check(_lib.lib().bta_gemm(flags, M, N, K, 1.0, _lib.ptr(Ad), K, _lib.ptr(Bd), K, 0.0,
                                       None, 0, _lib.ptr(C), N, 1, 0, 0, 0, 0, _lib.stream()))
        # FP64 sums of 256-long FP32 segments: |err| <= 256 u sum_k |a_k b_k| + u |c|
        # (u = 2^-24); plain FP32 accumulation over K = 2048 has the bound 2048 u.
        u = 2.0 ** -24
        bound = 256 * u * (A.abs() @ B.abs().t()) + u * ref.abs()
        ratio = ((C.double().cpu() - ref).abs() / bound).max().item()
        assert ratio < 1.0, ratio
